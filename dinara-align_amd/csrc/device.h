// Device contexts and memory of the library: one stream per GPU, every call ordered on it.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <mutex>
#include <vector>

namespace dinara {

struct DeviceContext {
    int id = 0;
    hipStream_t stream = nullptr;
    hipEvent_t start = nullptr, stop = nullptr, sweep_start = nullptr, sweep_stop = nullptr;
    size_t lds_limit = 0;  // dynamic LDS one workgroup may take
    int compute_units = 0;
    std::vector<hipEvent_t> pool;  // timing events, made as needed
    hipEvent_t event(size_t i);
    // Further streams for launches that may run side by side, forked from and joined back to `stream` by events;
    // made as needed, default priority.
    std::vector<hipStream_t> sides;
    std::vector<hipEvent_t> marks;  // events that only order streams (no timing)
    hipStream_t side(size_t i);
    hipEvent_t mark(size_t i);
    std::mutex lock;  // one caller's work on the stream at a time (the events are shared)
    // The unit-cost path's decision arena, lent to each alignment run under `lock` (a run walks its decisions before
    // it returns, so no batch holds it between runs): allocated once, grown when a run wants more, kept until the
    // process ends, so a fresh batch does not pay for mapping tens of gigabytes.
    struct DeviceBuffer *edit_arena = nullptr;
};

// The context of GPU `gpu`, made on first use; null when there is no such GPU.
DeviceContext *device_context(int64_t gpu);
int64_t device_count();

// Makes the context's GPU current for the calling thread, and restores the previous one after.
struct DeviceGuard {
    int previous = -1;
    explicit DeviceGuard(const DeviceContext *ctx);
    ~DeviceGuard();
};

// A fixed device region handed out front to back, 256-byte aligned: the memory of calls that may not allocate.
struct Carve {
    uint8_t *base = nullptr;
    size_t size = 0;
    size_t used = 0;
    bool exhausted = false;    // a reservation found no room
    void *take(size_t bytes);  // null (and `exhausted` set) when the region has no room left
};

// While one of these lives on a thread, every DeviceBuffer that thread reserves comes from `carve` and no device
// memory is allocated; such a buffer is never freed (the region outlives it).
struct CarveScope {
    Carve *previous;
    explicit CarveScope(Carve *carve);
    ~CarveScope();
};

// Device memory that grows on demand and is freed with its owner.
struct DeviceBuffer {
    void *ptr = nullptr;
    size_t size = 0;
    bool carved = false;  // taken from a Carve: not ours to free
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    ~DeviceBuffer();
    bool reserve(size_t bytes);                                          // at least `bytes`, contents dropped
    void release();                                                      // frees it now
    bool upload(DeviceContext *ctx, const void *host, size_t bytes);     // reserve, copy, wait
    bool write(DeviceContext *ctx, const void *host, size_t bytes);      // copy into what is there, wait
    bool read(DeviceContext *ctx, void *host, size_t bytes) const;       // copy out after the stream's work
};

}  // namespace dinara
