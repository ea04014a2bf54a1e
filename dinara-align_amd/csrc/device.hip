// Device contexts and memory (see device.h).
#include "device.h"

#include <mutex>
#include <vector>

namespace dinara {

namespace {
std::mutex contexts_mutex;
std::vector<DeviceContext *> contexts;
}  // namespace

int64_t device_count() {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess) return 0;
    return count;
}

DeviceContext *device_context(int64_t gpu) {
    std::lock_guard<std::mutex> lock(contexts_mutex);
    const int64_t count = device_count();
    if (gpu < 0 || gpu >= count) return nullptr;
    if ((int64_t)contexts.size() < count) contexts.resize(count, nullptr);
    if (contexts[gpu]) return contexts[gpu];
    int previous = -1;
    hipGetDevice(&previous);
    if (hipSetDevice((int)gpu) != hipSuccess) return nullptr;
    auto *ctx = new DeviceContext();
    ctx->id = (int)gpu;
    hipDeviceProp_t props;
    bool ok = hipGetDeviceProperties(&props, (int)gpu) == hipSuccess;
    if (ok) {
        // Dynamic LDS up to 64 KiB needs no opt-in attribute for a launch.
        ctx->lds_limit = props.sharedMemPerBlock < 65536 ? props.sharedMemPerBlock : 65536;
        ctx->compute_units = props.multiProcessorCount;
    }
    // Default priority only: one ordinary stream per device.
    ok = ok && hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipEventCreate(&ctx->start) == hipSuccess && hipEventCreate(&ctx->stop) == hipSuccess;
    ok = ok && hipEventCreate(&ctx->sweep_start) == hipSuccess && hipEventCreate(&ctx->sweep_stop) == hipSuccess;
    if (previous >= 0) hipSetDevice(previous);
    if (!ok) {
        delete ctx;
        return nullptr;
    }
    contexts[gpu] = ctx;
    return ctx;
}

hipEvent_t DeviceContext::event(size_t i) {
    while (pool.size() <= i) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        pool.push_back(e);
    }
    return pool[i];
}

hipStream_t DeviceContext::side(size_t i) {
    while (sides.size() <= i) {
        hipStream_t st = nullptr;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return nullptr;
        sides.push_back(st);
    }
    return sides[i];
}

hipEvent_t DeviceContext::mark(size_t i) {
    while (marks.size() <= i) {
        hipEvent_t e = nullptr;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
        marks.push_back(e);
    }
    return marks[i];
}

DeviceGuard::DeviceGuard(const DeviceContext *ctx) {
    hipGetDevice(&previous);
    if (ctx && ctx->id != previous) hipSetDevice(ctx->id);
}

DeviceGuard::~DeviceGuard() {
    if (previous >= 0) hipSetDevice(previous);
}

namespace {
thread_local Carve *active_carve = nullptr;
}  // namespace

void *Carve::take(size_t bytes) {
    const size_t at = (used + 255) & ~size_t(255);
    if (at > size || bytes > size - at) {
        exhausted = true;
        return nullptr;
    }
    used = at + bytes;
    return base + at;
}

CarveScope::CarveScope(Carve *carve) : previous(active_carve) { active_carve = carve; }

CarveScope::~CarveScope() { active_carve = previous; }

DeviceBuffer::~DeviceBuffer() {
    if (ptr && !carved) (void)hipFree(ptr);
}

void DeviceBuffer::release() {
    if (ptr && !carved) (void)hipFree(ptr);
    ptr = nullptr;
    size = 0;
    carved = false;
}

bool DeviceBuffer::reserve(size_t bytes) {
    if (bytes <= size && ptr) return true;
    release();
    if (Carve *carve = active_carve) {
        // A call that may not allocate: the bytes come from its region, or the reservation fails.
        ptr = carve->take(bytes);
        if (!ptr) return false;
        carved = true;
        size = bytes;
        return true;
    }
    if (hipMalloc(&ptr, bytes) != hipSuccess) {
        ptr = nullptr;
        return false;
    }
    size = bytes;
    return true;
}

bool DeviceBuffer::upload(DeviceContext *ctx, const void *host, size_t bytes) {
    if (!reserve(bytes < 4 ? 4 : bytes)) return false;
    if (bytes == 0) return true;
    return hipMemcpyAsync(ptr, host, bytes, hipMemcpyHostToDevice, ctx->stream) == hipSuccess &&
           hipStreamSynchronize(ctx->stream) == hipSuccess;
}

bool DeviceBuffer::write(DeviceContext *ctx, const void *host, size_t bytes) {
    if (bytes == 0) return true;
    if (bytes > size) return false;
    return hipMemcpyAsync(ptr, host, bytes, hipMemcpyHostToDevice, ctx->stream) == hipSuccess &&
           hipStreamSynchronize(ctx->stream) == hipSuccess;
}

bool DeviceBuffer::read(DeviceContext *ctx, void *host, size_t bytes) const {
    if (bytes == 0) return true;
    if (bytes > size) return false;
    return hipMemcpyAsync(host, ptr, bytes, hipMemcpyDeviceToHost, ctx->stream) == hipSuccess &&
           hipStreamSynchronize(ctx->stream) == hipSuccess;
}

}  // namespace dinara
