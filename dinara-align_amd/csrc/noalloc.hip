// dinara_gpu_distances_static: unit-cost distances for callers that cannot allocate device memory (code under
// graph capture, serving loops, harnesses that forbid hipMalloc). Every device buffer the batch needs -- its packed
// letters, queues, counters and results -- is carved from `scratch`, one fixed array in this library's code
// object, of which each GPU holds its own copy from the library's first use on it until the process ends.
#include <algorithm>
#include <mutex>
#include <vector>

#include "../../include/dinara.h"
#include "carved.h"
#include "device.h"

namespace dinara {
namespace {

constexpr size_t SCRATCH_BYTES = size_t(DINARA_STATIC_SCRATCH_BYTES);
// The queues, counters and results a pair takes on the bit-parallel path, besides its letters.
constexpr size_t BYTES_PER_PAIR = 444;

// The region itself. In host code this name is a host-side stand-in, not the device address: only the address
// hipGetSymbolAddress returns (kept per device below) is ever handed to a kernel or a copy.
static __device__ __align__(256) uint8_t scratch[SCRATCH_BYTES];

struct DeviceScratch {
    std::mutex lock;          // one call on a device's scratch at a time
    uint8_t *ptr = nullptr;   // the device's copy of `scratch`, looked up once
    bool looked_up = false;
};

std::mutex table_lock;
std::vector<DeviceScratch *> table;

// The device's scratch entry, its address looked up on first use with that device current; null when the lookup
// fails.
DeviceScratch *scratch_of(DeviceContext *ctx) {
    DeviceScratch *entry;
    {
        std::lock_guard<std::mutex> hold(table_lock);
        if ((int)table.size() <= ctx->id) table.resize(ctx->id + 1, nullptr);
        if (!table[ctx->id]) table[ctx->id] = new DeviceScratch();
        entry = table[ctx->id];
    }
    std::lock_guard<std::mutex> hold(entry->lock);
    if (!entry->looked_up) {
        DeviceGuard guard(ctx);
        void *address = nullptr;
        if (hipGetSymbolAddress(&address, HIP_SYMBOL(scratch)) != hipSuccess || !address) return nullptr;
        entry->ptr = static_cast<uint8_t *>(address);
        entry->looked_up = true;
    }
    return entry;
}

}  // namespace
}  // namespace dinara

using namespace dinara;

extern "C" {

int64_t dinara_gpu_distances_static(int64_t gpu, int64_t pairs, const char *const *references,
                                    const int64_t *reference_lengths, const char *const *queries,
                                    const int64_t *query_lengths, const dinara_costs *costs, const dinara_mode *mode,
                                    const dinara_options *options, int64_t threads, int64_t *results) {
    if (pairs <= 0) return 0;
    // A first bound before any work: every pair takes its letters at 2 bits at least and a few hundred bytes of
    // queues and counters.
    size_t least = 0;
    for (int64_t i = 0; i < pairs; ++i)
        least += (size_t)(std::max<int64_t>(reference_lengths[i], 0) + std::max<int64_t>(query_lengths[i], 0)) / 4 +
                 BYTES_PER_PAIR;
    if (least > SCRATCH_BYTES) return DINARA_SCRATCH_FULL;
    DeviceContext *ctx = device_context(gpu);
    if (!ctx) return DINARA_NO_DEVICE;
    DeviceScratch *entry = scratch_of(ctx);
    if (!entry) return DINARA_NO_DEVICE;
    std::lock_guard<std::mutex> hold(entry->lock);
    Carve carve;
    carve.base = entry->ptr;
    carve.size = SCRATCH_BYTES;
    return carved_unit_distances(&carve, gpu, pairs, references, reference_lengths, queries, query_lengths, costs,
                                 mode, options, threads, results);
}

}  // extern "C"
