// Roofline denominators measured on the device itself: the int32 VALU rate and the HBM copy rate. Not part of the
// alignment library; bench.py loads it to price its kernels, and its copy kernel is the plain copy of known size the
// HBM byte counters are calibrated against.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

constexpr int CHAINS = 8;  // independent dependency chains a lane, so issue never waits on a result

// Each iteration: CHAINS x (one v_add_u32 and one v_xor_b32), every one on a full 64-lane wave. The loop counter
// and branch are scalar instructions. The result is stored so nothing is dropped.
__global__ __launch_bounds__(256) void valu_chains(uint32_t *out, int iterations, uint32_t seed) {
    uint32_t a[CHAINS], b[CHAINS];
#pragma unroll
    for (int i = 0; i < CHAINS; ++i) {
        a[i] = seed + threadIdx.x * (i + 1);
        b[i] = seed ^ (blockIdx.x + i);
    }
    for (int it = 0; it < iterations; ++it) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
#pragma unroll
            for (int i = 0; i < CHAINS; ++i) {
                a[i] += b[i];
                b[i] ^= a[i];
            }
        }
    }
    uint32_t x = 0;
#pragma unroll
    for (int i = 0; i < CHAINS; ++i) x ^= a[i] + b[i];
    out[blockIdx.x * blockDim.x + threadIdx.x] = x;
}

// A plain copy, 16 bytes a lane a step, grid-strided.
__global__ __launch_bounds__(256) void copy16(const uint4 *__restrict__ in, uint4 *__restrict__ out, size_t count) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) out[i] = in[i];
}

}  // namespace

extern "C" {

// Int32 VALU ops a second (tera) over `repeats` launches of the chains on every compute unit; 0 on failure.
// `ops_per_launch` receives the ops one launch counts: lanes x iterations x 16 x CHAINS x 2.
double probe_valu(int64_t blocks_per_cu, int64_t iterations, int64_t repeats, double *ops_per_launch) {
    int device = 0;
    hipDeviceProp_t props;
    if (hipGetDevice(&device) != hipSuccess || hipGetDeviceProperties(&props, device) != hipSuccess) return 0;
    const int blocks = props.multiProcessorCount * (int)blocks_per_cu;
    uint32_t *out = nullptr;
    if (hipMalloc(&out, (size_t)blocks * 256 * 4) != hipSuccess) return 0;
    hipEvent_t start, stop;
    (void)hipEventCreate(&start);
    (void)hipEventCreate(&stop);
    hipLaunchKernelGGL(valu_chains, dim3(blocks), dim3(256), 0, 0, out, (int)iterations, 1u);  // warm
    (void)hipEventRecord(start, 0);
    for (int64_t r = 0; r < repeats; ++r)
        hipLaunchKernelGGL(valu_chains, dim3(blocks), dim3(256), 0, 0, out, (int)iterations, (uint32_t)r);
    (void)hipEventRecord(stop, 0);
    const bool ok = hipEventSynchronize(stop) == hipSuccess;
    float ms = 0;
    (void)hipEventElapsedTime(&ms, start, stop);
    (void)hipEventDestroy(start);
    (void)hipEventDestroy(stop);
    (void)hipFree(out);
    const double ops = (double)blocks * 256 * (double)iterations * 16 * CHAINS * 2;
    if (ops_per_launch) *ops_per_launch = ops;
    return ok && ms > 0 ? ops * (double)repeats / (ms / 1e3) / 1e12 : 0;
}

// HBM bytes a second (giga: `bytes` read plus `bytes` written a copy) over `repeats` copies of `bytes`; 0 on
// failure. `copy_ms` receives one copy's mean time.
double probe_copy(int64_t bytes, int64_t repeats, double *copy_ms) {
    const size_t count = (size_t)bytes / 16;
    uint4 *in = nullptr, *out = nullptr;
    if (hipMalloc(&in, count * 16) != hipSuccess) return 0;
    if (hipMalloc(&out, count * 16) != hipSuccess) {
        (void)hipFree(in);
        return 0;
    }
    (void)hipMemset(in, 1, count * 16);
    int device = 0;
    hipDeviceProp_t props;
    (void)hipGetDevice(&device);
    (void)hipGetDeviceProperties(&props, device);
    const int blocks = props.multiProcessorCount * 8;
    hipEvent_t start, stop;
    (void)hipEventCreate(&start);
    (void)hipEventCreate(&stop);
    hipLaunchKernelGGL(copy16, dim3(blocks), dim3(256), 0, 0, in, out, count);  // warm
    (void)hipEventRecord(start, 0);
    for (int64_t r = 0; r < repeats; ++r) hipLaunchKernelGGL(copy16, dim3(blocks), dim3(256), 0, 0, in, out, count);
    (void)hipEventRecord(stop, 0);
    const bool ok = hipEventSynchronize(stop) == hipSuccess;
    float ms = 0;
    (void)hipEventElapsedTime(&ms, start, stop);
    (void)hipEventDestroy(start);
    (void)hipEventDestroy(stop);
    (void)hipFree(in);
    (void)hipFree(out);
    if (copy_ms) *copy_ms = ms / (double)repeats;
    return ok && ms > 0 ? 2.0 * (double)count * 16 * (double)repeats / (ms / 1e3) / 1e9 : 0;
}

}  // extern "C"
