// The learned band proposal on MI355X (gfx950): k-mer features of a batch, and a small network that predicts each
// pair's cost from them, trained and run on the device by hand-written bf16 MFMA kernels with fp32 accumulation.
// No counterpart in the reference; the band it proposes is only a first guess: the band certificate still proves
// every result, and a pair it does not prove goes on to the band its cost proves, so every result stays exact.
//
// The network: 288 inputs (258 features and zero padding) -> 256 -> 256 -> 1, ReLU, bf16 weights and activations,
// fp32 accumulation and fp32 master weights updated by Adam. It predicts log(1 + cost).
//
// GEMM: one wave a 32 x 32 tile of C, v_mfma_f32_32x32x8_bf16 over K in steps of 8. Operands as the instruction
// reads them: lane l holds A[l % 32][4(l / 32) .. 4(l / 32) + 3] and B[4(l / 32) .. + 3][l % 32]; the 16 results
// of lane l are C[8(r / 4) + 4(l / 32) + r % 4][l % 32], r = 0..15. Either operand may be read transposed, so the
// backward products (X^T dZ, dZ W^T) use the same kernel. The products are small (256 wide); their tiles are read
// straight from global memory, which the caches serve.

#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "../../include/dinara.h"
#include "device.h"
#include "learned.h"

namespace dinara {

// ------------------------------------------------------------------ features

// 4-mer counts of 64 slots a workgroup, a thread a slot, in LDS (the thread's own 257-word row: the padding word
// spreads the rows over the banks). A 2-bit tape is read a word (16 letters) at a time: the 16 windows that end in
// the word come from the word and the one before it as 64 bits, and each is counted with an LDS add whose result
// is not waited for. The windows are counted by their tape bits (first letter lowest); a forward 4-mer code puts
// the first letter highest, so a forward tape's counts are read back with the four 2-bit fields reversed. The
// feature rows are then written a slot at a time by the whole wave, 64 consecutive floats a store.
__global__ __launch_bounds__(64) void kmer_features(Tape tape, int slots, int reversed, float *out) {
    constexpr int ROW = 257;
    __shared__ int32_t hist[64 * ROW];
    const int t = threadIdx.x, slot = blockIdx.x * 64 + t;
    int32_t *mine = hist + t * ROW;
    for (int k = 0; k < 256; ++k) mine[k] = 0;
    if (slot < slots) {
        const int g = slot >> 6, col = slot & 63;
        const uint32_t *ref_words = tape.words + tape.group_base[g] + col;
        const uint32_t *qry_words = ref_words + (size_t)tape.group_ref_words[g] * 64;
        const int bits = tape.bits, per = 32 / bits;
        auto count = [&](const uint32_t *words, int len, int delta) {
            if (bits == 2) {
                // Eight words a step, the next eight loaded while these are counted.
                const int total = (len + 15) >> 4;
                auto load = [&](uint32_t (&to)[8], int w0) {
#pragma unroll
                    for (int k = 0; k < 8; ++k) to[k] = w0 + k < total ? words[(size_t)(w0 + k) * 64] : 0u;
                };
                uint32_t now[8], next[8];
                load(now, 0);
                uint32_t before = 0;
                for (int w0 = 0; w0 < total; w0 += 8) {
                    load(next, w0 + 8);
#pragma unroll
                    for (int k = 0; k < 8; ++k) {
                        const int w = w0 + k;
                        const uint64_t pair = (uint64_t)before | ((uint64_t)now[k] << 32);
                        // Window j starts at letter 16w + j - 3 and ends at letter 16w + j.
                        const int first = w == 0 ? 3 : 0, last = len - 16 * w;
#pragma unroll
                        for (int j = 0; j < 16; ++j) {
                            if (j >= first && j < last) atomicAdd(&mine[(uint32_t)((pair >> (26 + 2 * j)) & 255u)], delta);
                        }
                        before = now[k];
                    }
#pragma unroll
                    for (int k = 0; k < 8; ++k) now[k] = next[k];
                }
                return;
            }
            uint32_t window = 0;  // the last four letters, first lowest
            int seen = 0;
            for (int w = 0; w * per < len; ++w) {
                uint32_t word = words[(size_t)w * 64];
                const int here = min(per, len - w * per);
                for (int i = 0; i < here; ++i) {
                    window = (window >> 2) | ((word & 3u) << 6);
                    word >>= bits;
                    if (++seen >= 4) atomicAdd(&mine[window], delta);
                }
            }
        };
        count(ref_words, tape.rows[slot], 1);
        count(qry_words, tape.cols[slot], -1);
    }
    __syncthreads();
    // A reversed tape holds each sequence back to front: its windows' tape bits are then the forward codes.
    const int first = blockIdx.x * 64, here = min(64, slots - first);
    for (int s = 0; s < here; ++s) {
        const int n = tape.rows[first + s], m = tape.cols[first + s];
        const float scale = 8.0f / sqrtf(1.0f + (float)n + (float)m);
        float *row = out + (size_t)(first + s) * FEATURE_STRIDE;
        for (int k = t; k < FEATURE_STRIDE; k += 64) {
            float v = 0.0f;
            if (k < 256) {
                const int at = reversed ? k : ((k & 3) << 6) | ((k & 12) << 2) | ((k & 48) >> 2) | (k >> 6);
                v = (float)hist[s * ROW + at] * scale;
            } else if (k == 256) {
                v = (logf(1.0f + (float)n) - 7.0f) * 0.5f;
            } else if (k == 257) {
                v = (logf(1.0f + (float)m) - 7.0f) * 0.5f;
            }
            row[k] = v;
        }
    }
}

void launch_features(const Tape &tape, int slots, int reversed, float *out, hipStream_t stream) {
    if (slots <= 0) return;
    hipLaunchKernelGGL(kmer_features, dim3((slots + 63) / 64), dim3(64), 0, stream, tape, slots, reversed, out);
}

// ------------------------------------------------------------------ bf16 and the MFMA product

// fp32 to bf16, round to nearest even (finite values; the network never makes others).
__device__ inline uint16_t to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ inline float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// C (M x N) = A (M x K) . B (K x N) + bias, in 32 x 32 tiles; A read as A[i][k] = a[i * lda + k] (or a[k * lda + i]
// transposed), B as B[k][j] = b[k * ldb + j] (or b[j * ldb + k]). Writes the fp32 product to `c` when non-null and
// its bf16 to `cb` when non-null: after ReLU when `relu`, zeroed where mask[i][j] <= 0 when `mask` is non-null.
template <bool TA, bool TB>
__global__ __launch_bounds__(64) void gemm_bf16(const uint16_t *a, int lda, const uint16_t *b, int ldb, int K,
                                                const float *bias, int relu, const float *mask, int ldm, float *c,
                                                int ldc, uint16_t *cb, int ldcb) {
    const int lane = threadIdx.x;
    const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
    const int row = lane & 31, kq = (lane >> 5) * 4;
    f32x16 acc = {};
    for (int k0 = 0; k0 < K; k0 += 8) {
        bf16x4 va, vb;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k = k0 + kq + j;
            va[j] = (short)(TA ? a[(size_t)k * lda + i0 + row] : a[(size_t)(i0 + row) * lda + k]);
            vb[j] = (short)(TB ? b[(size_t)(j0 + row) * ldb + k] : b[(size_t)k * ldb + j0 + row]);
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(va, vb, acc, 0, 0, 0);
    }
    const int j = j0 + row;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int i = i0 + 8 * (r / 4) + 4 * (lane >> 5) + (r % 4);
        float v = acc[r] + (bias ? bias[j] : 0.0f);
        if (c) c[(size_t)i * ldc + j] = v;
        if (cb) {
            if (relu) v = v > 0.0f ? v : 0.0f;
            if (mask && mask[(size_t)i * ldm + j] <= 0.0f) v = 0.0f;
            cb[(size_t)i * ldcb + j] = to_bf16(v);
        }
    }
}

template <bool TA, bool TB>
static void gemm(hipStream_t stream, const uint16_t *a, int lda, const uint16_t *b, int ldb, int M, int N, int K,
                 const float *bias, int relu, const float *mask, int ldm, float *c, int ldc, uint16_t *cb,
                 int ldcb) {
    hipLaunchKernelGGL((gemm_bf16<TA, TB>), dim3(M / 32, N / 32), dim3(64), 0, stream, a, lda, b, ldb, K, bias,
                       relu, mask, ldm, c, ldc, cb, ldcb);
}

// ------------------------------------------------------------------ prediction in one pass

// (rows x cols) bf16 into (cols x rows).
__global__ void transpose_bf16(const uint16_t *in, int rows, int cols, uint16_t *out) {
    const int at = blockIdx.x * blockDim.x + threadIdx.x;
    if (at >= rows * cols) return;
    out[(size_t)(at % cols) * rows + at / cols] = in[at];
}

constexpr int PREDICT_HIDDEN = 256;

// The network's prediction for 32 rows a wave, both hidden layers and the output in one pass: the fp32 features
// are rounded to bf16 as they are read, each layer is 8 MFMA tiles of 32 x 32 (the weights read transposed, 4
// bf16 a lane), and a layer's activations go through LDS to become the next product's left operand. The products
// take their K steps in the order `gemm_bf16` takes them, so the prediction is the same as the layer-by-layer one.
__global__ __launch_bounds__(64) void predict_rows(const float *features, int rows, const uint16_t *w1t,
                                                   const float *b1, const uint16_t *w2t, const float *b2,
                                                   const uint16_t *w3, const float *b3, float *y) {
    __shared__ uint16_t act[32 * PREDICT_HIDDEN];
    const int lane = threadIdx.x, row = lane & 31, kq = (lane >> 5) * 4;
    const int i0 = blockIdx.x * 32;
    f32x16 acc[8];
#pragma unroll
    for (int nt = 0; nt < 8; ++nt) acc[nt] = f32x16{};
    const bool real = i0 + row < rows;
    const float *x = features + (size_t)(real ? i0 + row : 0) * FEATURE_STRIDE;
    for (int k0 = 0; k0 < FEATURE_STRIDE; k0 += 8) {
        const float4 f = real ? *reinterpret_cast<const float4 *>(x + k0 + kq) : make_float4(0.f, 0.f, 0.f, 0.f);
        bf16x4 va;
        va[0] = (short)to_bf16(f.x);
        va[1] = (short)to_bf16(f.y);
        va[2] = (short)to_bf16(f.z);
        va[3] = (short)to_bf16(f.w);
#pragma unroll
        for (int nt = 0; nt < 8; ++nt) {
            const bf16x4 vb = *reinterpret_cast<const bf16x4 *>(w1t + (size_t)(nt * 32 + row) * FEATURE_STRIDE + k0 + kq);
            acc[nt] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(va, vb, acc[nt], 0, 0, 0);
        }
    }
    // Bias, ReLU and bf16 into LDS, row-major: MFMA result r of lane l is row 8(r/4) + 4(l/32) + r%4, column l%32.
    auto keep = [&](const float *bias) {
#pragma unroll
        for (int nt = 0; nt < 8; ++nt) {
            const int j = nt * 32 + row;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int i = 8 * (r / 4) + 4 * (lane >> 5) + (r % 4);
                const float v = acc[nt][r] + bias[j];
                act[i * PREDICT_HIDDEN + j] = to_bf16(v > 0.0f ? v : 0.0f);
            }
            acc[nt] = f32x16{};
        }
    };
    keep(b1);
    __syncthreads();
    for (int k0 = 0; k0 < PREDICT_HIDDEN; k0 += 8) {
        const bf16x4 va = *reinterpret_cast<const bf16x4 *>(act + row * PREDICT_HIDDEN + k0 + kq);
#pragma unroll
        for (int nt = 0; nt < 8; ++nt) {
            const bf16x4 vb = *reinterpret_cast<const bf16x4 *>(w2t + (size_t)(nt * 32 + row) * PREDICT_HIDDEN + k0 + kq);
            acc[nt] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(va, vb, acc[nt], 0, 0, 0);
        }
    }
    __syncthreads();
    keep(b2);
    __syncthreads();
    // The output: one lane a row, summed in column order like `head`.
    if (lane < 32 && real) {
        float sum = b3[0];
        for (int k = 0; k < PREDICT_HIDDEN; ++k) sum += from_bf16(act[row * PREDICT_HIDDEN + k]) * from_bf16(w3[k]);
        y[i0 + row] = sum;
    }
}

// ------------------------------------------------------------------ the network's small kernels

__global__ void to_bf16_kernel(const float *in, uint16_t *out, int64_t count) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) out[i] = to_bf16(in[i]);
}

// Rows `index[r]` of `in` (FEATURE_STRIDE floats each) as bf16 rows of `out`; rows past `count` zero.
__global__ void gather_rows(const float *in, const int32_t *index, int count, uint16_t *out, int rows) {
    const int r = blockIdx.x, k = threadIdx.x;
    if (r >= rows || k >= FEATURE_STRIDE) return;
    const float v = r < count ? in[(size_t)index[r] * FEATURE_STRIDE + k] : 0.0f;
    out[(size_t)r * FEATURE_STRIDE + k] = to_bf16(v);
}

// The output layer and the loss: y = A2 . w3 + b3 per row; with targets, dy = 2 (y - t) / count (zero past
// `count`), the row's squared error into `sq`, and dZ2 = dy w3 where Z2 > 0, as bf16.
__global__ void head(const uint16_t *a2, const uint16_t *w3, const float *b3, int hidden, int rows, int count,
                     const float *targets, float *y, float *dy, float *sq, const float *z2, uint16_t *dz2) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    float s = b3[0];
    for (int k = 0; k < hidden; ++k) s += from_bf16(a2[(size_t)r * hidden + k]) * from_bf16(w3[k]);
    if (y) y[r] = s;
    if (!targets) return;
    const float d = r < count ? 2.0f * (s - targets[r]) / (float)count : 0.0f;
    dy[r] = d;
    sq[r] = r < count ? (s - targets[r]) * (s - targets[r]) : 0.0f;
    for (int k = 0; k < hidden; ++k)
        dz2[(size_t)r * hidden + k] = to_bf16(z2[(size_t)r * hidden + k] > 0.0f ? d * from_bf16(w3[k]) : 0.0f);
}

// Column sums: out[j] = sum over rows of x[r][j] (bf16 x, or fp32 v[r] * bf16 x when `weights`).
__global__ void column_sums(const uint16_t *x, int rows, int cols, const float *weights, float *out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cols) return;
    float s = 0.0f;
    for (int r = 0; r < rows; ++r) s += (weights ? weights[r] : 1.0f) * from_bf16(x[(size_t)r * cols + j]);
    out[j] = s;
}

__global__ void sum_kernel(const float *x, int count, float *out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    float s = 0.0f;
    for (int i = 0; i < count; ++i) s += x[i];
    out[0] = s;
}

// One Adam step on `count` parameters: m, v the moments, `scale` = lr * sqrt(1 - b2^t) / (1 - b1^t).
__global__ void adam(float *w, const float *g, float *m, float *v, int64_t count, float b1, float b2, float scale,
                     float eps, uint16_t *wb) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - scale * mi / (sqrtf(vi) + eps);
    w[i] = wi;
    if (wb) wb[i] = to_bf16(wi);
}

}  // namespace dinara

using namespace dinara;

// ------------------------------------------------------------------ the network and its C ABI

struct dinara_mlp {
    DeviceContext *ctx = nullptr;
    int hidden = 256;
    int rows_cap = 0;   // rows the work buffers hold
    int64_t step = 0;   // Adam's step count
    // fp32 master weights, their bf16 copies, gradients, Adam moments (one flat buffer each: w1, b1, w2, b2, w3, b3)
    DeviceBuffer master, bf, grad, m1, m2;
    // per-step work
    DeviceBuffer x, z1, a1, z2, a2, y, dy, sq, dz2, dz1, targets, index, loss, features;
    DeviceBuffer identity;  // 0, 1, 2, ... for gathers of consecutive rows
    DeviceBuffer transposed;  // w1 and w2 as bf16, transposed, for the one-pass prediction
    DeviceBuffer predicted;   // the one-pass prediction's output for `dinara_mlp_predict`
    int identity_rows = 0;
    size_t off_w1, off_b1, off_w2, off_b2, off_w3, off_b3, total;
};

namespace {

using Net = dinara_mlp;

float *f32(DeviceBuffer &b, size_t offset = 0) { return static_cast<float *>(b.ptr) + offset; }
uint16_t *h16(DeviceBuffer &b, size_t offset = 0) { return static_cast<uint16_t *>(b.ptr) + offset; }

bool reserve_rows(Net *net, int rows) {
    if (rows <= net->rows_cap) return true;
    const size_t r = (size_t)rows, h = (size_t)net->hidden;
    if (!net->x.reserve(r * FEATURE_STRIDE * 2) || !net->z1.reserve(r * h * 4) || !net->a1.reserve(r * h * 2) ||
        !net->z2.reserve(r * h * 4) || !net->a2.reserve(r * h * 2) || !net->y.reserve(r * 4) ||
        !net->dy.reserve(r * 4) || !net->sq.reserve(r * 4) || !net->dz2.reserve(r * h * 2) ||
        !net->dz1.reserve(r * h * 2) || !net->targets.reserve(r * 4) || !net->loss.reserve(16))
        return false;
    net->rows_cap = rows;
    return true;
}

// Forward over `rows` rows of bf16 inputs in net->x: Z1, A1, Z2, A2 and y.
void forward(Net *net, int rows, hipStream_t s) {
    const int h = net->hidden;
    gemm<false, false>(s, h16(net->x), FEATURE_STRIDE, h16(net->bf, net->off_w1), h, rows, h, FEATURE_STRIDE,
                       f32(net->master, net->off_b1), 1, nullptr, 0, f32(net->z1), h, h16(net->a1), h);
    gemm<false, false>(s, h16(net->a1), h, h16(net->bf, net->off_w2), h, rows, h, h, f32(net->master, net->off_b2),
                       1, nullptr, 0, f32(net->z2), h, h16(net->a2), h);
}

// Forward and backward over `rows` rows (the first `count` real), targets in net->targets: the gradients into
// net->grad and the summed squared error into net->loss[0].
void forward_backward(Net *net, int rows, int count, hipStream_t s) {
    const int h = net->hidden;
    forward(net, rows, s);
    hipLaunchKernelGGL(head, dim3((rows + 255) / 256), dim3(256), 0, s, h16(net->a2), h16(net->bf, net->off_w3),
                       f32(net->master, net->off_b3), h, rows, count, f32(net->targets), f32(net->y), f32(net->dy),
                       f32(net->sq), f32(net->z2), h16(net->dz2));
    hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(1), 0, s, f32(net->sq), rows, f32(net->loss));
    // Output layer: dw3 = A2^T dy, db3 = sum dy.
    hipLaunchKernelGGL(column_sums, dim3((h + 255) / 256), dim3(256), 0, s, h16(net->a2), rows, h, f32(net->dy),
                       f32(net->grad, net->off_w3));
    hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(1), 0, s, f32(net->dy), rows, f32(net->grad, net->off_b3));
    // Second layer: dW2 = A1^T dZ2, db2 = sum dZ2, dZ1 = (dZ2 W2^T) where Z1 > 0.
    gemm<true, false>(s, h16(net->a1), h, h16(net->dz2), h, h, h, rows, nullptr, 0, nullptr, 0,
                      f32(net->grad, net->off_w2), h, nullptr, 0);
    hipLaunchKernelGGL(column_sums, dim3((h + 255) / 256), dim3(256), 0, s, h16(net->dz2), rows, h, nullptr,
                       f32(net->grad, net->off_b2));
    gemm<false, true>(s, h16(net->dz2), h, h16(net->bf, net->off_w2), h, rows, h, h, nullptr, 0, f32(net->z1), h,
                      nullptr, 0, h16(net->dz1), h);
    // First layer: dW1 = X^T dZ1, db1 = sum dZ1.
    gemm<true, false>(s, h16(net->x), FEATURE_STRIDE, h16(net->dz1), h, FEATURE_STRIDE, h, rows, nullptr, 0,
                      nullptr, 0, f32(net->grad, net->off_w1), h, nullptr, 0);
    hipLaunchKernelGGL(column_sums, dim3((h + 255) / 256), dim3(256), 0, s, h16(net->dz1), rows, h, nullptr,
                       f32(net->grad, net->off_b1));
}

void refresh_bf16(Net *net, hipStream_t s) {
    hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)((net->total + 255) / 256)), dim3(256), 0, s,
                       f32(net->master), h16(net->bf), (int64_t)net->total);
}

// Uploads `rows` fp32 feature rows (FEATURE_STRIDE each) as bf16 into net->x, padded to a multiple of 32.
bool upload_rows(Net *net, int64_t rows, const float *features, int *padded) {
    const int p = (int)((rows + 31) / 32 * 32);
    if (!reserve_rows(net, std::max(p, 32))) return false;
    if (!net->features.reserve((size_t)std::max<int64_t>(rows, 1) * FEATURE_STRIDE * 4)) return false;
    if (rows > 0 && !net->features.write(net->ctx, features, (size_t)rows * FEATURE_STRIDE * 4)) return false;
    std::vector<int32_t> idx((size_t)std::max<int64_t>(rows, 1));
    for (int64_t i = 0; i < rows; ++i) idx[i] = (int32_t)i;
    if (!net->index.upload(net->ctx, idx.data(), idx.size() * 4)) return false;
    hipLaunchKernelGGL(gather_rows, dim3(p), dim3(FEATURE_STRIDE), 0, net->ctx->stream, f32(net->features),
                       static_cast<const int32_t *>(net->index.ptr), (int)rows, h16(net->x), p);
    *padded = p;
    return true;
}

}  // namespace

int64_t mlp_predict_device(dinara_mlp *net, const float *features, int64_t rows, float *out) {
    if (!net) return DINARA_INVALID_MODE;
    if (rows <= 0) return 0;
    DeviceGuard guard(net->ctx);
    hipStream_t s = net->ctx->stream;
    const int64_t chunk = 1 << 16;
    if (!reserve_rows(net, (int)chunk) || !net->index.reserve((size_t)chunk * 4)) return DINARA_OUT_OF_MEMORY;
    std::vector<int32_t> idx((size_t)chunk);
    for (int64_t first = 0; first < rows; first += chunk) {
        const int64_t count = std::min(chunk, rows - first);
        const int padded = (int)((count + 31) / 32 * 32);
        for (int64_t i = 0; i < count; ++i) idx[i] = (int32_t)(first + i);
        if (!net->index.write(net->ctx, idx.data(), (size_t)count * 4)) return DINARA_OUT_OF_MEMORY;
        hipLaunchKernelGGL(gather_rows, dim3(padded), dim3(FEATURE_STRIDE), 0, s, features,
                           static_cast<const int32_t *>(net->index.ptr), (int)count, h16(net->x), padded);
        forward(net, padded, s);
        hipLaunchKernelGGL(head, dim3((padded + 255) / 256), dim3(256), 0, s, h16(net->a2), h16(net->bf, net->off_w3),
                           f32(net->master, net->off_b3), net->hidden, padded, (int)count, nullptr, f32(net->y),
                           nullptr, nullptr, nullptr, nullptr);
        if (!net->y.read(net->ctx, out + first, (size_t)count * 4)) return DINARA_OUT_OF_MEMORY;
    }
    return 0;
}

int64_t mlp_predict_into(dinara_mlp *net, const float *features, int64_t rows, float *out) {
    if (!net) return DINARA_INVALID_MODE;
    if (rows <= 0) return 0;
    DeviceGuard guard(net->ctx);
    hipStream_t s = net->ctx->stream;
    if (net->hidden == PREDICT_HIDDEN && rows <= INT32_MAX / FEATURE_STRIDE) {
        const size_t w1 = (size_t)FEATURE_STRIDE * PREDICT_HIDDEN, w2 = (size_t)PREDICT_HIDDEN * PREDICT_HIDDEN;
        if (!net->transposed.reserve((w1 + w2) * 2)) return DINARA_OUT_OF_MEMORY;
        uint16_t *w1t = h16(net->transposed), *w2t = w1t + w1;
        hipLaunchKernelGGL(transpose_bf16, dim3((unsigned)((w1 + 255) / 256)), dim3(256), 0, s,
                           h16(net->bf, net->off_w1), FEATURE_STRIDE, PREDICT_HIDDEN, w1t);
        hipLaunchKernelGGL(transpose_bf16, dim3((unsigned)((w2 + 255) / 256)), dim3(256), 0, s,
                           h16(net->bf, net->off_w2), PREDICT_HIDDEN, PREDICT_HIDDEN, w2t);
        hipLaunchKernelGGL(predict_rows, dim3((unsigned)((rows + 31) / 32)), dim3(64), 0, s, features, (int)rows, w1t,
                           f32(net->master, net->off_b1), w2t, f32(net->master, net->off_b2),
                           h16(net->bf, net->off_w3), f32(net->master, net->off_b3), out);
        return hipPeekAtLastError() == hipSuccess ? 0 : DINARA_OUT_OF_MEMORY;
    }
    const int chunk = 1 << 16;
    if (!reserve_rows(net, chunk)) return DINARA_OUT_OF_MEMORY;
    if (net->identity_rows < chunk) {
        std::vector<int32_t> idx((size_t)chunk);
        for (int i = 0; i < chunk; ++i) idx[i] = i;
        if (!net->identity.upload(net->ctx, idx.data(), idx.size() * 4)) return DINARA_OUT_OF_MEMORY;
        net->identity_rows = chunk;
    }
    for (int64_t first = 0; first < rows; first += chunk) {
        const int count = (int)std::min<int64_t>(chunk, rows - first);
        const int padded = (count + 31) / 32 * 32;
        hipLaunchKernelGGL(gather_rows, dim3(padded), dim3(FEATURE_STRIDE), 0, s, features + first * FEATURE_STRIDE,
                           static_cast<const int32_t *>(net->identity.ptr), count, h16(net->x), padded);
        forward(net, padded, s);
        hipLaunchKernelGGL(head, dim3((padded + 255) / 256), dim3(256), 0, s, h16(net->a2), h16(net->bf, net->off_w3),
                           f32(net->master, net->off_b3), net->hidden, padded, count, nullptr, f32(net->y),
                           nullptr, nullptr, nullptr, nullptr);
        if (hipMemcpyAsync(out + first, net->y.ptr, (size_t)count * 4, hipMemcpyDeviceToDevice, s) != hipSuccess)
            return DINARA_OUT_OF_MEMORY;
    }
    return hipPeekAtLastError() == hipSuccess ? 0 : DINARA_OUT_OF_MEMORY;
}

extern "C" {

dinara_mlp *dinara_mlp_new(int64_t gpu, uint64_t seed) {
    DeviceContext *ctx = device_context(gpu);
    if (!ctx) return nullptr;
    DeviceGuard guard(ctx);
    auto *net = new Net();
    net->ctx = ctx;
    const size_t h = (size_t)net->hidden;
    net->off_w1 = 0;
    net->off_b1 = net->off_w1 + (size_t)FEATURE_STRIDE * h;
    net->off_w2 = net->off_b1 + h;
    net->off_b2 = net->off_w2 + h * h;
    net->off_w3 = net->off_b2 + h;
    net->off_b3 = net->off_w3 + h;
    net->total = net->off_b3 + 1;
    // He-normal weights for the ReLU layers (the padding inputs' rows zero), a small output layer, zero biases.
    std::vector<float> w(net->total, 0.0f);
    std::mt19937_64 rng(seed);
    std::normal_distribution<float> n1(0.0f, std::sqrt(2.0f / FEATURES)), n2(0.0f, std::sqrt(2.0f / (float)h)),
        n3(0.0f, std::sqrt(1.0f / (float)h));
    for (size_t k = 0; k < (size_t)FEATURES; ++k)
        for (size_t j = 0; j < h; ++j) w[net->off_w1 + k * h + j] = n1(rng);
    for (size_t k = 0; k < h * h; ++k) w[net->off_w2 + k] = n2(rng);
    for (size_t k = 0; k < h; ++k) w[net->off_w3 + k] = n3(rng);
    w[net->off_b3] = 7.0f;  // log(1 + cost) of a typical pair, where training starts
    const size_t bytes = net->total * 4;
    std::vector<float> zero(net->total, 0.0f);
    if (!net->master.upload(ctx, w.data(), bytes) || !net->bf.reserve(net->total * 2) || !net->grad.reserve(bytes) ||
        !net->m1.upload(ctx, zero.data(), bytes) || !net->m2.upload(ctx, zero.data(), bytes)) {
        delete net;
        return nullptr;
    }
    refresh_bf16(net, ctx->stream);
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
        delete net;
        return nullptr;
    }
    return net;
}

void dinara_mlp_free(dinara_mlp *net) {
    if (!net) return;
    DeviceGuard guard(net->ctx);
    delete net;
}

int64_t dinara_mlp_parameter_count(dinara_mlp *net) { return net ? (int64_t)net->total : 0; }

// The fp32 master weights, flat: w1 (288 x 256, row k input k), b1 (256), w2 (256 x 256), b2, w3 (256), b3 (1).
int64_t dinara_mlp_get_weights(dinara_mlp *net, float *out) {
    if (!net) return DINARA_INVALID_MODE;
    DeviceGuard guard(net->ctx);
    return net->master.read(net->ctx, out, net->total * 4) ? 0 : DINARA_OUT_OF_MEMORY;
}

int64_t dinara_mlp_set_weights(dinara_mlp *net, const float *in) {
    if (!net) return DINARA_INVALID_MODE;
    DeviceGuard guard(net->ctx);
    if (!net->master.write(net->ctx, in, net->total * 4)) return DINARA_OUT_OF_MEMORY;
    refresh_bf16(net, net->ctx->stream);
    return hipStreamSynchronize(net->ctx->stream) == hipSuccess ? 0 : DINARA_OUT_OF_MEMORY;
}

// The gradients of the mean squared error over `rows` rows (features FEATURE_STRIDE floats a row, targets one a
// row) at the current weights, flat as the weights, and the loss.
int64_t dinara_mlp_gradients(dinara_mlp *net, int64_t rows, const float *features, const float *targets,
                             float *gradients, double *loss) {
    if (!net || rows <= 0) return DINARA_INVALID_MODE;
    DeviceGuard guard(net->ctx);
    int padded = 0;
    if (!upload_rows(net, rows, features, &padded)) return DINARA_OUT_OF_MEMORY;
    std::vector<float> t((size_t)padded, 0.0f);
    std::memcpy(t.data(), targets, (size_t)rows * 4);
    if (!net->targets.write(net->ctx, t.data(), t.size() * 4)) return DINARA_OUT_OF_MEMORY;
    forward_backward(net, padded, (int)rows, net->ctx->stream);
    float sq = 0;
    if (!net->grad.read(net->ctx, gradients, net->total * 4) || !net->loss.read(net->ctx, &sq, 4))
        return DINARA_OUT_OF_MEMORY;
    if (loss) *loss = sq / (double)rows;
    return 0;
}

// Predictions (log(1 + cost)) for `rows` rows.
int64_t dinara_mlp_predict(dinara_mlp *net, int64_t rows, const float *features, float *out) {
    if (!net) return DINARA_INVALID_MODE;
    if (rows <= 0) return 0;
    DeviceGuard guard(net->ctx);
    if (net->hidden == PREDICT_HIDDEN && rows <= INT32_MAX / FEATURE_STRIDE) {  // the one-pass prediction
        const size_t bytes = (size_t)rows * FEATURE_STRIDE * 4;
        if (!net->features.reserve(bytes) || !net->features.write(net->ctx, features, bytes) ||
            !net->predicted.reserve((size_t)rows * 4))
            return DINARA_OUT_OF_MEMORY;
        if (const int64_t code = mlp_predict_into(net, f32(net->features), rows, f32(net->predicted))) return code;
        return net->predicted.read(net->ctx, out, (size_t)rows * 4) ? 0 : DINARA_OUT_OF_MEMORY;
    }
    const int64_t chunk = 1 << 16;
    for (int64_t first = 0; first < rows; first += chunk) {
        const int64_t count = std::min(chunk, rows - first);
        int padded = 0;
        if (!upload_rows(net, count, features + first * FEATURE_STRIDE, &padded)) return DINARA_OUT_OF_MEMORY;
        forward(net, padded, net->ctx->stream);
        hipLaunchKernelGGL(head, dim3((padded + 255) / 256), dim3(256), 0, net->ctx->stream, h16(net->a2),
                           h16(net->bf, net->off_w3), f32(net->master, net->off_b3), net->hidden, padded, count,
                           nullptr, f32(net->y), nullptr, nullptr, nullptr, nullptr);
        if (!net->y.read(net->ctx, out + first, (size_t)count * 4)) return DINARA_OUT_OF_MEMORY;
    }
    return 0;
}

// Trains on `rows` rows for `epochs` passes in shuffled batches of `batch` rows (a multiple of 32) by Adam at
// learning rate `lr`: every step's forward, backward and update on the device. Writes the last epoch's mean loss.
int64_t dinara_mlp_train(dinara_mlp *net, int64_t rows, const float *features, const float *targets, int64_t epochs,
                         int64_t batch, double lr, uint64_t seed, double *final_loss) {
    if (!net || rows <= 0 || batch <= 0 || batch % 32 != 0) return DINARA_INVALID_MODE;
    DeviceContext *ctx = net->ctx;
    DeviceGuard guard(ctx);
    hipStream_t s = ctx->stream;
    if (!reserve_rows(net, (int)batch)) return DINARA_OUT_OF_MEMORY;
    DeviceBuffer all_targets;
    if (!net->features.upload(ctx, features, (size_t)rows * FEATURE_STRIDE * 4) ||
        !all_targets.upload(ctx, targets, (size_t)rows * 4) || !net->index.reserve((size_t)rows * 4))
        return DINARA_OUT_OF_MEMORY;
    std::vector<int32_t> order((size_t)rows);
    for (int64_t i = 0; i < rows; ++i) order[i] = (int32_t)i;
    std::mt19937_64 rng(seed);
    std::vector<float> batch_targets((size_t)batch);
    double epoch_loss = 0;
    const float b1 = 0.9f, b2 = 0.999f;
    for (int64_t e = 0; e < epochs; ++e) {
        std::shuffle(order.begin(), order.end(), rng);
        if (!net->index.write(ctx, order.data(), order.size() * 4)) return DINARA_OUT_OF_MEMORY;
        std::vector<float> shuffled((size_t)rows);
        for (int64_t i = 0; i < rows; ++i) shuffled[i] = targets[order[i]];
        DeviceBuffer shuffled_targets;
        if (!shuffled_targets.upload(ctx, shuffled.data(), shuffled.size() * 4)) return DINARA_OUT_OF_MEMORY;
        double sum = 0;
        int64_t seen = 0;
        for (int64_t first = 0; first < rows; first += batch) {
            const int count = (int)std::min<int64_t>(batch, rows - first);
            hipLaunchKernelGGL(gather_rows, dim3((unsigned)batch), dim3(FEATURE_STRIDE), 0, s, f32(net->features),
                               static_cast<const int32_t *>(net->index.ptr) + first, count, h16(net->x), (int)batch);
            if (hipMemcpyAsync(net->targets.ptr, f32(shuffled_targets) + first, (size_t)count * 4,
                               hipMemcpyDeviceToDevice, s) != hipSuccess)
                return DINARA_OUT_OF_MEMORY;
            if (count < batch && hipMemsetAsync(f32(net->targets) + count, 0, (size_t)(batch - count) * 4, s) != hipSuccess)
                return DINARA_OUT_OF_MEMORY;
            forward_backward(net, (int)batch, count, s);
            ++net->step;
            const double bias1 = 1.0 - std::pow((double)b1, (double)net->step);
            const double bias2 = 1.0 - std::pow((double)b2, (double)net->step);
            const float scale = (float)(lr * std::sqrt(bias2) / bias1);
            hipLaunchKernelGGL(adam, dim3((unsigned)((net->total + 255) / 256)), dim3(256), 0, s, f32(net->master),
                               f32(net->grad), f32(net->m1), f32(net->m2), (int64_t)net->total, b1, b2, scale, 1e-8f,
                               h16(net->bf));
            float sq = 0;
            if (!net->loss.read(ctx, &sq, 4)) return DINARA_OUT_OF_MEMORY;
            sum += sq;
            seen += count;
        }
        epoch_loss = sum / (double)seen;
    }
    if (final_loss) *final_loss = epoch_loss;
    return hipStreamSynchronize(s) == hipSuccess ? 0 : DINARA_OUT_OF_MEMORY;
}

}  // extern "C"
