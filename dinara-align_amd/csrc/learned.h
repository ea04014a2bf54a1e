// The learned band proposal: k-mer features of a resident batch and the small network that predicts each pair's
// cost from them (learned.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.h"

namespace dinara {

constexpr int FEATURES = 258;        // 256 4-mer count differences (reference less query), 2 log-lengths
constexpr int FEATURE_STRIDE = 288;  // a pair's features padded to whole MFMA steps (a multiple of 32)

// Writes each slot's features, fp32, FEATURE_STRIDE a slot, into `out` (device memory), from the batch's tape:
// `reversed` when the tape holds both sequences back to front.
void launch_features(const Tape &tape, int slots, int reversed, float *out, hipStream_t stream);

}  // namespace dinara

struct dinara_mlp;

// The network's predictions (log(1 + cost)) for `rows` rows of fp32 features already in device memory, into `out`
// (host memory). 0, or a DINARA_* code.
int64_t mlp_predict_device(dinara_mlp *net, const float *features, int64_t rows, float *out);

// The same into `out` in device memory, ordered on the network's stream with no wait for the host.
int64_t mlp_predict_into(dinara_mlp *net, const float *features, int64_t rows, float *out);
