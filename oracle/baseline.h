/*
 * The CPU baseline: a C restatement of the reference's CPU batch path for global gap-affine alignments
 * (reference lanes.mojo: inter-sequence SIMD over a band its cost proves, then a second pass, DP flags and the
 * walk back for Ties.LEFT). It is timed beside the GPU as the reported CPU baseline and, as a second CPU
 * implementation, cross-checked against the oracle. Test and measurement infrastructure only; the product
 * never links it.
 *
 * Differences from the reference, stated: 16-bit lanes only (the reference first tries bytes, 64 pairs to an
 * AVX-512 register, then 16 bits), 32 pairs a group, every flag of a group's band kept (the reference keeps a
 * row every sqrt(rows) and sweeps stretches again), and a shared atomic group counter for the threads.
 */
#ifndef DINARA_BASELINE_H
#define DINARA_BASELINE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Aligns `pairs` pairs under costs (mismatch, open, extend) on `threads` threads: each pair's cost into costs[i]
 * (or -1 when a lane cannot hold it) and, when `cigars` is non-null, its Ties.LEFT CIGAR NUL-terminated at
 * cigars + i * cigar_stride. Returns the pairs settled. */
int64_t baseline_alignments(int64_t pairs, const char *const *references, const int64_t *reference_lengths,
                            const char *const *queries, const int64_t *query_lengths, int64_t mismatch,
                            int64_t open, int64_t extend, int64_t threads, int64_t *costs, char *cigars,
                            int64_t cigar_stride);

#ifdef __cplusplus
}
#endif

#endif
