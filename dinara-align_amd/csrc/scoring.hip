// The C ABI's substitution-table calls (dinara_gpu_scores, dinara_gpu_scoring_alignments).
#include "../../include/dinara.h"

extern "C" {

int64_t dinara_gpu_scores(int64_t, int64_t, const char *const *, const int64_t *, const char *const *,
                          const int64_t *, const dinara_scoring *, const dinara_mode *, int64_t, int64_t *) {
    return DINARA_INVALID_MODE;
}

int64_t dinara_gpu_scoring_alignments(int64_t, int64_t, const char *const *, const int64_t *, const char *const *,
                                      const int64_t *, const dinara_scoring *, const dinara_mode *, int64_t, int64_t,
                                      dinara_alignment *, int64_t *) {
    return DINARA_INVALID_MODE;
}

}  // extern "C"
