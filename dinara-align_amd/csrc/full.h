// The whole-matrix path (full.hip): a pair no band class takes (wider than the widest band, or too many letters for
// a kernel that stages them in LDS) is swept whole, strip after strip of columns in a pipeline of waves, and walked
// back by a whole wave.
#pragma once

#include "common.h"

namespace dinara {

// Columns a lane holds, and a strip (one wave) of them.
constexpr int FULL_C = 8;
constexpr int FULL_STRIP = 64 * FULL_C;
// The persistent sweep's workgroups (a wave each): strips beyond them are claimed as waves finish theirs.
constexpr int FULL_GRID = 2048;
// A wave's wait for its left strip's boundary column gives up after this many polls (the sweep then reports the
// pair unsettled instead of hanging).
constexpr int FULL_POLL_LIMIT = 1 << 22;

struct FullArgs {
    SweepArgs s;                 // the batch's tape, costs, results and CIGAR output
    int pairs;                   // the pairs this launch sweeps
    const int32_t *pair_slot;    // per pair: its slot of the tape
    const int32_t *strip_first;  // per pair: its first strip among the launch's (ceil(cols / FULL_STRIP), at least
                                 // one, a pair), then the launch's strips
    const int64_t *decision_at;  // per pair: its first decision record
    const int64_t *carry_at;     // per pair: its first carry word
    int record;                  // keep every cell's decision and walk it
    void *decisions;             // per pair, strips x rows x 64 lane records: 4 bytes (one piece) or 8 (two)
    unsigned long long *carry;   // per pair and strip, rows + 1 boundary records (two words with two pieces), zero =
                                 // not yet written
    int *claim;                  // the strips claimed so far (zero at launch)
    int *failed;                 // per pair: set when a wait gave up
};

// Host-side entry points of full.hip: the sweep of every listed pair's strips (costs and decisions, settling the
// slots), then the walks and CIGARs, a wave a pair.
void launch_full_sweep(const FullArgs &a, int strips, hipStream_t stream);
void launch_full_walk(const FullArgs &a, hipStream_t stream);
// Lists the slots whose status is not settled (0) into `list`, their number into `count` (zero at launch).
void launch_full_collect(const int32_t *status, int slots, int32_t *list, int32_t *count, hipStream_t stream);

}  // namespace dinara
