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
    SweepArgs s;             // the batch's tape, costs, results and CIGAR output
    int slot;                // the pair (a slot of the tape)
    int rows, cols;          // its reference and query letters
    int strips;              // ceil(cols / FULL_STRIP), at least one
    int record;              // keep every cell's decision and walk it
    void *decisions;         // strips x rows x 64 lane records: 4 bytes (one piece) or 8 (two) of 4-bit / 8-bit flags
    unsigned long long *carry;  // per strip, rows + 1 boundary records (two words with two pieces), zero = not yet
    int *claim;              // the strips claimed so far (zero at launch)
    int *failed;             // set when a wait gave up
};

// Host-side entry points of full.hip: the sweep (cost and decisions, settling the slot), then the walk and CIGAR.
void launch_full_sweep(const FullArgs &a, hipStream_t stream);
void launch_full_walk(const FullArgs &a, hipStream_t stream);
// Lists the slots whose status is not settled (0) into `list`, their number into `count` (zero at launch).
void launch_full_collect(const int32_t *status, int slots, int32_t *list, int32_t *count, hipStream_t stream);

}  // namespace dinara
