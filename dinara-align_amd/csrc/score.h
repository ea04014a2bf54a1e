// Shared device/host definitions of the substitution-table (Scoring) kernels, score.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dinara {

// Below every reachable score: the host refuses pairs whose scores could come within a quarter of the 32-bit
// range (reference scoring.mojo:216-238, ScoreReach), so the sentinel plus a few steps never wraps or reads real.
constexpr int SCORE_NEG = -(1 << 28);

// What one launch of score_wave computes.
enum ScoreMode : int {
    SCORE_GLOBAL = 0,  // the corner's score
    SCORE_LOCAL = 1,   // the best cell, floored at zero, and where it lies
    SCORE_REACH = 2,   // over both prefixes reversed from a local end: the first cell reaching a target
    SCORE_TRACE = 3,   // the corner's score and every cell's decision, for the walk
};

// The columns a lane holds: a strip of 64C columns a wave, C = 1..8.
constexpr int SCORE_MAX_C = 8;
constexpr int SCORE_STRIP_MAX = 64 * SCORE_MAX_C;

// A reference too long to stage whole in LDS is read through a ring of this many chunks of this many letters.
constexpr int SCORE_RING_CHUNK = 256;
constexpr int SCORE_RING_CHUNKS = 4;

struct ScoreArgs {
    const uint8_t *codes;    // every sequence's letter codes (the alphabet's places)
    const int64_t *ref_at;   // per slot: where its reference's codes start
    const int64_t *qry_at;   // per slot: where its query's codes start
    const int32_t *view;     // per slot: reference begin, reference letters, query begin, query letters
    int reversed;            // the views read both sequences back to front from their ends
    int ring;                // the references go through an LDS ring rather than being staged whole
    const int32_t *list;     // the slots of this launch, `count` of them
    int count;
    const int8_t *table;     // table[a * 32 + b]: reference letter a against query letter b
    int letters;
    int open, extend;        // a gap's first letter, each later one (scores, zero or less)
    int mode;
    int end_bits;            // SCORE_LOCAL: > 0 keys each column's best as score << end_bits | row (rows and scores
                             // fit the 32 bits), 0 compares and selects
    int c;                   // columns a lane
    int32_t *score;          // per slot
    int32_t *end_row;        // SCORE_LOCAL: the best cell's row and column; SCORE_REACH: the found cell's
    int32_t *end_col;
    const int32_t *target;   // SCORE_REACH: per slot, the score to reach
    uint32_t *arena;         // SCORE_TRACE: `arena_stride` words a listed pair
    int64_t arena_stride;
    uint64_t *walk_base;     // SCORE_TRACE: per slot, the address of its decisions
    int32_t *carry;          // the last column of a strip, two ints a row, `carry_stride` ints a listed pair
    int64_t carry_stride;
    // The walk's CIGARs (SCORE_TRACE only)
    int eqx;
    char *op_scratch;        // per slot, the walk's moves 2 bits each
    const int64_t *op_base;
    char *cigars;
    unsigned long long *cigar_used;
    int64_t *cigar_offset;
    int32_t *cigar_length;
    int slots;
};

// Host-side entry points of score.hip. `lds` is the bytes a block stages: the profile, then the reference's letters
// (or, with `ring`, the ring they pass through).
size_t score_lds(int letters, int rows, bool ring);
void launch_score(const ScoreArgs &a, size_t lds, hipStream_t stream);
void launch_score_walk(const ScoreArgs &a, hipStream_t stream);

}  // namespace dinara
