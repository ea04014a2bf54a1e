// Shared device/host definitions of the MI355X alignment kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dinara {

// A cell no path inside the band reaches. Every real cost stays below REAL_LIMIT (the host refuses pairs
// whose dearest path could pass it), so INF plus a few moves never wraps and never reads as real.
constexpr int INF = 1 << 28;
constexpr int REAL_LIMIT = 1 << 26;

// Gap-affine costs in the sweep's terms (reference lanes.mojo:187-259, modes.mojo:35-195): a deletion, a
// reference letter alone, of k letters costs del_open + k * del_ext; an insertion ins_open + k * ins_ext.
struct CostParams {
    int mismatch;
    int del_open, del_ext, ins_open, ins_ext;
    int del_first, ins_first;  // a gap's first letter: opening plus extension
    // Two-piece costs (reference lanes.mojo:187-345): a gap costs the less of its two pieces; with one piece
    // (pieces == 1) the second's fields are unused.
    int pieces;
    int del_open2, del_ext2, ins_open2, ins_ext2;
    int del_first2, ins_first2;
};

// Zero words after the tape's last group: a letter cursor may read two words past its sequence (64 words a word
// index across the transposed groups).
constexpr int TAPE_PADDING_WORDS = 4 * 64;

// Which pairs a launch takes and where their letters lie. The tape holds every sequence 2 or 8 bits a
// letter, 16 or 4 to a 32-bit word, transposed by groups of 64 slots: slot s (the s-th pair in length
// order) has word w of its reference at tape[group_base[s / 64] + w * 64 + s % 64], its query words after
// group_ref_words[s / 64] reference words.
struct Tape {
    const uint32_t *words;
    const int64_t *group_base;
    const int32_t *group_ref_words;
    const int32_t *rows;  // reference letters per slot
    const int32_t *cols;  // query letters per slot
    int bits;
};

// The sweep classes: four sixteen-lane-group bands (32, 64, 128, 256 diagonals, group.hip); wave-per-pair bands of
// 64Q (Q = 1..32, sweep.hip), which also take every pair under a user band; and block bands, one pair a block of G
// waves with S diagonal pairs a lane, 128SG diagonals (group.hip): 512 to 16384, in steps of at most a third.
constexpr int CLASS_GROUP32 = 0;
constexpr int CLASS_GROUP256 = 3;
constexpr int CLASS_WAVE_FIRST = 4;  // Q = 1, 2, 4, 8, 16, 32 for classes 4..9
constexpr int CLASS_WAVE_LAST = 9;
constexpr int CLASS_BLOCK_FIRST = 10;
constexpr int BLOCK_CLASSES = 12;
constexpr int CLASSES = CLASS_BLOCK_FIRST + BLOCK_CLASSES;
constexpr int CLASS_TOO_WIDE = -1;

// A block class's diagonal pairs a lane (S) and waves (G), block class b = cls - CLASS_BLOCK_FIRST: 512 (4 x 1),
// 1024 (8 x 1), 1536 (4 x 3), 2048 (8 x 2), 3072 (8 x 3), then 8 x 4..8 (4096..8192), 12288 (12 x 8), 16384 (16 x 8).
__host__ __device__ constexpr int block_s(int cls) {
    return (cls - CLASS_BLOCK_FIRST == 0 || cls - CLASS_BLOCK_FIRST == 2) ? 4
           : cls - CLASS_BLOCK_FIRST == 10                              ? 12
           : cls - CLASS_BLOCK_FIRST == 11                              ? 16
                                                                        : 8;
}
__host__ __device__ constexpr int block_waves(int cls) {
    return cls - CLASS_BLOCK_FIRST == 0   ? 1
           : cls - CLASS_BLOCK_FIRST == 1 ? 1
           : cls - CLASS_BLOCK_FIRST == 2 ? 3
           : cls - CLASS_BLOCK_FIRST == 3 ? 2
           : cls - CLASS_BLOCK_FIRST == 4 ? 3
           : cls - CLASS_BLOCK_FIRST <= 9 ? cls - CLASS_BLOCK_FIRST - 1
                                          : 8;
}

// The diagonals a class's band holds.
__host__ __device__ constexpr int band_width(int cls) {
    return cls <= CLASS_GROUP256    ? 32 << cls
           : cls <= CLASS_WAVE_LAST ? 64 << (cls - CLASS_WAVE_FIRST)
                                    : 128 * block_s(cls) * block_waves(cls);
}
__host__ __device__ constexpr bool band_is_wave(int cls) { return cls >= CLASS_WAVE_FIRST && cls <= CLASS_WAVE_LAST; }
__host__ __device__ constexpr bool band_is_block(int cls) { return cls >= CLASS_BLOCK_FIRST; }


// Everything a sweep launch reads and writes.
struct SweepArgs {
    Tape tape;
    const int32_t *list;  // slots to sweep, `count` of them
    int count;
    int cls;              // the launch's band class
    int slots;            // slots in the batch
    int32_t *want_low;    // per slot: the lane diagonals (query less reference) the band must hold
    int32_t *want_high;
    CostParams costs;
    int strict;           // 1 for alignments: no path off the band may cost as little
    int64_t cap;          // the cost cap, or a value past every cost
    int user_band;        // 1 when diagonals outside [user_low, user_high] (lane terms) are barred
    int doubling;         // 1: an unproven pair's next band is its band twice as wide, not the band its cost proves
    int user_low, user_high;
    int eqx;
    int forward_cigar;    // 1 when the tape holds both sequences reversed (Ties.LEFT): moves come in order
    int record;
    int lds_limit;        // letters one wave kernel's LDS holds
    // results per slot
    int32_t *cost;
    int32_t *status;      // 0 settled, 1 filed for a wider band, 2 wider than any band class
    int32_t *walk_low;    // the settled band's lowest diagonal and its class, for the walk
    int32_t *walk_class;
    // decisions: a group band's 4 bits a cell, a word for each diagonal pair every 8 anti-diagonals; a wave band's
    // a byte a cell, W bytes a row
    uint8_t *arena;          // this launch's decisions, `arena_stride` bytes a listed pair
    int64_t arena_stride;
    uint64_t *walk_flags;    // per slot: the address of its settled band's decisions
    int group_letters;       // LDS bytes a pair takes in a group launch (four pairs a block)
    int groups_fit;          // 1 when four pairs' letters fit one block's LDS
    int group_max_letters;   // the longest pair (both sequences) a group launch stages; longer pairs take blocks
    int group16_top;         // > 0: group launches of two slots a lane and more hold their steady cells as 16-bit
                             // halves, `group16_top` the unreachable value (letters then take 2 bytes each in LDS)
    int block_mask;          // bit c - CLASS_BLOCK_FIRST: block class c's letters fit its LDS and its costs its tags
    // CIGARs: each slot's scratch run of 2(n+m)+2 bytes from cigar_scratch_base, compacted into `cigars` at an
    // offset taken from `cigar_used`
    char *cigar_scratch;
    const int64_t *cigar_scratch_base;
    char *cigars;
    unsigned long long *cigar_used;
    int64_t *cigar_offset;
    int32_t *cigar_length;
    // unproven slots, filed by the class of the band that will prove them
    int32_t *retry_lists;  // CLASSES lists of `retry_capacity` slots each
    int32_t *retry_counts;
    int retry_capacity;
    // Free ends (wave sweeps only). Each slot sweeps a view of its pair: view_rows reference letters from
    // view_ref_at and view_cols query letters from view_qry_at, read back to front when `view_reversed`; free[4 s ..
    // 4 s + 3] the letters it may leave out for nothing at the start and end of the view's reference, then at those
    // of its query. A settled slot's end cell (the least cost; ties to the highest diagonal, or the lowest with
    // `lowest_end`) goes to end_row / end_col.
    int free_ends;
    int view_reversed;
    int lowest_end;
    const int32_t *view_rows, *view_cols, *view_ref_at, *view_qry_at;
    const int32_t *free;
    int32_t *end_row, *end_col;
};

// Host-side entry points of sweep.hip.
int class_width(int cls);
bool class_is_wave(int cls);
void launch_sweep(const SweepArgs &a, size_t lds, hipStream_t stream);
void launch_walk(const SweepArgs &a, hipStream_t stream);
void launch_group(int s, const SweepArgs &a, size_t lds, hipStream_t stream);
void launch_block(int s, int waves, const SweepArgs &a, size_t lds, hipStream_t stream);

}  // namespace dinara
