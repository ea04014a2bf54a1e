// The unit-cost path: global edit distance and its CIGAR by a banded bit-parallel sweep (edit.hip).
//
// A band of W = 64 x lanes x words diagonals is a bit-vector window over one column's rows that moves down one row
// a column, so bit i always stands for diagonal high - i. Each column takes one step of the bit-vector recurrence for
// edit distance (Myers 1999, with Hyyro's banded form): Pv/Mv hold the column's vertical differences (+1/-1),
// Ph/Mh the row's horizontal ones. The row above the window and the cell left of the window's new bottom row lie
// outside the band; the sweep stands in a value one more than the in-band neighbour they share (a +1 carry into the
// top, a +1 vertical difference under the bottom), which the diagonal move always matches or beats, so every in-band
// value is the least cost of a path inside the band. Rows above the matrix are the boundary's continuation (cost
// column - row) and rows below it feed nothing, so their letters may be anything.
//
// Pairs are binned by band width. A wave holds 64 / lanes pairs, each on `lanes` adjacent lanes, and every pair
// group takes its next pair from its bin's queue the moment it finishes the last one (a persistent launch): pairs of
// very different lengths share waves without waiting for each other. Every bin's first queue is swept at once, the
// bins side by side on streams of their own (a long pair's tail in one bin overlaps the others' work); a pair its
// band does not prove is filed into a retry queue of a wider bin. The retry queues are swept side by side too,
// filing into a second retry set, which is swept last, narrowest bin first. The launches have a fixed shape, so one
// hipGraph replays them.
#pragma once

#include "common.h"

namespace dinara {

// Bins: 1, 2, 3, 4, 6, 8, 12, 16, 21, 32 and 64 lanes of one 64-bit word each (64 to 4096 diagonals), then 64 lanes
// of 2 and 4 words. A lane keeps one word wherever it can, since the words of a lane are swept one after another and
// a long pair's column-after-column chain is what the run waits on. Between the powers of two, the lane counts fill
// most of a wave (3 lanes: 21 pairs a wave, 6: 10, 12: 5, 21: 3), so a band is seldom much wider than its pair asks:
// a wave's cost is its column steps, whatever its lanes hold.
constexpr int EDIT_BINS = 13;
__host__ __device__ constexpr int edit_lanes(int bin) {
    switch (bin) {
        case 0: return 1;
        case 1: return 2;
        case 2: return 3;
        case 3: return 4;
        case 4: return 6;
        case 5: return 8;
        case 6: return 12;
        case 7: return 16;
        case 8: return 21;
        case 9: return 32;
        default: return 64;
    }
}
__host__ __device__ constexpr int edit_words(int bin) { return bin < 11 ? 1 : (bin == 11 ? 2 : 4); }
__host__ __device__ constexpr int edit_width(int bin) { return 64 * edit_lanes(bin) * edit_words(bin); }
// Plain band doubling (the policy the learned band is measured against) keeps to the bins whose widths are powers of
// two: the first at least `need` wide from `from` on, and then the one twice as wide as `bin` (from a seeded band of
// another width, the first at least twice as wide; -1 past the widest).
__host__ __device__ constexpr bool edit_power_bin(int bin) { return (edit_width(bin) & (edit_width(bin) - 1)) == 0; }
__host__ __device__ inline int edit_power_bin_for(int need, int from) {
    for (int bin = from; bin < EDIT_BINS; ++bin)
        if (need <= edit_width(bin) && edit_power_bin(bin)) return bin;
    return -1;
}
__host__ __device__ inline int edit_doubled(int bin) {
    for (int next = bin + 1; next < EDIT_BINS; ++next)
        if (edit_width(next) >= 2 * edit_width(bin)) return next;  // from a power of two: exactly twice
    return -1;
}

struct EditArgs {
    SweepArgs s;            // tape, results, certificate policy (unit costs), CIGAR output
    int scale;              // the batch's costs are `scale` times unit costs
    int capacity;           // slots a list holds
    int32_t *band_low;      // per slot: the lowest diagonal of the band its next sweep takes
    const int32_t *first_band_low;  // per slot: its first band's, copied into `band_low` at the start of a run
    int32_t *lists;         // the queues a launch sweeps: EDIT_BINS of `capacity` slots
    int32_t *counts;        // [0, EDIT_BINS) queue lengths, [EDIT_BINS, 2 EDIT_BINS) queue heads
    int32_t *file_lists;    // the queues a launch files unproven pairs into (laid out as `lists`)
    int32_t *file_counts;
    int32_t *spare_counts;  // a further queue set's counts, emptied with `file_counts` at the start of a run
    int32_t *first_lists;   // the first sweep's queues, copied into `lists` at the start of a run
    const int32_t *first_counts;
    int32_t *first_counts_out;  // the same counts, written by the seeding kernel
    int32_t *first_low_out;     // the first bands' lowest diagonals, written by the seeding kernel
    // alignments: the band each settled pair is traced in, its decisions' place in the arena, and the queues of the
    // tracing sweep and of the walk
    int32_t *trace_lists;   // EDIT_BINS queues of `capacity` slots
    int32_t *trace_counts;  // lengths, then heads
    int64_t *trace_offset;  // per slot: byte offset of its decisions in the arena
    int32_t *walk_ready;    // per slot: 1 when its decisions are kept and it waits for the walk
    int32_t *walk_moves;    // per slot: the moves the last walk made (-1: not walked by it), for its CIGAR text
    int32_t *walk_early;    // per slot: 1 when it was ready at the first pass's end and walked while retries ran
    int32_t *deferred;      // settled pairs whose decisions found no room in the arena this pass
    int32_t *deferred_count;
    uint8_t *arena;
    unsigned long long *arena_used;
    unsigned long long arena_capacity;
    uint32_t *stats;        // sweeps begun by bin, then trace sweeps by bin, then pairs their first band left unproven,
                            // then pairs proven with their decisions kept
    unsigned long long *work;  // word-columns swept (a pair's columns times its band's words), by bin, then traced
    int one_pair;           // 1: each pair group sweeps one pair and stops (no refill), for comparison
    int keep_first;         // 1 (alignments): every sweep keeps its decisions, and a pair it proves walks them
};

// The bytes a pair of `cols` query letters traced in bin `bin` keeps: its columns eight at a time, each word's eight
// columns 128 contiguous bytes (two 64-bit planes a column), the pair's words side by side; one line more, for a
// run of lines that starts its columns at a place past the first.
__host__ __device__ inline unsigned long long edit_trace_bytes(int cols, int bin) {
    return (unsigned long long)((cols + 7) / 8 + 1) * edit_lanes(bin) * edit_words(bin) * 128ull;
}

// Pairs' longer sides split at 256, 1024, 4096 and 16384 letters: each class has its own least first bin, the
// tuned start of its band doubling.
constexpr int EDIT_LENGTH_CLASSES = 5;
__host__ __device__ inline int edit_length_class(int longer) {
    return longer < 256 ? 0 : longer < 1024 ? 1 : longer < 4096 ? 2 : longer < 16384 ? 3 : 4;
}

// The bin whose band holds `need` diagonals, at least `from`; -1 when none does.
__host__ __device__ inline int edit_bin_for(int need, int from) {
    for (int bin = from; bin < EDIT_BINS; ++bin)
        if (need <= edit_width(bin)) return bin;
    return -1;
}

// The widest bin a pair of `letters` letters (both sequences) can reach under either retry policy: a band placed
// inside the matrix's diagonals and letters + 3 wide proves any pair, and band doubling reaches the first power
// of two at least that wide.
__host__ __device__ inline int edit_widest(int letters) {
    const int need = letters + 3 < edit_width(EDIT_BINS - 1) ? letters + 3 : edit_width(EDIT_BINS - 1);
    int power = 64;
    while (power < need) power *= 2;
    const int a = edit_bin_for(need, 0), b = edit_bin_for(power, 0);
    return a > b ? a : b;
}

// The first pass's bin edges: `bins` (bit b for bin b) is the set of bins a first band may take. A pair's first
// band takes the narrowest bin of the set that holds `need` diagonals, from `from` on, and no wider than `widest`
// (the widest bin its launches reach); past it, the narrowest of all bins that holds it. Retries take every bin.
constexpr uint32_t EDIT_ALL_BINS = (1u << EDIT_BINS) - 1;
__host__ __device__ inline int edit_first_bin(int need, int from, uint32_t bins, int widest) {
    for (int bin = from; bin <= widest && bin < EDIT_BINS; ++bin)
        if (((bins >> bin) & 1u) && need <= edit_width(bin)) return bin;
    return edit_bin_for(need, from);
}

// Where a band of `width` diagonals holding [want_low, want_high] starts: centred, kept inside the matrix's
// diagonals (-rows - 1 .. cols + 1) where it can be.
__host__ __device__ inline int edit_place(int want_low, int want_high, int width, int rows, int cols) {
    int low = want_low - (width - (want_high - want_low + 1)) / 2;
    if (low + width - 1 > cols + 1) low = cols + 2 - width;
    if (low < -rows - 1) low = -rows - 1;
    if (low > want_low) low = want_low;
    if (low + width - 1 < want_high) low = want_high - width + 1;
    return low;
}

// Each length class's least first bin, and the bins first bands may take.
struct EditStart {
    int bin[EDIT_LENGTH_CLASSES];
    uint32_t bins;
};

// Host-side entry points of edit.hip. `grid` is the workgroups of each persistent launch (256 threads each).
void launch_edit_reset(const EditArgs &a, int slots, hipStream_t stream);
void launch_edit_sweep(const EditArgs &a, int bin, bool trace, int grid, hipStream_t stream);
// The walk, in two launches that may run at once: the first `long_slots` slots (those whose longer side is at least
// EDIT_WALK_LONG letters: slots are longest first) a wave a pair, and slots `first` .. `slots` - 1 a lane a pair.
// `early`: walk the slots ready at the first pass's end (copied by launch_edit_snapshot), while the retries sweep on
// other streams; it writes no flag those sweeps write. Otherwise walk every ready slot not walked early, leaving the
// early slots' moves as they are. Then, once both have ended, the CIGAR text of the slots they walked (the early
// walk's text may be made while the last walk runs).
void launch_edit_walk_long(const EditArgs &a, int long_slots, hipStream_t stream, bool early = false);
void launch_edit_walk_short(const EditArgs &a, int first, int slots, hipStream_t stream, bool early = false);
void launch_edit_text(const EditArgs &a, int slots, hipStream_t stream, bool early = false);
constexpr int EDIT_WALK_LONG = 4096;
void launch_edit_snapshot(const EditArgs &a, int slots, hipStream_t stream);
// Every slot's first band and first queue from predictions of log(1 + cost) in device memory (learned seeds).
void launch_edit_seed(const EditArgs &a, const float *prediction, double margin, double slope, const EditStart &start,
                      int slots, hipStream_t stream);
// Empties the arena and the trace and walk queues, then queues `count` deferred pairs (`held`, a copy of the
// deferred list) for a further tracing pass.
void launch_edit_requeue(const EditArgs &a, const int32_t *held, int count, hipStream_t stream);

}  // namespace dinara
