// The unit-cost path's kernels: the persistent banded bit-parallel sweep, the walk over its 2-bit decisions, and
// the small kernels that reset and requeue its lists (the scheme is described in edit.h).
//
// Reference: the unit-cost distance is the reference's lanes path at costs (1, 0, 1) (lanes.mojo:950-1004) and its
// device kernel's bit-parallel recurrence (device_edit.mojo:29-77); the CIGAR follows the reference's bit-parallel
// traceback rule (traceback.mojo:400-575): from the corner, a substitution when one is optimal, else a reference
// letter alone, else a query letter alone, and two equal letters only when no edit is; Ties.LEFT walks both
// sequences reversed (edit_distance.mojo:295-297). At unit costs that rule and the reference's DP-flag rule
// (lanes.mojo:905-923) give the same CIGAR.
#include <type_traits>

#include "bands.h"
#include "edit.h"

namespace dinara {
namespace {

// Lane t receives lane t - 1's value (FROM_LOWER) or lane t + 1's; the wave's end lane receives 0 (its callers
// replace the end lane's value, so no copy of the old value is needed).
template <bool FROM_LOWER>
__device__ __forceinline__ uint32_t wave_neighbour(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, FROM_LOWER ? 0x138 : 0x130, 0xF, 0xF, true);
}

// The even bits of a 32-bit word packed into its low 16.
__device__ __forceinline__ uint32_t even_bits(uint32_t x) {
    x &= 0x55555555u;
    x = (x | (x >> 1)) & 0x33333333u;
    x = (x | (x >> 2)) & 0x0F0F0F0Fu;
    x = (x | (x >> 4)) & 0x00FF00FFu;
    return (x | (x >> 8)) & 0x0000FFFFu;
}

// Bit b of byte q of a 32-bit word, for q = 0..3, packed into its low 4 bits.
__device__ __forceinline__ uint32_t byte_bits(uint32_t x, int b) {
    x = (x >> b) & 0x01010101u;
    return (x | (x >> 7) | (x >> 14) | (x >> 21)) & 0xFu;
}

__device__ __forceinline__ void place(uint64_t &plane, uint64_t bits, int shift) {
    if (shift >= 64 || shift <= -64) return;
    plane |= shift >= 0 ? bits << shift : bits >> -shift;
}

// The code planes of 64 letters of one sequence from letter `first` on: plane b bit i is bit b of letter
// first + i's code, 0 for letters outside 0..n-1. `text` holds word w of the sequence at text[64 w].
template <int BITS>
__device__ __forceinline__ void load_planes(const uint32_t *text, int n, int first, uint64_t (&planes)[BITS]) {
    constexpr int PER = 32 / BITS;
#pragma unroll
    for (int b = 0; b < BITS; ++b) planes[b] = 0;
    const int w0 = first >= 0 ? first / PER : -((PER - 1 - first) / PER);
    const int w1 = (first + 63) >= 0 ? (first + 63) / PER : -1;
    for (int w = w0 < 0 ? 0 : w0; w <= w1 && w * PER < n; ++w) {
        const uint32_t word = text[(size_t)w * 64];
        const int valid = n - w * PER < PER ? n - w * PER : PER;
        const uint64_t keep = (uint64_t(1) << valid) - 1;
        const int shift = w * PER - first;
        if constexpr (BITS == 2) {
            place(planes[0], even_bits(word) & keep, shift);
            place(planes[1], even_bits(word >> 1) & keep, shift);
        } else {
#pragma unroll
            for (int b = 0; b < BITS; ++b) place(planes[b], byte_bits(word, b) & keep, shift);
        }
    }
}

// One sequence read a letter at a time from letter `from` >= 0 on: its code, then the next's. `cur` holds the current
// word shifted so its next code is in the low bits; the word after it is loaded a word ahead, so a column never waits
// on its own load. The refill load has no bounds check, so its result lands straight in the register it is kept in
// (a checked load's result would be copied there at once, the copy waiting for the load): a cursor reads at most two
// words past its sequence's last, which the tape's padding and the next group's words hold, and letters past the end
// are never used (columns stop at the query's end; rows below the reference's feed no row above them). A lane with
// no pair is parked: it never loads. All unsigned: no sign fix-ups in the column step.
template <int BITS>
struct Cursor {
    static constexpr uint32_t PER = 32 / BITS;
    static constexpr uint32_t MASK = BITS == 8 ? 0xFFu : 3u;
    const uint32_t *text;
    uint32_t cur, ahead, left, next_word;

    __device__ void park(const uint32_t *) {
        text = nullptr;
        cur = ahead = 0;
        left = 0xFFFFFFFFu;  // never refills
        next_word = 0;
    }
    __device__ void begin(const uint32_t *t, int from, int) {
        text = t;
        const uint32_t w = (uint32_t)from / PER, r = (uint32_t)from % PER;
        cur = text[(size_t)w * 64] >> (BITS * r);
        left = PER - r;
        ahead = text[(size_t)(w + 1) * 64];
        next_word = w + 2;
    }
    __device__ uint32_t next() {
        const uint32_t code = cur & MASK;
        cur >>= BITS;
        if (--left == 0) {
            cur = ahead;
            left = PER;
            ahead = text[(size_t)next_word * 64];
            ++next_word;
        }
        return code;
    }
    __device__ void refill() {}  // refills on its own, letter by letter
};

// One sequence of 2-bit letters read a letter a column, refilled only at the steps the whole wave shares (its
// column steps t with t mod 8 = 7, the ends of its eight-column periods): `now` holds the letters up to the period's
// end, low bits first; at the period's last column the next eight are cut from the two words loaded at the period's
// start (`word0`, `word1`, the words that hold letters `ahead` .. `ahead` + 7), and the words for the eight after
// are loaded. No lane branches on its own position and no load is used fewer than eight columns after it is made,
// so a column never waits on memory (the stores the wave makes at a period's end included: a wait for a load also
// waits for every store made before it). A refill reads at most one word past the one holding letter `ahead` + 7,
// where `ahead` is at most one past the last letter a column used; a parked lane stays on the tape's first words.
struct Letters2 {
    const uint32_t *text;
    uint32_t now, word0, word1, ahead, live;

    __device__ void park(const uint32_t *tape) {
        text = tape;
        now = word0 = word1 = ahead = live = 0;
    }
    // Letters from `from` on, the first taken at a column whose place in its period is `place`.
    __device__ void begin(const uint32_t *t, uint32_t from, int place) {
        text = t;
        const uint32_t w = from / 16;
        now = __builtin_amdgcn_alignbit(text[(size_t)(w + 1) * 64], text[(size_t)w * 64], 2 * (from % 16));
        ahead = from + (uint32_t)(8 - place);
        live = 8;
        word0 = text[(size_t)(ahead / 16) * 64];
        word1 = text[(size_t)(ahead / 16 + 1) * 64];
    }
    __device__ uint32_t next() {
        const uint32_t code = now & 3u;
        now >>= 2;
        return code;
    }
    // At the period's last column, after its letter is taken.
    __device__ void refill() {
        now = __builtin_amdgcn_alignbit(word1, word0, 2 * (ahead % 16));
        ahead += live;
        word0 = text[(size_t)(ahead / 16) * 64];
        word1 = text[(size_t)(ahead / 16 + 1) * 64];
    }
};

template <int BITS>
using LetterSource = typename std::conditional<BITS == 2, Letters2, Cursor<BITS>>::type;


// A plane of K 64-bit words shifted down one bit (bit i takes bit i + 1), each word's bit 63 taken from the next
// word's bit 0 and the last word's from bit 0 of `incoming`: two alignbits a word.
template <int K>
__device__ __forceinline__ void slide_down(uint64_t (&x)[K], uint32_t incoming) {
#pragma unroll
    for (int kk = 0; kk < K; ++kk) {
        const uint32_t lo = (uint32_t)x[kk], hi = (uint32_t)(x[kk] >> 32);
        const uint32_t in = kk + 1 < K ? (uint32_t)x[kk + 1 < K ? kk + 1 : kk] : incoming;
        const uint32_t new_lo = __builtin_amdgcn_alignbit(hi, lo, 1);
        const uint32_t new_hi = __builtin_amdgcn_alignbit(in, hi, 1);
        x[kk] = ((uint64_t)new_hi << 32) | new_lo;
    }
}

// Where the HOLD columns HOLD q .. HOLD q + HOLD - 1 of word w of a pair's `words` words lie in its kept decisions:
// eight columns of a word take 128 contiguous bytes (two 64-bit planes a column), the words of an eight-column block
// side by side; HOLD is 8 (a whole block) or 4 (half of one).
template <int HOLD>
__device__ __forceinline__ size_t held_chunk_at(int q, int words, int w) {
    return HOLD == 8 ? ((size_t)q * words + w) * 128 : ((size_t)(q >> 1) * words + w) * 128 + (size_t)(q & 1) * 64;
}

// ------------------------------------------------------------------ settling and filing (first lanes only)

__device__ void settle(const EditArgs &a, int slot, int n, int m, int64_t found) {
    const SweepArgs &s = a.s;
    s.cost[slot] = (int32_t)(found * a.scale);
    s.status[slot] = 0;
    if (!s.record || found > s.cap) return;  // distances, or above the cap: no CIGAR asked
    if (n == 0 || m == 0) {                  // all gaps: nothing to trace
        s.walk_low[slot] = 0;
        s.walk_class[slot] = 0;
        a.walk_ready[slot] = 1;
        return;
    }
    // The narrowest band the strict certificate proves at the found cost: it holds every optimal path.
    const int end = m - n;
    const int64_t target = found + 1;
    int h = end > 0 ? end : 0, l = end < 0 ? end : 0;
    while (h < m && off_band(s.costs, h + 1, end) < target) ++h;
    while (l > -n && off_band(s.costs, l - 1, end) < target) --l;
    const int bin = edit_bin_for(h - l + 1, 0);
    const int low = edit_place(l, h, edit_width(bin), n, m);
    s.walk_low[slot] = low;
    s.walk_class[slot] = bin;
    const unsigned long long bytes = edit_trace_bytes(m, bin);
    const unsigned long long at = atomicAdd(a.arena_used, bytes);
    if (at + bytes > a.arena_capacity) {
        a.deferred[atomicAdd(a.deferred_count, 1)] = slot;
        return;
    }
    a.trace_offset[slot] = (int64_t)at;
    a.trace_lists[(size_t)bin * a.capacity + atomicAdd(&a.trace_counts[bin], 1)] = slot;
}

__device__ void refile(const EditArgs &a, int slot, int n, int m, int lo, int hi, int want_low, int want_high,
                       int bin) {
    if (a.s.status[slot] < 0) atomicAdd(&a.stats[2 * EDIT_BINS], 1u);  // left unproven by its first band
    if (a.s.doubling) {
        // Plain band doubling: the next bin, twice as wide, centred on the band just swept.
        want_low = lo;
        want_high = hi;
    }
    const int next = a.s.doubling ? edit_doubled(bin) : edit_bin_for(want_high - want_low + 1, bin + 1);
    if (next < 0) {
        a.s.status[slot] = 2;
        return;
    }
    a.band_low[slot] = edit_place(want_low, want_high, edit_width(next), n, m);
    a.s.status[slot] = 1;
    a.file_lists[(size_t)next * a.capacity + atomicAdd(&a.file_counts[next], 1)] = slot;
}

// ------------------------------------------------------------------ the sweep

// The sweep's three uses: a pass that proves bands (SWEEP), the second pass of an alignment batch that sweeps each
// settled pair's traced band again keeping its decisions (TRACE), and a pass that proves bands and keeps the
// decisions of each band as it goes (SWEEP_TRACED), so a pair its first band proves goes straight to the walk.
constexpr int SWEEP = 0, TRACE = 1, SWEEP_TRACED = 2;

// Bin `bin`'s persistent sweep: L lanes of K words a pair, letters BITS a code. Kept decisions are 2 bits a cell, the
// walk's first true choice in the order two unequal letters (3: both planes), a reference letter alone (1: plane 0),
// a query letter alone (2: plane 1), two equal letters (0), in two bit-planes: word w's columns 8q..8q+7 in the 128
// bytes at arena + offset + (q words + w) 128, 16 a column.
//
// A lane of one word (K = 1) writes whole lines in step with its wave: a pair claimed when its wave has made t0
// column steps starts its columns at place `phase` = t0 mod 8 of its first line (its column j at place phase + j of
// the line run), so every lane of the wave fills a line at the same step (t mod 8 = 7). Each column goes to LDS; at
// those steps the wave writes the lanes' lines together, eight lanes a line, eight whole lines a store instruction,
// and a pair ending mid-line writes its last line itself. The phase rides in the low bits of the pair's decision
// offset (which is a multiple of 128). Lanes of more words hold their last four columns in registers and write
// their 64 bytes at once, phase 0.
// Sixteen bytes of kept decisions in global memory, stored through an address held as an integer (a global store,
// not a flat one).
typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));
typedef u64x2 __attribute__((address_space(1))) global_decisions;

__device__ __forceinline__ u64x2 planes_of(ulonglong2 v) { return u64x2{v.x, v.y}; }

template <int L, int K, int BITS, int MODE>
__global__ __launch_bounds__(256) void edit_sweep(EditArgs a, int bin) {
    constexpr bool TRACE = MODE == 1;          // sweeps settled pairs' traced bands again
    constexpr bool RECORD = MODE != 0;         // keeps decisions
    constexpr int W = 64 * L * K;
    const int lane = threadIdx.x & 63;
    // A wave holds 64 / L pair groups of L lanes; the lanes past the last whole group never take a pair.
    const int k = lane % L;
    const bool in_group = lane < (64 / L) * L;
    const bool first_lane = k == 0 && in_group, last_lane = k == L - 1 && in_group;
    const unsigned long long inner_lanes = ~__ballot(last_lane);  // lanes whose carry goes on to the next lane
    const int32_t *queue = (TRACE ? a.trace_lists : a.lists) + (size_t)bin * a.capacity;
    int32_t *counts = TRACE ? a.trace_counts : a.counts;
    const int count = counts[bin];
    if (count <= 0) return;  // an empty queue (the tracing pass, when every pair kept its decisions)
    int32_t *head = counts + EDIT_BINS + bin;
    // Narrow bins hold many short pairs a wave: a fixed priority by width. Waves of 16 lanes a pair and more set
    // theirs from the longest pair they have left, at each claim (below).
    if constexpr (L < 16) {
        if (W >= 512)
            __builtin_amdgcn_s_setprio(2);
        else if (W >= 128)
            __builtin_amdgcn_s_setprio(1);
    }

    int slot = -1;
    bool drained = false;
    int n = 0, m = 0, lo = 0, hi = 0, j = 0, top = 0;
    uint64_t pv[K], mv[K], pl[BITS][K];
    LetterSource<BITS> bottom, query;
    bottom.park(a.s.tape.words);
    query.park(a.s.tape.words);
    uint8_t *trace = nullptr;
    bool recording = false;    // this group keeps the current pair's decisions
    long long trace_at = 0;    // and where, from the arena's start
    uint32_t line = 0;         // one-word lanes: the lane's first line, in 128-byte lines from the arena's start
                               // (the arena is under 2^39 bytes: at most 0.6 of the device's memory)
    constexpr int WL = L * K;  // the pair's words
    // One word a lane: columns staged in LDS, eight places a lane, lane l's place p at index l ^ p of place row p, so
    // the wave's line reads (eight lanes a line, one place each) fall in different banks with no padding: 32 KB a
    // workgroup, five to a compute unit.
    constexpr bool STAGED = RECORD && K == 1;
    __shared__ ulonglong2 stage[STAGED ? 4 : 1][STAGED ? 8 : 1][STAGED ? 64 : 1];
    ulonglong2 (*const places)[STAGED ? 64 : 1] = stage[STAGED ? threadIdx.x >> 6 : 0];
    int t = 0;       // the wave's column steps so far (the same in every lane)
    int t0 = 0;      // the step count when the lane's pair was claimed
    // More words a lane: the last HOLD columns' decision planes in registers, oldest first.
    constexpr int HOLD = STAGED ? 1 : 4;
    uint64_t held0[HOLD][K], held1[HOLD][K];
#pragma unroll
    for (int q = 0; q < HOLD; ++q)
#pragma unroll
        for (int kk = 0; kk < K; ++kk) held0[q][kk] = held1[q][kk] = 0;
#pragma unroll
    for (int kk = 0; kk < K; ++kk) {
        pv[kk] = mv[kk] = 0;
#pragma unroll
        for (int b = 0; b < BITS; ++b) pl[b][kk] = 0;
    }

    while (true) {
        // Pair groups with nothing to sweep claim the next pairs: one atomic a wave.
        const bool asking = first_lane && slot < 0 && !drained;
        if (a.one_pair) drained = true;  // a group asks once
        const unsigned long long askers = __ballot(asking);
        if (askers) {
            const int leader = __builtin_ctzll(askers);
            int base = 0;
            // A drained queue is seen by a plain read (the head only grows): no atomic for it.
            if (lane == leader) {
                base = __hip_atomic_load(head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (base < count) base = atomicAdd(head, __popcll(askers));
            }
            base = __shfl(base, leader);
            const int mine = base + __popcll(askers & ((1ull << lane) - 1ull));
            const int idx = __shfl(mine, lane - k);
            const bool group_asked = __shfl((int)asking, lane - k) != 0;
            if (group_asked) {
                if (idx < count) {
                    slot = queue[idx];
                    n = a.s.tape.rows[slot];
                    m = a.s.tape.cols[slot];
                    lo = TRACE ? a.s.walk_low[slot] : a.band_low[slot];
                    hi = lo + W - 1;
                    j = 0;
                    top = 0;
                    const int g = slot >> 6;
                    const uint32_t *ref = a.s.tape.words + a.s.tape.group_base[g] + (slot & 63);
                    const uint32_t *qry = ref + (size_t)a.s.tape.group_ref_words[g] * 64;
#pragma unroll
                    for (int kk = 0; kk < K; ++kk) {
                        const int w = k * K + kk;
                        // Column 0: cost |row|, so +1 below row 0 and -1 from it up (rows above the matrix).
                        const int t = hi + 1 - 64 * w;
                        pv[kk] = t <= 0 ? ~0ull : t >= 64 ? 0ull : ~0ull << t;
                        mv[kk] = ~pv[kk];
                        uint64_t planes[BITS];
                        load_planes<BITS>(ref, n, 64 * w - hi - 1, planes);
#pragma unroll
                        for (int b = 0; b < BITS; ++b) pl[b][kk] = planes[b];
                    }
                    bottom.begin(ref, -lo, t & 7);
                    query.begin(qry, 0, t & 7);
                    t0 = t;
                    if (TRACE) {
                        trace_at = a.trace_offset[slot] & ~127ll;
                        recording = true;
                        // The walk reads where this pass puts the pair's columns.
                        if (first_lane) a.trace_offset[slot] = trace_at | (STAGED ? (t0 & 7) : 0);
                    } else if (RECORD) {
                        // Room for this band's decisions, if the arena has it; else the pair sweeps without.
                        long long at = -1;
                        if (first_lane && m > 0) {
                            const unsigned long long bytes = edit_trace_bytes(m, bin);
                            const unsigned long long got = atomicAdd(a.arena_used, bytes);
                            if (got + bytes <= a.arena_capacity) at = (long long)got;
                        }
                        const int at_lo = __shfl((int)(at & 0xFFFFFFFF), lane - k);
                        const int at_hi = __shfl((int)(at >> 32), lane - k);
                        trace_at = ((long long)at_hi << 32) | (long long)(unsigned)at_lo;
                        recording = trace_at >= 0;
                    }
                    trace = recording ? a.arena + trace_at : a.arena;
                    line = recording ? (uint32_t)(trace_at >> 7) + (uint32_t)k : 0u;
                    if (first_lane) {
                        atomicAdd(&a.stats[(TRACE ? EDIT_BINS : 0) + bin], 1u);
                        atomicAdd(&a.work[(TRACE ? EDIT_BINS : 0) + bin], (unsigned long long)m * (L * K));
                    }
                } else {
                    drained = true;
                }
            }
            // The wave's longest pair left, column after column, is the run's longest chain when it is long: such
            // waves issue first where they share a SIMD (set again at each claim, as pairs end and new ones start).
            if constexpr (L >= 16) {
                int left = slot >= 0 ? m - j : 0;
#pragma unroll
                for (int d = 1; d < 64; d <<= 1) {
                    const int other = __shfl_xor(left, d);
                    left = other > left ? other : left;
                }
                if (left > 12000)
                    __builtin_amdgcn_s_setprio(3);
                else if (left > 4000)
                    __builtin_amdgcn_s_setprio(2);
                else if (left > 1000)
                    __builtin_amdgcn_s_setprio(1);
                else
                    __builtin_amdgcn_s_setprio(0);
            }
        }
        if (!__ballot(slot >= 0)) break;

        // Pairs at their last column: the cost at (n, m), then settle, file for a wider band, or queue the walk.
        const bool finishing = slot >= 0 && j == m;
        if (__ballot(finishing)) {
            const int i_n = hi - (m - n);  // the window bit of row n
            int part = 0;
#pragma unroll
            for (int kk = 0; kk < K; ++kk) {
                const int w = k * K + kk;
                const int from = (64 * w > 1 ? 64 * w : 1) - 64 * w;
                const int to = (i_n < 64 * w + 63 ? i_n : 64 * w + 63) - 64 * w;
                if (from <= to) {
                    const uint64_t mask = (to == 63 ? ~0ull : ((1ull << (to + 1)) - 1ull)) & (~0ull << from);
                    part += __popcll(pv[kk] & mask) - __popcll(mv[kk] & mask);
                }
            }
            // The group's sum, at its first lane (the only one that uses it).
#pragma unroll
            for (int d = 1; d < L; d <<= 1) {
                const int other = __shfl_down(part, d);
                part += k + d < L ? other : 0;
            }
            const int found = hi + m + top + part;  // `top` is the first lane's
            if constexpr (STAGED) {
                // A pair ending before its line's last place: that line from LDS, by its own lane.
                const int last = (t - 1) & 7;  // the place of its last column
                if (finishing && recording && m > 0 && last != 7) {
                    uint8_t *line = trace + ((size_t)(((t - 1) >> 3) - (t0 >> 3)) * WL + k) * 128;
                    for (int q = 0; q <= last; ++q)
                        reinterpret_cast<global_decisions *>((uint64_t)(uintptr_t)line)[q] =
                            planes_of(places[q][lane ^ q]);
                }
            } else if constexpr (RECORD) {
                // The last, partly filled columns: shifted down to the chunk's start, then written whole.
                if (finishing && recording && (m % HOLD)) {
                    for (int pad = m % HOLD; pad < HOLD; ++pad) {
#pragma unroll
                        for (int kk = 0; kk < K; ++kk) {
#pragma unroll
                            for (int q = 0; q + 1 < HOLD; ++q) {
                                held0[q][kk] = held0[q + 1][kk];
                                held1[q][kk] = held1[q + 1][kk];
                            }
                        }
                    }
#pragma unroll
                    for (int kk = 0; kk < K; ++kk) {
                        global_decisions *chunk = reinterpret_cast<global_decisions *>((uint64_t)(uintptr_t)trace + held_chunk_at<HOLD>(m / HOLD, WL, k * K + kk));
#pragma unroll
                        for (int q = 0; q < HOLD; ++q) chunk[q] = u64x2{held0[q][kk], held1[q][kk]};
                    }
                }
            }
            if (first_lane && finishing) {
                if (TRACE) {
                    // The traced band was proven in the first pass: its cost must agree.
                    if ((int64_t)found * a.scale != (int64_t)a.s.cost[slot]) a.s.status[slot] = 3;
                    a.walk_ready[slot] = 1;
                } else {
                    int want_low = 0, want_high = 0;
                    if (!certify(a.s, n, m, found, lo, hi, &want_low, &want_high)) {
                        refile(a, slot, n, m, lo, hi, want_low, want_high, bin);
                    } else if (RECORD && recording && found <= a.s.cap) {
                        // Proven with its decisions kept: walk this band.
                        a.s.cost[slot] = (int32_t)(found * a.scale);
                        a.s.status[slot] = 0;
                        a.s.walk_low[slot] = lo;
                        a.s.walk_class[slot] = bin;
                        a.trace_offset[slot] = trace_at | (STAGED ? (t0 & 7) : 0);
                        a.walk_ready[slot] = 1;
                        atomicAdd(&a.stats[2 * EDIT_BINS + 1], 1u);
                    } else {
                        settle(a, slot, n, m, found);
                    }
                }
            }
            if (finishing) {
                slot = -1;
                bottom.park(a.s.tape.words);
                query.park(a.s.tape.words);
            }
            continue;
        }

        // Columns until the wave's next pair end: no group claims or finishes before then.
        int steps = slot >= 0 ? m - j : 0x7FFFFFFF;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int other = __shfl_xor(steps, d);
            steps = other < steps ? other : steps;
        }
        // Every lane holds the same count: a scalar loop counter.
        steps = __builtin_amdgcn_readfirstlane(steps);
        const bool active = slot >= 0;
        for (int step = 0; step < steps; ++step) {
        // One column. The window moves down a row: each word takes the bit of the word below; the band's bottom
        // takes +1 under it (Pv), and the new bottom row's letter.
        ++j;
        const int place = t & 7;  // this column's place in its line (every lane alike)
        ++t;
        if constexpr (BITS == 2) {
            // Each plane slides down a row: its top word's bit 63 comes from bit 0 of the next lane's lowest word (one
            // DPP move a plane), or at the band's bottom from the stand-ins (+1 vertical, no -1) and the new bottom
            // row's letter; two alignbits shift a 64-bit word and take that bit in.
            uint32_t ipv = L > 1 ? wave_neighbour<false>((uint32_t)pv[0]) : 0u;
            uint32_t imv = L > 1 ? wave_neighbour<false>((uint32_t)mv[0]) : 0u;
            uint32_t ip0 = L > 1 ? wave_neighbour<false>((uint32_t)pl[0][0]) : 0u;
            uint32_t ip1 = L > 1 ? wave_neighbour<false>((uint32_t)pl[1][0]) : 0u;
            const uint32_t code = bottom.next();  // every lane takes it; the band's bottom lane uses it
            if (last_lane) {
                ipv = 1u;
                imv = 0u;
                ip0 = code;
                ip1 = code >> 1;
            }
            slide_down<K>(pv, ipv);
            slide_down<K>(mv, imv);
            slide_down<K>(pl[0], ip0);
            slide_down<K>(pl[1], ip1);
        } else {
            uint32_t low = (uint32_t)(pv[0] & 1) | ((uint32_t)(mv[0] & 1) << 1);
#pragma unroll
            for (int b = 0; b < BITS; ++b) low |= (uint32_t)(pl[b][0] & 1) << (2 + b);
            uint32_t in = L > 1 ? wave_neighbour<false>(low) : 0u;
            if (last_lane) in = 1u | (bottom.next() << 2);
            slide_down<K>(pv, in);
            slide_down<K>(mv, in >> 1);
#pragma unroll
            for (int b = 0; b < BITS; ++b) slide_down<K>(pl[b], in >> (2 + b));
        }
        // The rows whose letter equals the query's letter j - 1.
        const uint32_t c = query.next();
        if (place == 7) {  // the period's last column: both sequences' next letters
            bottom.refill();
            query.refill();
        }
        uint64_t eq[K];
#pragma unroll
        for (int kk = 0; kk < K; ++kk) {
            uint64_t differ = 0;
#pragma unroll
            for (int b = 0; b < BITS; ++b) {
                const uint32_t m = 0u - ((c >> b) & 1u);  // bit b of the letter, in every bit (both halves alike)
                differ |= pl[b][kk] ^ (((uint64_t)m << 32) | m);
            }
            eq[kk] = ~differ;
        }
        // (Eq & Pv) + Pv over the pair's whole window: carries ripple through a lane's words, and across lanes by
        // generate / propagate ballots (a carry into lane t is bit t of (G + (G | P)) ^ P).
        uint64_t sum[K];
        unsigned long long gen = 0, prop = 0;  // lanes whose words carry out, lanes whose words pass a carry on
        if constexpr (K == 1) {
            // One word: the ballots of the compares themselves, the groups' last lanes masked off in the scalar unit.
            sum[0] = (eq[0] & pv[0]) + pv[0];
            // The compares' lane masks (llvm.amdgcn.icmp, predicates 36 unsigned less, 32 equal).
            gen = __builtin_amdgcn_uicmpl(sum[0], pv[0], 36) & inner_lanes;
            prop = __builtin_amdgcn_uicmpl(sum[0], ~0ull, 32) & inner_lanes;
        } else {
            uint32_t carry = 0;
            bool all_ones = true;
#pragma unroll
            for (int kk = 0; kk < K; ++kk) {
                const uint64_t x = (eq[kk] & pv[kk]) + pv[kk];
                const uint32_t c1 = x < pv[kk];
                sum[kk] = x + carry;
                carry = c1 | (uint32_t)(carry && x == ~0ull);
                all_ones = all_ones && sum[kk] == ~0ull;
            }
            gen = __builtin_amdgcn_ballot_w64(carry != 0) & inner_lanes;
            prop = __builtin_amdgcn_ballot_w64(all_ones) & inner_lanes;
        }
        if constexpr (L > 1) {
            const unsigned long long into = (gen + (gen | prop)) ^ prop;
            if ((into >> lane) & 1ull) {
                uint32_t up = 1;
#pragma unroll
                for (int kk = 0; kk < K; ++kk) {
                    sum[kk] += up;
                    up = up && sum[kk] == 0;
                }
            }
        }
        uint64_t ph[K], mh[K], xv[K];
#pragma unroll
        for (int kk = 0; kk < K; ++kk) {
            xv[kk] = eq[kk] | mv[kk];
            const uint64_t xh = (sum[kk] ^ pv[kk]) | eq[kk];
            ph[kk] = mv[kk] | ~(xh | pv[kk]);
            mh[kk] = pv[kk] & xh;
        }
        // Horizontal differences one row down: a +1 carries into the band's top (the stand-in above it).
        uint32_t carried = (uint32_t)(ph[K - 1] >> 63) | ((uint32_t)(mh[K - 1] >> 63) << 1);
        carried = L > 1 ? wave_neighbour<true>(carried) : 0u;
        if (first_lane) carried = 1u;
#pragma unroll
        for (int kk = K - 1; kk >= 0; --kk) {
            const uint64_t phs = (ph[kk] << 1) | (kk ? ph[kk > 0 ? kk - 1 : 0] >> 63 : (uint64_t)(carried & 1));
            const uint64_t mhs = (mh[kk] << 1) | (kk ? mh[kk > 0 ? kk - 1 : 0] >> 63 : (uint64_t)((carried >> 1) & 1));
            const uint64_t npv = mhs | ~(xv[kk] | phs);
            const uint64_t nmv = phs & xv[kk];
            if constexpr (STAGED) {
                // X: unequal letters whose diagonal neighbour is one less (vertical + horizontal difference 1:
                // one of them +1, neither -1). The planes: X or a reference letter alone; X or a query letter alone.
                const uint64_t unequal_step = ~eq[kk] & (npv ^ phs) & ~(nmv | mhs);
                places[place][lane ^ place] = make_ulonglong2(unequal_step | npv, unequal_step | (ph[kk] & ~npv));
                if (place == 7) {
                    // Every lane's line is full: the wave writes them, lanes 8g .. 8g + 7 the eight places of lane
                    // g' = 8i + g's line in instruction i.
                    const unsigned long long writers = __ballot(active && recording);
                    if (writers) {
                        __builtin_amdgcn_wave_barrier();
                        const uint32_t mine = line + (uint32_t)(((t - 1) >> 3) - (t0 >> 3)) * WL;
                        const int piece = lane & 7, row = lane >> 3;
                        const uint64_t base = (uint64_t)(uintptr_t)a.arena + (uint64_t)piece * 16;
                        if constexpr (L % 8 == 0) {
                            // Lanes 8i .. 8i + 7 are one pair's, so lane 8i + g's line is lane 8i's plus g, and
                            // the eight write or none do. With every pair writing, four lines' places are read
                            // before their four stores, so the reads' waits overlap.
                            if (writers == ~0ull) {
#pragma unroll
                                for (int half = 0; half < 2; ++half) {
                                    ulonglong2 v[4];
                                    uint32_t at[4];
#pragma unroll
                                    for (int i = 0; i < 4; ++i) {
                                        const int g = 4 * half + i;
                                        at[i] = __builtin_amdgcn_readlane(mine, 8 * g) + (uint32_t)row;
                                        v[i] = places[piece][(8 * g + row) ^ piece];
                                    }
#pragma unroll
                                    for (int i = 0; i < 4; ++i)
                                        *reinterpret_cast<global_decisions *>(base + (uint64_t)at[i] * 128) =
                                            planes_of(v[i]);
                                }
                            } else {
#pragma unroll
                                for (int g = 0; g < 8; ++g) {
                                    if (((writers >> (8 * g)) & 0xFFull) == 0) continue;
                                    const uint32_t at = __builtin_amdgcn_readlane(mine, 8 * g) + (uint32_t)row;
                                    *reinterpret_cast<global_decisions *>(base + (uint64_t)at * 128) =
                                        planes_of(places[piece][(8 * g + row) ^ piece]);
                                }
                            }
                        } else {
#pragma unroll
                            for (int g = 0; g < 8; ++g) {
                                if (((writers >> (8 * g)) & 0xFFull) == 0) continue;
                                const int from = 8 * g + row;
                                const uint32_t at = __shfl(mine, from);
                                const ulonglong2 v = places[piece][from ^ piece];
                                if ((writers >> from) & 1ull)
                                    *reinterpret_cast<global_decisions *>(base + (uint64_t)at * 128) = planes_of(v);
                            }
                        }
                        __builtin_amdgcn_wave_barrier();
                    }
                }
            } else if constexpr (RECORD) {
                // X: unequal letters whose diagonal neighbour is one less (vertical + horizontal difference 1:
                // one of them +1, neither -1).
                const uint64_t unequal_step = ~eq[kk] & (npv ^ phs) & ~(nmv | mhs);
#pragma unroll
                for (int q = 0; q + 1 < HOLD; ++q) {
                    held0[q][kk] = held0[q + 1][kk];
                    held1[q][kk] = held1[q + 1][kk];
                }
                held0[HOLD - 1][kk] = unequal_step | npv;
                held1[HOLD - 1][kk] = unequal_step | (ph[kk] & ~npv);
                if (active && recording && (j % HOLD) == 0) {
                    global_decisions *chunk = reinterpret_cast<global_decisions *>((uint64_t)(uintptr_t)trace + held_chunk_at<HOLD>(j / HOLD - 1, WL, k * K + kk));
#pragma unroll
                    for (int q = 0; q < HOLD; ++q) chunk[q] = u64x2{held0[q][kk], held1[q][kk]};
                }
            }
            pv[kk] = npv;
            mv[kk] = nmv;
        }
        top += (int)((uint32_t)pv[0] & 1u) + __builtin_amdgcn_sbfe((int)(uint32_t)mv[0], 0, 1);
        }
    }
}

// ------------------------------------------------------------------ the walk

// The 2-bit choices of one diagonal in the sixteen columns of a span, column q's in bits 2q and 2q + 1, from the
// 32-bit halves (of the word's two decision planes) that hold the diagonal's bit, `shift` within them.
__device__ __forceinline__ uint32_t diagonal_choices(const uint32_t (&x)[16], const uint32_t (&y)[16], uint32_t shift) {
    uint32_t packed = 0;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        packed |= __builtin_amdgcn_ubfe(x[q], shift, 1u) << (2 * q);
        packed |= __builtin_amdgcn_ubfe(y[q], shift, 1u) << (2 * q + 1);
    }
    return packed;
}

// Sixteen bytes of kept decisions in global memory, reached from an address held as an integer (one lane loads
// for another): marked global so the load is a global load, not a flat one.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef const u32x4 __attribute__((address_space(1))) global_chunk;

// What a lane with no span to walk loads, so that every lane makes the same loads.
__device__ ulonglong2 walk_blank[8];  // zeros: global memory, like the arena (one address space, global loads)

// One thread a slot, walking it if it is ready: from (n, m) back to the origin over the 2-bit decisions. Slots are
// in length order, longest first, so a wave's pairs walk about as far as each other. `early`: the slots ready at
// the first pass's end (see launch_edit_walk); otherwise every ready slot not walked early.
//
// The walk goes by rounds, the wave's lanes together. In a round a lane stays in one span (32 columns of one
// word), sixteen columns at a time: it takes the 2-bit choices of its diagonal in those sixteen columns at once,
// makes every diagonal step up to the next gap or the half span's edge in one go, makes the gap, and takes the
// choices again for the new diagonal, until it leaves the span. The moves go to the slot's tape; the CIGAR text is
// made from it afterwards (edit_text).
__global__ __launch_bounds__(64) void edit_walk(EditArgs a, int first, int slots, int early) {
    __shared__ u32x4 span_stage[64][32];     // each lane's span, put together by the lanes that loaded it
    __shared__ ulonglong2 span_where[64][2];  // each lane's span: its four chunks' addresses
    const int t = first + blockIdx.x * 64 + threadIdx.x;
    if (first + (int)blockIdx.x * 64 >= slots) return;  // whole waves only: the rounds are the wave's
    int slot = -1;
    if (t < slots) {
        if (early) {
            slot = a.walk_early[t] ? t : -1;  // read only: the retries' sweeps write `walk_ready` meanwhile
        } else if (a.walk_early[t]) {
            a.walk_ready[t] = 0;  // walked early
        } else if (a.walk_ready[t]) {
            slot = t;
            a.walk_ready[t] = 0;
        }
    }
    // A wave walking a long pair issues first: its rounds, one after another, are the walk's longest chain. Not in
    // the early walk, which has until the retries end, whose sweeps are then the longest chains.
    if (!early) {
        int letters = slot >= 0 ? a.s.tape.rows[slot] + a.s.tape.cols[slot] : 0;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int other = __shfl_xor(letters, d);
            letters = other > letters ? other : letters;
        }
        if (letters > 16000)
            __builtin_amdgcn_s_setprio(3);
        else if (letters > 6000)
            __builtin_amdgcn_s_setprio(2);
        else if (letters > 2000)
            __builtin_amdgcn_s_setprio(1);
    }
    OpTape ops{};
    int r = 0, c = 0, i = 0, width = 0, words = 0;
    int phase = 0;  // the place of the pair's column 0 in its first line
    const uint8_t *base = a.arena;
    uint32_t unequal_marks = 0;  // without eqx two unequal letters are 'M' like two equal ones
    bool outside = false;
    if (slot >= 0) {
        r = a.s.tape.rows[slot];
        c = a.s.tape.cols[slot];
        const int bin = a.s.walk_class[slot];
        width = edit_width(bin);
        words = edit_lanes(bin) * edit_words(bin);
        i = a.s.walk_low[slot] + width - 1 - c + r;  // bit of the band's window: row r's diagonal c - r
        if (r > 0 && c > 0) {
            const int64_t at = a.trace_offset[slot];
            base += at & ~int64_t(127);
            phase = (int)(at & 7);
        }
        ops = op_tape(a.s, slot);
        unequal_marks = a.s.eqx ? 0x55555555u : 0u;
    }
    bool live = slot >= 0 && r > 0 && c > 0;
    if (live && (i < 0 || i >= width)) {  // a proven band holds every step; this pair's did not
        a.s.status[slot] = 3;
        outside = true;
        live = false;
    }
    // A lane's span (its next 32 columns of one word: four 128-byte chunks) is loaded by 32 lanes of its wave, eight
    // to a chunk, so that each load instruction reads eight whole lines rather than a piece of 64; the pieces go to
    // the lane through LDS. A round hands over the pieces loaded a round before (waiting for them), starts the loads
    // of the span the lane expects next (the same word, 32 columns back), and walks while they are in flight, sixteen
    // columns (a half span) at a time from LDS. A lane whose next span is not the one loaded (its diagonal moved into
    // another word) sits out a round while it comes. Span s of word w is tagged s·words + w.
    u32x4 pieces[32];
    uint32_t half_x[16], half_y[16];  // sixteen columns of the span's two decision planes: the 32-bit halves that
    int half = 0, part = 0;           // hold the lane's diagonal (`half`), in the span's `part` (0 low, 1 high)
    int now_tag = -1, loaded_tag = -1;
    int need = 0;  // the span the lane walks next
    // Columns are read at their places: column c - 1 (from 0) is place c - 1 + phase of the pair's line run.
    const int chunks = (c + phase + 7) >> 3;
    const int lane = threadIdx.x;
    const uint64_t blank = (uint64_t)(uintptr_t)walk_blank;
    // Every lane of the wave: publish the span's four chunk addresses, then load piece `lane % 32` of the spans of
    // lanes 2j + lane / 32, j = 0..31.
    const auto issue = [&](int tag) __attribute__((always_inline)) {
        uint64_t at[4] = {blank, blank, blank, blank};
        if (tag >= 0) {
            const int span = tag / words, word = tag - span * words;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int chunk = 4 * span + k < chunks ? 4 * span + k : 4 * span;  // past the pair: any chunk
                at[k] = (uint64_t)(uintptr_t)(base + ((size_t)chunk * words + word) * 128);
            }
        }
        span_where[lane][0] = make_ulonglong2(at[0], at[1]);
        span_where[lane][1] = make_ulonglong2(at[2], at[3]);
        __builtin_amdgcn_wave_barrier();
        const int piece = lane & 31, which = piece >> 3;
#pragma unroll
        for (int j = 0; j < 32; ++j) {
            const ulonglong2 where = span_where[2 * j + (lane >> 5)][which >> 1];
            const uint64_t line = (which & 1) ? where.y : where.x;
            pieces[j] = *reinterpret_cast<global_chunk *>(line + (uint64_t)(piece & 7) * 16);
        }
        loaded_tag = tag;
    };
    // This lane's span in LDS (column q: plane 0's 64 bits, then plane 1's): half `h` of both planes, columns
    // 16p .. 16p + 15.
    const uint32_t *const mine = reinterpret_cast<const uint32_t *>(&span_stage[lane][0]);
    const auto read_half = [&](int h, int p) __attribute__((always_inline)) {
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            half_x[q] = mine[4 * (16 * p + q) + h];
            half_y[q] = mine[4 * (16 * p + q) + 2 + h];
        }
        half = h;
        part = p;
    };
    // Every lane of the wave: hand the loaded pieces to their lanes through LDS (their registers are then free for
    // the next loads).
    const auto hand_over = [&]() __attribute__((always_inline)) {
#pragma unroll
        for (int j = 0; j < 32; ++j) span_stage[2 * j + (lane >> 5)][lane & 31] = pieces[j];
        __builtin_amdgcn_wave_barrier();
        now_tag = loaded_tag;
    };
    if (live) need = ((c - 1 + phase) >> 5) * words + (i >> 6);
    issue(live ? need : -1);
    while (__any(live)) {
        hand_over();
        const bool walking = live && now_tag == need;
        const int expected = walking ? need - words : need;
        issue(live && expected >= 0 ? expected : -1);
        read_half((i & 63) >> 5, ((c - 1 + phase) >> 4) & 1);
        uint32_t choices = diagonal_choices(half_x, half_y, (uint32_t)i & 31u);
        if (walking) {
            const int span = (c - 1 + phase) >> 5, word = i >> 6;
            for (;;) {
                // The diagonal run from column q down to the nearest gap (choice 1 'D' or 2 'I') or the half span's
                // edge, then the gap: one batch of moves. Choice 3 (two unequal letters) is 'X' with eqx, 0 '='.
                const int q = (c - 1 + phase) & 15;
                const uint32_t upto = q == 15 ? 0xFFFFFFFFu : (4u << (2 * q)) - 1u;
                const uint32_t gaps = (choices ^ (choices >> 1)) & 0x55555555u & upto;
                const int stop = gaps ? (31 - __builtin_clz(gaps)) >> 1 : -1;
                // No further than the rows left, nor than the columns left (the places before the first are none).
                const int run = q - stop < r ? q - stop : r;
                const int steps = run < c ? run : c;
                const uint32_t unequal = choices & (choices >> 1) & unequal_marks;
                const uint32_t walked = (__builtin_bitreverse32(unequal) >> 1) >> (2 * (15 - q));
                uint64_t codes = steps >= 16 ? (uint64_t)walked : (uint64_t)(walked & ((1u << (2 * steps)) - 1u));
                r -= steps;
                c -= steps;
                const bool gap = stop >= 0 && r > 0 && c > 0;
                const uint32_t choice = (choices >> (2 * (stop & 15))) & 3u;
                codes |= gap ? (uint64_t)(choice + 1u) << (2 * steps) : 0ull;  // 'D' (code 2) or 'I' (code 3)
                ops.put_many(codes, steps + (gap ? 1 : 0));
                const int deletion = gap && choice == 1u, insertion = gap && choice != 1u;
                r -= deletion;
                c -= insertion;
                i += insertion - deletion;
                const bool finished = r == 0 || c == 0;
                const bool astray = !finished && (i < 0 || i >= width);
                if (finished || astray || (i >> 6) != word || ((c - 1 + phase) >> 5) != span) {
                    if (astray) {  // a proven band holds every step; this pair's did not
                        a.s.status[slot] = 3;
                        outside = true;
                    }
                    live = !(finished || astray);
                    break;
                }
                // Into the span's low half, or the diagonal across bit 31/32: those columns' halves from LDS.
                if (((i & 63) >> 5) != half || (((c - 1 + phase) >> 4) & 1) != part)
                    read_half((i & 63) >> 5, ((c - 1 + phase) >> 4) & 1);
                choices = diagonal_choices(half_x, half_y, (uint32_t)i & 31u);
            }
            if (live) need = ((c - 1 + phase) >> 5) * words + (i >> 6);
        }
    }
    hand_over();  // the last loads: every load the wave made is waited for before it ends
    if (slot >= 0 && !outside) {
        for (; r > 0; --r) ops.put(2u);
        for (; c > 0; --c) ops.put(3u);
        ops.close();
    }
    // The last walk leaves the slots walked early as they are: their text may still be being made.
    if (t < slots && (early || !a.walk_early[t])) a.walk_moves[t] = slot >= 0 && !outside ? (int32_t)ops.count : -1;
}

// The walk of a pair whose longer side is at least EDIT_WALK_LONG letters: one pair a wave, every lane on it. The
// wave keeps 1024 columns (128 lines) of the word its path is in in LDS, and loads the 1024 before them meanwhile.
// Along a diagonal, lane q takes the choice of the column q before the current one: two ballots give the next 64
// columns' choices, the diagonal run is the count of trailing zeros of their gaps (choice 1 'D' or 2 'I', the two
// planes unequal), and a gap turns the diagonal, for which the lanes read again. Every lane holds the same walk state
// and writes the same moves. The moves and the rule are edit_walk's.
constexpr int LONG_LINES = 128;
constexpr int LONG_COLUMNS = 8 * LONG_LINES;

// 16 bits spread to the even bits of 32: code 1 for each set bit, two bits a move.
__device__ __forceinline__ uint32_t spread_bits(uint32_t x) {
    x = (x | (x << 8)) & 0x00FF00FFu;
    x = (x | (x << 4)) & 0x0F0F0F0Fu;
    x = (x | (x << 2)) & 0x33333333u;
    return (x | (x << 1)) & 0x55555555u;
}

__global__ __launch_bounds__(64) void edit_walk_long(EditArgs a, int count, int early) {
    __shared__ u32x4 block[LONG_COLUMNS];  // the resident block: a place's two planes, 16 bytes, from its first line's
    const int t = blockIdx.x, lane = threadIdx.x;
    if (t >= count) return;
    int slot = -1;
    if (early) {
        slot = a.walk_early[t] ? t : -1;  // read only: the retries' sweeps write `walk_ready` meanwhile
    } else if (a.walk_early[t]) {
        if (lane == 0) a.walk_ready[t] = 0;  // walked early
    } else if (a.walk_ready[t]) {
        slot = t;
        if (lane == 0) a.walk_ready[t] = 0;
    }
    if (slot < 0) {
        // The last walk leaves the slots walked early as they are: their text may still be being made.
        if (lane == 0 && (early || !a.walk_early[t])) a.walk_moves[t] = -1;
        return;
    }
    int r = a.s.tape.rows[slot], c = a.s.tape.cols[slot];
    const int bin = a.s.walk_class[slot];
    const int width = edit_width(bin), words = edit_lanes(bin) * edit_words(bin);
    int i = a.s.walk_low[slot] + width - 1 - c + r;  // bit of the band's window: row r's diagonal c - r
    const uint8_t *base = a.arena;
    int phase = 0;  // the place of the pair's column 0 in its first line
    if (r > 0 && c > 0) {
        const int64_t at = a.trace_offset[slot];
        base += at & ~int64_t(127);
        phase = (int)(at & 7);
    }
    OpTape ops = op_tape(a.s, slot);
    const bool eqx = a.s.eqx != 0;  // without eqx two unequal letters are 'M' like two equal ones
    bool live = r > 0 && c > 0, outside = false;
    if (live && (i < 0 || i >= width)) {  // a proven band holds every step; this pair's did not
        if (lane == 0) a.s.status[slot] = 3;
        outside = true;
        live = false;
    }
    const int lines = (c + phase + 7) >> 3;  // the pair's lines of each word
    int held_word = -1, held_low = 0;        // the resident block: its word and first line
    int next_word = -1, next_low = 0;        // the block in flight
    u32x4 pieces[LONG_LINES / 8];
    const auto issue = [&](int word, int low) __attribute__((always_inline)) {
#pragma unroll
        for (int j = 0; j < LONG_LINES / 8; ++j) {
            int line = low + 8 * j + (lane >> 3);
            if (line >= lines) line = low;  // past the pair: any line of it
            pieces[j] = *reinterpret_cast<global_chunk *>(
                (uint64_t)(uintptr_t)(base + ((size_t)line * words + word) * 128) + (uint64_t)(lane & 7) * 16);
        }
        next_word = word;
        next_low = low;
    };
    const auto hand_over = [&]() __attribute__((always_inline)) {
#pragma unroll
        for (int j = 0; j < LONG_LINES / 8; ++j) block[64 * j + lane] = pieces[j];
        __builtin_amdgcn_wave_barrier();
        held_word = next_word;
        held_low = next_low;
        next_word = -1;
    };
    while (live) {
        const int word = i >> 6;
        const int line = (c - 1 + phase) >> 3;  // the current column's
        if (held_word != word || line < held_low || line >= held_low + LONG_LINES) {
            const int low = line - (LONG_LINES - 1) > 0 ? line - (LONG_LINES - 1) : 0;
            if (next_word != word || next_low != low) issue(word, low);
            hand_over();
            if (low > 0) issue(word, low - LONG_LINES > 0 ? low - LONG_LINES : 0);
        }
        // Diagonal runs and gaps in the resident block, until the path leaves it or its word, or ends.
        for (;;) {
            const int place = c - 1 + phase - lane;  // lane q: the column q before the current one
            const bool valid = lane < c && place >= held_low * 8;
            uint32_t xb = 0, yb = 0;
            if (valid) {
                const uint32_t *column = reinterpret_cast<const uint32_t *>(&block[place - held_low * 8]);
                const int h = (i & 63) >> 5;
                xb = (column[h] >> (i & 31)) & 1u;
                yb = (column[2 + h] >> (i & 31)) & 1u;
            }
            const unsigned long long xs = __ballot(xb != 0), ys = __ballot(yb != 0), vs = __ballot(valid);
            const unsigned long long stops = (xs ^ ys) | ~vs;  // a gap, or the block's or the pair's edge
            const int stop = stops ? __builtin_ctzll(stops) : 64;
            int steps = stop < r ? stop : r;
            steps = steps < c ? steps : c;
            const unsigned long long unequal = eqx ? xs & ys : 0ull;
            for (int done = 0; done < steps; done += 16) {
                const int moves = steps - done < 16 ? steps - done : 16;
                const uint32_t bits = (uint32_t)(unequal >> done) & (moves == 16 ? 0xFFFFu : (1u << moves) - 1u);
                ops.put_many((uint64_t)spread_bits(bits), moves);
            }
            r -= steps;
            c -= steps;
            if (r == 0 || c == 0) {
                live = false;
                break;
            }
            if (stop == 64) continue;          // 64 diagonal moves: on along the diagonal
            if (!((vs >> stop) & 1ull)) break;  // the block's edge: the block before it
            const uint32_t choice = (uint32_t)((xs >> stop) & 1ull) | ((uint32_t)((ys >> stop) & 1ull) << 1);
            ops.put(choice + 1u);  // 'D' (code 2) or 'I' (code 3)
            const int deletion = choice == 1u, insertion = choice != 1u;
            r -= deletion;
            c -= insertion;
            i += insertion - deletion;
            if (r == 0 || c == 0) {
                live = false;
                break;
            }
            if (i < 0 || i >= width) {  // a proven band holds every step; this pair's did not
                if (lane == 0) a.s.status[slot] = 3;
                outside = true;
                live = false;
                break;
            }
            if ((i >> 6) != word || ((c - 1 + phase) >> 3) < held_low) break;
        }
    }
    if (next_word >= 0) hand_over();  // the last loads: every load the wave made is waited for before it ends
    if (!outside) {
        for (; r > 0; --r) ops.put(2u);
        for (; c > 0; --c) ops.put(3u);
        ops.close();
    }
    if (lane == 0) a.walk_moves[t] = outside ? -1 : (int32_t)ops.count;
}

// The CIGAR text of the pairs the walk just made moves for: a wave a pair, all 64 lanes on its tape, four pairs a
// workgroup, their room in the output claimed with one atomic.
// `early`: the slots the early walk took; otherwise the others (the two may run at once).
__global__ __launch_bounds__(256) void edit_text(EditArgs a, int slots, int early) {
    __shared__ unsigned long long room[4];
    __shared__ unsigned long long start;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int slot = blockIdx.x * 4 + wave;
    const int moves = slot < slots && (a.walk_early[slot] != 0) == (early != 0) ? a.walk_moves[slot] : -1;
    const uint64_t *words = moves >= 0 ? op_tape(a.s, slot).words : nullptr;
    const int64_t length = moves >= 0 ? wave_cigar_text(a.s.eqx, a.s.forward_cigar, words, moves, nullptr) : 0;
    if (lane == 0) room[wave] = (unsigned long long)length;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long total = room[0] + room[1] + room[2] + room[3];
        start = total ? atomicAdd(a.s.cigar_used, total) : 0ull;
    }
    __syncthreads();
    if (moves < 0) return;
    unsigned long long at = start;
    for (int w = 0; w < wave; ++w) at += room[w];
    wave_cigar_text(a.s.eqx, a.s.forward_cigar, words, moves, a.s.cigars + at);
    if (lane == 0) {
        a.s.cigar_offset[slot] = (int64_t)at;
        a.s.cigar_length[slot] = (int32_t)length;
    }
}

// ------------------------------------------------------------------ lists

// The start of a run: every slot unsettled, the first queues in place, every count and the arena empty.
__global__ void edit_reset(EditArgs a, int slots) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t < slots) {
        a.s.status[t] = -1;
        a.band_low[t] = a.first_band_low[t];
        a.walk_ready[t] = 0;
        a.walk_early[t] = 0;
    }
    for (int bin = 0; bin < EDIT_BINS; ++bin)
        if (t < a.first_counts[bin])
            a.lists[(size_t)bin * a.capacity + t] = a.first_lists[(size_t)bin * a.capacity + t];
    if (t < EDIT_BINS) {
        a.counts[t] = a.first_counts[t];
        a.counts[EDIT_BINS + t] = 0;
        a.file_counts[t] = 0;
        a.file_counts[EDIT_BINS + t] = 0;
        a.spare_counts[t] = 0;
        a.spare_counts[EDIT_BINS + t] = 0;
        a.trace_counts[t] = 0;
        a.trace_counts[EDIT_BINS + t] = 0;
    }
    if (t <= 2 * EDIT_BINS + 1) a.stats[t] = 0;
    if (t < 2 * EDIT_BINS) a.work[t] = 0;
    if (t == 0) {
        *a.deferred_count = 0;
        *a.arena_used = 0;
        if (a.s.cigar_used) *a.s.cigar_used = 0;
    }
}

// The slots ready at the first pass's end, for the early walk.
__global__ void edit_snapshot(EditArgs a, int slots) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t < slots) a.walk_early[t] = a.walk_ready[t];
}

// A further tracing pass: the arena and the trace and walk queues emptied (one workgroup, before the requeue).
__global__ void edit_clear(EditArgs a) {
    const int t = threadIdx.x;
    if (t < 2 * EDIT_BINS) a.trace_counts[t] = 0;
    if (t == 0) {
        *a.deferred_count = 0;
        *a.arena_used = 0;
    }
}

// Queues the pairs `held` (deferred by the pass before, `count` of them) for tracing, in the band `settle` chose,
// deferring again those the arena has no room for.
__global__ void edit_requeue(EditArgs a, const int32_t *held, int count) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= count) return;
    const int slot = held[t];
    const int bin = a.s.walk_class[slot];
    const unsigned long long bytes = edit_trace_bytes(a.s.tape.cols[slot], bin);
    const unsigned long long at = atomicAdd(a.arena_used, bytes);
    if (at + bytes > a.arena_capacity) {
        a.deferred[atomicAdd(a.deferred_count, 1)] = slot;
        return;
    }
    a.trace_offset[slot] = (int64_t)at;
    a.trace_lists[(size_t)bin * a.capacity + atomicAdd(&a.trace_counts[bin], 1)] = slot;
}

// Each slot's first band from the network's prediction of its cost (log(1 + cost)): the guess e^(prediction +
// margin + slope ln(max(longer side, 1000) / 1000)) - 1 (a long pair's failed band costs a long re-sweep), and the
// band out from the start and end diagonals to the last diagonal a path of the guess's cost could visit (unit costs:
// a path through diagonal k > max(0, end) costs at least 2k - end, through k < min(0, end) at least end - 2k), one
// more either side; then the slot is filed in the first queue of the narrowest bin of the first pass's set that
// holds the band. A workgroup's slots (consecutive, so of like length) take one place in each queue with one atomic
// a bin.
__global__ void edit_seed(EditArgs a, const float *prediction, double margin, double slope, EditStart start,
                          int slots) {
    __shared__ int local[EDIT_BINS], base[EDIT_BINS];
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (threadIdx.x < EDIT_BINS) local[threadIdx.x] = 0;
    __syncthreads();
    int bin = -1, rank = 0, low = 0;
    if (t < slots) {
        const int n = a.s.tape.rows[t], m = a.s.tape.cols[t], end = m - n;
        const double longer_side = n > m ? n : m;
        const double guessed =
            expm1((double)prediction[t] + margin + slope * log((longer_side > 1000 ? longer_side : 1000) / 1000.0));
        // Clamped before the cast: a guess above n + m + 2 proves nothing more, and an infinite or huge prediction
        // from a badly trained network must not reach a conversion to an integer it does not fit (NaN fails both
        // tests and gives 0).
        const double most = (double)n + (double)m + 2.0;
        const long long g = guessed > 0 ? (long long)(guessed < most ? guessed : most) : 0;
        const long long from_up = end > 0 ? end : 0, from_down = end < 0 ? end : 0;
        long long up = (g + end) >= 0 ? (g + end) / 2 : -((-(g + end) + 1) / 2);  // floor((g + end) / 2)
        long long down = (end - g) >= 0 ? (end - g + 1) / 2 : -((g - end) / 2);   // ceil((end - g) / 2)
        up = up > from_up ? up : from_up;
        up = up < m + 1 ? up : m + 1;
        down = down < from_down ? down : from_down;
        down = down > -(n + 1) ? down : -(n + 1);
        long long hi = from_up + 1 > up + 1 ? from_up + 1 : up + 1;
        long long lo = from_down - 1 < down - 1 ? from_down - 1 : down - 1;
        hi = hi < m + 1 ? hi : m + 1;
        lo = lo > -n - 1 ? lo : -n - 1;
        const int longer = n > m ? n : m;
        const int from = start.bin[edit_length_class(longer)], widest = edit_widest(n + m);
        bin = edit_first_bin((int)(hi - lo + 1), from, start.bins, from > widest ? from : widest);
        if (bin >= 0) {
            low = edit_place((int)lo, (int)hi, edit_width(bin), n, m);
            rank = atomicAdd(&local[bin], 1);
        }
    }
    __syncthreads();
    if (threadIdx.x < EDIT_BINS)
        base[threadIdx.x] = local[threadIdx.x] ? atomicAdd(&a.first_counts_out[threadIdx.x], local[threadIdx.x]) : 0;
    __syncthreads();
    if (t < slots) {
        a.first_low_out[t] = low;
        if (bin >= 0) a.first_lists[(size_t)bin * a.capacity + base[bin] + rank] = t;
    }
}

__global__ void edit_seed_clear(int32_t *counts) {
    if (threadIdx.x < EDIT_BINS) counts[threadIdx.x] = 0;
}

template <int L, int K, int MODE>
void sweep_bits(const EditArgs &a, int bin, int grid, hipStream_t stream) {
    if (a.s.tape.bits == 2)
        hipLaunchKernelGGL((edit_sweep<L, K, 2, MODE>), dim3(grid), dim3(256), 0, stream, a, bin);
    else
        hipLaunchKernelGGL((edit_sweep<L, K, 8, MODE>), dim3(grid), dim3(256), 0, stream, a, bin);
}

template <int MODE>
void sweep_bin(const EditArgs &a, int bin, int grid, hipStream_t stream) {
    switch (bin) {
        case 0: sweep_bits<1, 1, MODE>(a, bin, grid, stream); break;
        case 1: sweep_bits<2, 1, MODE>(a, bin, grid, stream); break;
        case 2: sweep_bits<3, 1, MODE>(a, bin, grid, stream); break;
        case 3: sweep_bits<4, 1, MODE>(a, bin, grid, stream); break;
        case 4: sweep_bits<6, 1, MODE>(a, bin, grid, stream); break;
        case 5: sweep_bits<8, 1, MODE>(a, bin, grid, stream); break;
        case 6: sweep_bits<12, 1, MODE>(a, bin, grid, stream); break;
        case 7: sweep_bits<16, 1, MODE>(a, bin, grid, stream); break;
        case 8: sweep_bits<21, 1, MODE>(a, bin, grid, stream); break;
        case 9: sweep_bits<32, 1, MODE>(a, bin, grid, stream); break;
        case 10: sweep_bits<64, 1, MODE>(a, bin, grid, stream); break;
        case 11: sweep_bits<64, 2, MODE>(a, bin, grid, stream); break;
        default: sweep_bits<64, 4, MODE>(a, bin, grid, stream); break;
    }
}

}  // namespace

void launch_edit_snapshot(const EditArgs &a, int slots, hipStream_t stream) {
    if (slots > 0) hipLaunchKernelGGL(edit_snapshot, dim3((slots + 255) / 256), dim3(256), 0, stream, a, slots);
}

void launch_edit_reset(const EditArgs &a, int slots, hipStream_t stream) {
    const int threads = slots > 2 * EDIT_BINS + 2 ? slots : 2 * EDIT_BINS + 2;
    hipLaunchKernelGGL(edit_reset, dim3((threads + 255) / 256), dim3(256), 0, stream, a, slots);
}

void launch_edit_sweep(const EditArgs &a, int bin, bool trace, int grid, hipStream_t stream) {
    if (trace)
        sweep_bin<TRACE>(a, bin, grid, stream);
    else if (a.keep_first)
        sweep_bin<SWEEP_TRACED>(a, bin, grid, stream);
    else
        sweep_bin<SWEEP>(a, bin, grid, stream);
}

void launch_edit_seed(const EditArgs &a, const float *prediction, double margin, double slope, const EditStart &start,
                      int slots, hipStream_t stream) {
    hipLaunchKernelGGL(edit_seed_clear, dim3(1), dim3(64), 0, stream, a.first_counts_out);
    if (slots > 0)
        hipLaunchKernelGGL(edit_seed, dim3((slots + 255) / 256), dim3(256), 0, stream, a, prediction, margin, slope,
                           start, slots);
}

void launch_edit_walk_long(const EditArgs &a, int long_slots, hipStream_t stream, bool early) {
    if (long_slots > 0)
        hipLaunchKernelGGL(edit_walk_long, dim3(long_slots), dim3(64), 0, stream, a, long_slots, early ? 1 : 0);
}

void launch_edit_walk_short(const EditArgs &a, int first, int slots, hipStream_t stream, bool early) {
    if (slots > first)
        hipLaunchKernelGGL(edit_walk, dim3((slots - first + 63) / 64), dim3(64), 0, stream, a, first, slots,
                           early ? 1 : 0);
}

void launch_edit_text(const EditArgs &a, int slots, hipStream_t stream, bool early) {
    if (slots <= 0) return;
    hipLaunchKernelGGL(edit_text, dim3((slots + 3) / 4), dim3(256), 0, stream, a, slots, early ? 1 : 0);
}

void launch_edit_requeue(const EditArgs &a, const int32_t *held, int count, hipStream_t stream) {
    hipLaunchKernelGGL(edit_clear, dim3(1), dim3(64), 0, stream, a);
    if (count > 0)
        hipLaunchKernelGGL(edit_requeue, dim3((count + 255) / 256), dim3(256), 0, stream, a, held, count);
}

}  // namespace dinara
