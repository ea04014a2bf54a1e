// Device code the sweeps share: the band certificate, the filing of unproven pairs, and the walk back with its
// CIGAR text (reference lanes.mojo:333-345 off_band, :1337-1392 proof, :1410-1460 walked).
#pragma once

#include "cigar.h"
#include "common.h"

namespace dinara {

__device__ __forceinline__ int imin(int a, int b) { return a < b ? a : b; }

// ------------------------------------------------------------------ the band certificate

// A gap's cost at its cheaper piece, nothing for none (reference lanes.mojo:268-286).
__device__ __forceinline__ int64_t deleted_cost(const CostParams &c, int64_t letters) {
    if (letters <= 0) return 0;
    const int64_t one = (int64_t)c.del_open + letters * (int64_t)c.del_ext;
    if (c.pieces != 2) return one;
    const int64_t two = (int64_t)c.del_open2 + letters * (int64_t)c.del_ext2;
    return two < one ? two : one;
}

__device__ __forceinline__ int64_t inserted_cost(const CostParams &c, int64_t letters) {
    if (letters <= 0) return 0;
    const int64_t one = (int64_t)c.ins_open + letters * (int64_t)c.ins_ext;
    if (c.pieces != 2) return one;
    const int64_t two = (int64_t)c.ins_open2 + letters * (int64_t)c.ins_ext2;
    return two < one ? two : one;
}

// The least any path visiting lane diagonal k (query letters less reference letters) costs, its end on lane
// diagonal `end`: a gap out from the start's diagonal and one back to the end's (reference lanes.mojo:333-345).
__device__ __forceinline__ int64_t off_band(const CostParams &c, int64_t k, int64_t end) {
    if (k > (end > 0 ? end : 0)) return inserted_cost(c, k) + deleted_cost(c, k - end);
    if (k < (end < 0 ? end : 0)) return deleted_cost(c, -k) + inserted_cost(c, end - k);
    return 0;
}

// Whether `found`, the least cost over lane diagonals lo..hi, settles the pair (reference lanes.mojo:1360-1392);
// if not, the band every cheaper (strict: no dearer) path could visit, in *want_low..*want_high (or, under the
// doubling policy, the same band twice as wide).
__device__ __forceinline__ bool certify(const SweepArgs &a, int rows, int cols, int64_t found, int lo, int hi, int *want_low,
                        int *want_high) {
    const int end = cols - rows;
    int top = cols, bottom = -rows;
    if (a.user_band) {
        top = top < a.user_high ? top : a.user_high;
        bottom = bottom > a.user_low ? bottom : a.user_low;
        lo = lo > a.user_low ? lo : a.user_low;
        hi = hi < a.user_high ? hi : a.user_high;
    }
    const int64_t far = INT64_MAX / 4;
    const int64_t beyond_high = hi < top ? off_band(a.costs, hi + 1, end) : far;
    const int64_t beyond_low = lo > bottom ? off_band(a.costs, lo - 1, end) : far;
    const int64_t off = beyond_high < beyond_low ? beyond_high : beyond_low;
    const bool beaten = a.strict ? found < off : found <= off;
    if (beaten || off > a.cap) return true;
    if (a.doubling) {
        // Plain band doubling (the policy the learned band is measured against): the same band twice as wide.
        const int half = (hi - lo + 1) / 2 + 1;
        *want_low = lo - half > bottom ? lo - half : bottom;
        *want_high = hi + half < top ? hi + half : top;
        return false;
    }
    const int64_t capped = found < a.cap ? found : a.cap;
    const int64_t target = a.strict ? capped + 1 : (found <= a.cap ? found : a.cap + 1);
    int h = hi, l = lo;
    while (h < top && off_band(a.costs, h + 1, end) < target) ++h;
    while (l > bottom && off_band(a.costs, l - 1, end) < target) --l;
    *want_low = l;
    *want_high = h;
    return false;
}

// off_band with free letters (reference lanes.mojo:333-345 with LaneEnds, :150-185): a path may start on lane
// diagonals -start_rows .. start_columns and end on end - end_columns .. end + end_rows.
__device__ __forceinline__ int64_t off_band_ends(const CostParams &c, int64_t k, int64_t end, int start_rows,
                                                 int end_rows, int start_columns, int end_columns) {
    const int64_t start_low = -start_rows, start_high = start_columns;
    const int64_t end_low = end - end_columns, end_high = end + end_rows;
    const int64_t high = start_high > end_high ? start_high : end_high;
    const int64_t low = start_low < end_low ? start_low : end_low;
    if (k > high) return inserted_cost(c, k - start_high) + deleted_cost(c, k - end_high);
    if (k < low) return deleted_cost(c, start_low - k) + inserted_cost(c, end_low - k);
    return 0;
}

// certify with free letters (no user band): the same proof and the same next band, off_band_ends pricing.
__device__ __forceinline__ bool certify_ends(const SweepArgs &a, int rows, int cols, int64_t found, int lo, int hi,
                                             const int (&f)[4], int *want_low, int *want_high) {
    const int end = cols - rows;
    const int top = cols, bottom = -rows;
    const int64_t far = INT64_MAX / 4;
    auto off = [&](int64_t k) { return off_band_ends(a.costs, k, end, f[0], f[1], f[2], f[3]); };
    const int64_t beyond_high = hi < top ? off(hi + 1) : far;
    const int64_t beyond_low = lo > bottom ? off(lo - 1) : far;
    const int64_t least = beyond_high < beyond_low ? beyond_high : beyond_low;
    const bool beaten = a.strict ? found < least : found <= least;
    if (beaten || least > a.cap) return true;
    const int64_t capped = found < a.cap ? found : a.cap;
    const int64_t target = a.strict ? capped + 1 : (found <= a.cap ? found : a.cap + 1);
    int h = hi, l = lo;
    while (h < top && off(h + 1) < target) ++h;
    while (l > bottom && off(l - 1) < target) --l;
    *want_low = l;
    *want_high = h;
    return false;
}

// The class of band a pair of `letters` letters needing `need` diagonals takes next: the narrowest that holds them.
// A group or block band is placed on an even diagonal, so it holds one diagonal fewer than its width. Pairs under a
// user band take the wave bands, which bar diagonals outside it; pairs longer than a group launch stages take
// blocks.
__device__ __forceinline__ int class_for(int need, int letters, const SweepArgs &a) {
    if (!a.user_band && a.groups_fit && letters <= a.group_max_letters) {
        for (int cls = CLASS_GROUP32; cls <= CLASS_GROUP256; ++cls)
            if (need < band_width(cls)) return cls;
    }
    if (!a.user_band) {
        for (int cls = CLASS_BLOCK_FIRST; cls < CLASSES; ++cls)
            if (((a.block_mask >> (cls - CLASS_BLOCK_FIRST)) & 1) && need < band_width(cls)) return cls;
    }
    for (int cls = CLASS_WAVE_FIRST; cls <= CLASS_WAVE_LAST; ++cls)
        if (need <= band_width(cls)) return cls;
    return CLASS_TOO_WIDE;
}

__device__ __forceinline__ void file_retry(const SweepArgs &a, int slot, int letters, int want_low, int want_high) {
    a.want_low[slot] = want_low;
    a.want_high[slot] = want_high;
    const int cls = class_for(want_high - want_low + 1, letters, a);
    // A wave sweep stages both sequences in LDS: a pair they overflow has no class to go to.
    if (cls == CLASS_TOO_WIDE || (band_is_wave(cls) && letters > a.lds_limit)) {
        a.status[slot] = 2;
        return;
    }
    a.status[slot] = 1;
    const int at = atomicAdd(&a.retry_counts[cls], 1);
    a.retry_lists[(size_t)cls * a.retry_capacity + at] = slot;
}

// ------------------------------------------------------------------ the walk

__device__ __forceinline__ OpTape op_tape(const SweepArgs &a, int slot) {
    return OpTape{reinterpret_cast<uint64_t *>(a.cigar_scratch + a.cigar_scratch_base[slot]), 0, 0};
}

// The reference's walk (lanes.mojo:1410-1460) over one pair's decisions, `flag_at(r, j)` reading the decision
// of row r's diagonal lo + j: bits 0-1 the source (0 two equal letters, 3 two unequal, 1 the deletion layer, 2 the
// insertion layer), bit 2 the deletion extends into the cell, bit 3 the insertion does. Reaching row 0 or column
// 0, the rest is gaps.
template <typename FlagAt>
__device__ __forceinline__ int64_t walk_rows(const SweepArgs &a, int slot, int n, int m, int lo, FlagAt flag_at) {
    OpTape ops = op_tape(a, slot);
    const uint32_t unequal = a.eqx ? 1u : 0u;
    int r = n, c = m, layer = 0;  // the walk's start: the corner, or a free end's cell
    while (r > 0 && c > 0) {
        const uint32_t f = flag_at(r, c - r - lo);
        if (layer == 0) {
            const uint32_t src = f & 3;
            if (src == 0 || src == 3) {
                ops.put(src == 0 ? 0u : unequal);
                --r;
                --c;
                continue;
            }
            layer = (int)src;
        }
        const bool extended = layer == 1 ? ((f >> 2) & 1) : ((f >> 3) & 1);
        if (layer == 1) {
            ops.put(2u);
            --r;
        } else {
            ops.put(3u);
            --c;
        }
        if (!extended) layer = 0;
    }
    for (; r > 0; --r) ops.put(2u);
    for (; c > 0; --c) ops.put(3u);
    ops.close();
    return ops.count;
}

// The same walk over two-piece decisions, a byte a cell: bits 0-2 the source the aligned layer takes (0 two equal
// letters, 5 two unequal, 1 the first piece's deletions, 2 its insertions, 3 the second piece's deletions, 4 its
// insertions), bit 2 + L gap layer L (1..4) extends into the cell. Odd layers delete, even ones insert.
template <typename FlagAt>
__device__ __forceinline__ int64_t walk_rows2(const SweepArgs &a, int slot, int n, int m, int lo, FlagAt flag_at) {
    OpTape ops = op_tape(a, slot);
    const uint32_t unequal = a.eqx ? 1u : 0u;
    int r = n, c = m, layer = 0;
    while (r > 0 && c > 0) {
        const uint32_t f = flag_at(r, c - r - lo);
        if (layer == 0) {
            const uint32_t src = f & 7;
            if (src == 0 || src == 5) {
                ops.put(src == 0 ? 0u : unequal);
                --r;
                --c;
                continue;
            }
            layer = (int)src;
        }
        const bool extended = (f >> (2 + layer)) & 1;
        if (layer & 1) {
            ops.put(2u);
            --r;
        } else {
            ops.put(3u);
            --c;
        }
        if (!extended) layer = 0;
    }
    for (; r > 0; --r) ops.put(2u);
    for (; c > 0; --c) ops.put(3u);
    ops.close();
    return ops.count;
}

// The same walk over the group sweep's decisions, `nibble_at(r, c)` reading cell (r, c)'s: bits 3:2 (source + 3)
// mod 4 (source 0 two unequal letters, 1 the deletion layer, 2 the insertion layer, 3 two equal letters), bit 1
// D(r+1, c) opened, bit 0 I(r, c+1) opened. Every step makes one move: a gap that reaches (r, c) from below or
// from the right stays in its layer unless the nibble of (r, c) says its next cell opened.
template <typename NibbleAt>
__device__ __forceinline__ int64_t walk_tiles(const SweepArgs &a, int slot, int n, int m, NibbleAt nibble_at) {
    OpTape ops = op_tape(a, slot);
    const uint32_t unequal = a.eqx ? 1u : 0u;
    int r = n, c = m, layer = 0;  // 0 the aligned layer, 1 the deletion layer, 2 the insertion layer
    while (r > 0 && c > 0) {
        const uint32_t f = nibble_at(r, c);
        if ((layer == 1 && (f & 2u)) || (layer == 2 && (f & 1u))) layer = 0;
        uint32_t code;
        if (layer == 0) {
            const uint32_t stored = f >> 2;  // 3 two unequal letters, 0 deletion, 1 insertion, 2 two equal
            code = stored == 2 ? 0u : (stored == 3 ? unequal : stored + 2);
            layer = stored < 2 ? (int)stored + 1 : 0;
        } else {
            code = (uint32_t)layer + 1;
        }
        r -= code != 3u;
        c -= code != 2u;
        ops.put(code);
    }
    for (; r > 0; --r) ops.put(2u);
    for (; c > 0; --c) ops.put(3u);
    ops.close();
    return ops.count;
}

// Writes every walked lane's CIGAR into the shared output (`moves` below zero for a lane with none). Every lane of
// the wave must call it.
__device__ __forceinline__ void publish_cigars(const SweepArgs &a, int slot, int64_t moves) {
    const uint64_t *words = reinterpret_cast<const uint64_t *>(
        a.cigar_scratch + (moves >= 0 ? a.cigar_scratch_base[slot] : 0));
    publish_wave(a.eqx, a.forward_cigar, words, moves, a.cigars, a.cigar_used, a.cigar_offset + slot,
                 a.cigar_length + slot);
}

}  // namespace dinara
