// Device code the sweeps share: the band certificate, the filing of unproven pairs, and the walk back with its
// CIGAR text (reference lanes.mojo:333-345 off_band, :1337-1392 proof, :1410-1460 walked).
#pragma once

#include "common.h"

namespace dinara {

__device__ __forceinline__ int imin(int a, int b) { return a < b ? a : b; }

// ------------------------------------------------------------------ the band certificate

__device__ __forceinline__ int64_t deleted_cost(const CostParams &c, int64_t letters) {
    return letters <= 0 ? 0 : (int64_t)c.del_open + letters * (int64_t)c.del_ext;
}

__device__ __forceinline__ int64_t inserted_cost(const CostParams &c, int64_t letters) {
    return letters <= 0 ? 0 : (int64_t)c.ins_open + letters * (int64_t)c.ins_ext;
}

// The least any path visiting lane diagonal k (query letters less reference letters) costs, its end on lane
// diagonal `end`: a gap out from the start's diagonal and one back to the end's (reference lanes.mojo:333-345).
__device__ __forceinline__ int64_t off_band(const CostParams &c, int64_t k, int64_t end) {
    if (k > (end > 0 ? end : 0)) return inserted_cost(c, k) + deleted_cost(c, k - end);
    if (k < (end < 0 ? end : 0)) return deleted_cost(c, -k) + inserted_cost(c, end - k);
    return 0;
}

// Whether `found`, the least cost over lane diagonals lo..hi, settles the pair (reference lanes.mojo:1360-1392);
// if not, the band every cheaper (strict: no dearer) path could visit, in *want_low..*want_high.
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
    const int64_t capped = found < a.cap ? found : a.cap;
    const int64_t target = a.strict ? capped + 1 : (found <= a.cap ? found : a.cap + 1);
    int h = hi, l = lo;
    while (h < top && off_band(a.costs, h + 1, end) < target) ++h;
    while (l > bottom && off_band(a.costs, l - 1, end) < target) --l;
    *want_low = l;
    *want_high = h;
    return false;
}

// The class of band a pair needing `need` diagonals takes next: the narrowest that holds them. A group band
// is placed on an even diagonal, so it holds one diagonal fewer than its width.
__device__ __forceinline__ int class_for(int need, const SweepArgs &a) {
    if (!a.user_band && a.groups_fit) {
        for (int cls = CLASS_GROUP32; cls <= CLASS_GROUP256; ++cls)
            if (need < (32 << cls)) return cls;
    }
    for (int cls = CLASS_WAVE_FIRST; cls < CLASSES; ++cls)
        if (need <= (64 << (cls - CLASS_WAVE_FIRST))) return cls;
    return CLASS_TOO_WIDE;
}

__device__ __forceinline__ void file_retry(const SweepArgs &a, int slot, int letters, int want_low, int want_high) {
    a.want_low[slot] = want_low;
    a.want_high[slot] = want_high;
    const int cls = class_for(want_high - want_low + 1, a);
    // A wave sweep stages both sequences in LDS: a pair they overflow has no class to go to.
    if (cls == CLASS_TOO_WIDE || (cls >= CLASS_WAVE_FIRST && letters > a.lds_limit)) {
        a.status[slot] = 2;
        return;
    }
    a.status[slot] = 1;
    const int at = atomicAdd(&a.retry_counts[cls], 1);
    a.retry_lists[(size_t)cls * a.retry_capacity + at] = slot;
}

// ------------------------------------------------------------------ the walk and the CIGAR text

// A walk's moves, 2 bits each, 32 to a 64-bit word, in the order the walk makes them (from the pair's end):
// 0 two equal letters ('=', or 'M' for both kinds without eqx), 1 two unequal ('X'), 2 'D', 3 'I'.
struct OpTape {
    uint64_t *words;
    uint64_t held;
    int64_t count;

    __device__ void put(uint32_t code) {
        held |= (uint64_t)code << (2 * (count & 31));
        if ((++count & 31) == 0) {
            words[(count >> 5) - 1] = held;
            held = 0;
        }
    }
    __device__ void close() {
        if (count & 31) words[count >> 5] = held;
    }
};

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
    int r = n, c = m, layer = 0;
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

__device__ __forceinline__ int decimal_digits(uint32_t v) {
    return v < 10u ? 1 : v < 100u ? 2 : v < 1000u ? 3 : v < 10000u ? 4 : v < 100000u ? 5 : v < 1000000u ? 6
         : v < 10000000u ? 7 : v < 100000000u ? 8 : v < 1000000000u ? 9 : 10;
}

// The CIGAR text of a walk's moves: runs of one move, each its length then its letter, in the alignment's order
// (the moves' own order when the sequences were walked reversed, else back to front). Returns the text's length;
// writes it to `out` unless null. Runs are found a word at a time: the moves that differ from the run's are the
// nonzero 2-bit fields of the word xor the run's code repeated.
__device__ __forceinline__ int64_t cigar_text(const SweepArgs &a, const uint64_t *words, int64_t count, char *out) {
    const char letter_of[4] = {a.eqx ? '=' : 'M', 'X', 'D', 'I'};
    int64_t length = 0;
    auto emit = [&](uint32_t code, int64_t run) {
        const uint32_t v = (uint32_t)run;
        const int digits = decimal_digits(v);
        if (out) {
            uint32_t rest = v;
            for (int i = digits - 1; i >= 0; --i) {
                out[length + i] = (char)('0' + rest % 10u);
                rest /= 10u;
            }
            out[length + digits] = letter_of[code];
        }
        length += digits + 1;
    };
    auto code_at = [&](int64_t p) { return (uint32_t)(words[p >> 5] >> (2 * (p & 31))) & 3u; };
    if (a.forward_cigar) {
        int64_t p = 0;
        while (p < count) {
            const uint32_t code = code_at(p);
            const uint64_t repeated = (uint64_t)code * 0x5555555555555555ull;
            int64_t q = p;
            while (q < count) {
                const int off = (int)(q & 31);
                const uint64_t x = (words[q >> 5] ^ repeated) >> (2 * off);
                if (x == 0) {
                    q += 32 - off;
                    continue;
                }
                q += __builtin_ctzll(x) >> 1;
                break;
            }
            q = q < count ? q : count;
            emit(code, q - p);
            p = q;
        }
    } else {
        int64_t p = count - 1;
        while (p >= 0) {
            const uint32_t code = code_at(p);
            const uint64_t repeated = (uint64_t)code * 0x5555555555555555ull;
            int64_t q = p;
            while (q >= 0) {
                const int off = (int)(q & 31);
                const uint64_t x = (words[q >> 5] ^ repeated) << (2 * (31 - off));
                if (x == 0) {
                    q -= off + 1;
                    continue;
                }
                q -= __builtin_clzll(x) >> 1;
                break;
            }
            emit(code, p - q);
            p = q;
        }
    }
    return length;
}

// Claims room for every lane's CIGAR in the shared output with one atomic a wave: an exclusive prefix sum of the
// lengths across the wave, its total added to the counter by the last lane. Every lane of the wave must call it
// (a lane with nothing to publish brings length 0).
__device__ __forceinline__ unsigned long long wave_claim(unsigned long long *counter, unsigned long long length) {
    const int lane = threadIdx.x & 63;
    unsigned long long inclusive = length;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const unsigned long long other = __shfl_up(inclusive, d);
        if (lane >= d) inclusive += other;
    }
    unsigned long long base = 0;
    if (lane == 63 && inclusive > 0) base = atomicAdd(counter, inclusive);
    base = __shfl(base, 63);
    return base + inclusive - length;
}

// Writes every walked lane's CIGAR into the shared output (`moves` below zero for a lane with none). Every lane of
// the wave must call it.
__device__ __forceinline__ void publish_cigars(const SweepArgs &a, int slot, int64_t moves) {
    const uint64_t *words = reinterpret_cast<const uint64_t *>(
        a.cigar_scratch + (moves >= 0 ? a.cigar_scratch_base[slot] : 0));
    const int64_t length = moves >= 0 ? cigar_text(a, words, moves, nullptr) : 0;
    const unsigned long long at = wave_claim(a.cigar_used, (unsigned long long)length);
    if (moves < 0) return;
    cigar_text(a, words, moves, a.cigars + at);
    a.cigar_offset[slot] = (int64_t)at;
    a.cigar_length[slot] = (int32_t)length;
}

}  // namespace dinara
