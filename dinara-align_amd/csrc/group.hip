// Banded global alignment, sixteen lanes a pair, by anti-diagonal (MI355X, gfx950).
//
// The band's W = 32S diagonals are dealt out two-and-two: lane j (0..15) of a pair's group holds the S diagonal
// pairs k0 + 2(Sj + s) and k0 + 2(Sj + s) + 1, s < S. Cell (r, c) lies on diagonal k = c - r and anti-diagonal
// t = r + c, so on an even t every even diagonal has a cell and on an odd t every odd one: each half-step
// computes S cells a lane, every lane busy. Sixteen lanes a pair give the GPU sixteen times the waves of one
// thread a pair, which a batch of 100k short pairs needs to fill 1024 SIMDs. Letters sit in LDS, a byte each.
//
// The recurrence is Gotoh's in cost form (reference lanes.mojo:795-947), arranged so each cell hands its
// successors what they need: cell (r, c) computes its own H from H(r-1, c-1) (the slot's own value two steps
// back), D(r, c) and I(r, c), and then D(r+1, c) = min(H + del_open, D) + del_ext and I(r, c+1) = min(H +
// ins_open, I) + ins_ext. D(r+1, c) lies one diagonal down, I(r, c+1) one up: inside a lane those are its other
// slots, across lanes one DPP row shift a half-step brings the neighbour's (row_shl:1 the D from the lane above,
// row_shr:1 the I from the lane below). A group is one DPP row: an edge lane's shift has no source lane and keeps
// the register's old value, INF from the start, so the band's edges read INF without an extra instruction.
//
// Tags. Every value is cost * 16 + a 4-bit tag, so one min both finds the cost and breaks ties the reference's
// way (lanes.mojo:905-923): H's candidates carry the source in bits 3:2 (0 two unequal letters, 1 the deletion
// layer, 2 the insertion layer, 3 two equal letters: the reference's flag priority, lowest wins), D(r+1, c)'s
// candidates carry bit 1 (0 extended from D(r, c), 1 opened from H: extension wins a tie, as the reference's
// `extended <= opened`), I's bit 0 the same. A cell's decision nibble is bits 3:2 (source + 3) mod 4, bit 1 D(r+1,
// c) opened, bit 0 I(r, c+1) opened: one v_xad and one v_alignbit a cell. The walk (walk_bands) reads a cell's
// source from its own nibble, and whether a gap reaching it extended from the nibble of the cell it moves to.
//
// With both extensions equal (SH), values are held shifted, H' = H - e(r + c): an extension then costs nothing,
// a diagonal step its pair's cost less 2e, and the borders are flat; every comparison is between values of one
// cell, shifted alike, so the decisions are the same.
//
// Decisions are stored 4 bits a cell: for each diagonal pair q = (k - k0) / 2 one 32-bit word every 8
// anti-diagonals (a block), nibble t % 8 (even t the even diagonal's cell, odd t the odd one's); the four blocks
// of a tile of 32 anti-diagonals side by side, 16 bytes, and the tile's diagonal pairs side by side: word
// ((t / 32) * (W / 2) + q) * 4 + (t / 8) % 4. A lane writes a tile of a slot with one 16-byte store, and a walk
// along a diagonal reads 32 anti-diagonals with one 16-byte load.

#include "bands.h"

namespace dinara {

// Tagged unreachable: above every real tagged value (the host keeps costs times 16 under 2^30) with room to add.
constexpr int TAG_INF = 1 << 30;

// The neighbour one lane below: within a 16-lane row (row_shr:1), or across the wave (wave_shr:1). The first lane
// has no source and keeps the old value: its register never changes.
template <bool WAVE>
__device__ __forceinline__ int shift_from_lower(int old, int v) {
    return __builtin_amdgcn_update_dpp(old, v, WAVE ? 0x138 : 0x111, 0xF, 0xF, false);
}
// The neighbour one lane above: within the row (row_shl:1), or across the wave (wave_shl:1); the last lane keeps
// the old value.
template <bool WAVE>
__device__ __forceinline__ int shift_from_upper(int old, int v) {
    return __builtin_amdgcn_update_dpp(old, v, WAVE ? 0x130 : 0x101, 0xF, 0xF, false);
}

// The tagged constants a sweep adds.
struct GroupCosts {
    int sub_eq, sub_ne;  // H(r-1, c-1) to H(r, c): two equal letters (tag 3), two unequal (tag 0)
    int open_d, open_i;  // H to the next gap cell: the opening, tagged D opened / I opened
    int ext_d, ext_i;    // what a gap letter adds after the min (zero when shifted)
    int row0, col0;      // the borders' step: H(0, c) = ins_open + c * row0, H(r, 0) = del_open + r * col0
    int ins_open, del_open;
};

template <bool RECORD, bool SH>
__device__ __forceinline__ GroupCosts group_costs(const CostParams &k) {
    GroupCosts g;
    const int e = SH ? k.del_ext : 0;
    g.sub_eq = (-2 * e) * 16 + (RECORD ? 12 : 0);
    g.sub_ne = (k.mismatch - 2 * e) * 16;
    g.open_d = k.del_open * 16 + (RECORD ? 6 : 0);
    g.open_i = k.ins_open * 16 + (RECORD ? 9 : 0);
    g.ext_d = SH ? 0 : k.del_ext * 16;
    g.ext_i = SH ? 0 : k.ins_ext * 16;
    g.row0 = SH ? 0 : k.ins_ext * 16;
    g.col0 = SH ? 0 : k.del_ext * 16;
    g.ins_open = k.ins_open * 16;
    g.del_open = k.del_open * 16;
    return g;
}

// The cell after its H: D(r+1, c) and I(r, c+1) from the clean H and the gap candidates, and the decision nibble.
// (dn ^ in) has bits 3:0 = 1, 1, D opened, I opened; h's bits 1:0 are clear, so the low nibble of the sum is
// (source + 3) mod 4, D opened, I opened. alignbit shifts the nibble in from the word's top.
template <bool RECORD, bool SH>
__device__ __forceinline__ void group_tail(const GroupCosts &g, int h, int hc, int dc, int ic, int &dn, int &in,
                                           uint32_t &acc) {
    dn = min(hc + g.open_d, dc);
    in = min(hc + g.open_i, ic);
    if (!SH) {
        dn += g.ext_d;
        in += g.ext_i;
    }
    if (RECORD) acc = __builtin_amdgcn_alignbit((uint32_t)((dn ^ in) + h), acc, 4);
}

// One cell inside the matrix: `hd` the clean H(r-1, c-1), `dx` = D(r, c) and `ix` = I(r, c) as their producers
// tagged them. Returns the clean H in `hc`, D(r+1, c) in `dn` and I(r, c+1) in `in`; appends the decision.
template <bool RECORD, bool SH>
__device__ __forceinline__ void group_cell(const GroupCosts &g, bool eq, int hd, int dx, int ix, int &hc, int &dn,
                                           int &in, uint32_t &acc) {
    const int sub = hd + (eq ? g.sub_eq : g.sub_ne);
    // The gaps as H's candidates: tag D (4) or I (8), their open bits cleared, which also makes them the
    // extension candidates of the next gap cells.
    const int dc = RECORD ? (dx & ~2) : dx;
    const int ic = RECORD ? (ix & ~1) : ix;
    const int h = min(sub, min(dc, ic));
    hc = RECORD ? (h & ~15) : h;
    group_tail<RECORD, SH>(g, h, hc, dc, ic, dn, in, acc);
}

// The careful form: row 0 is an insertion of every column before, column 0 a deletion of every row, outside the
// matrix unreachable; the letters are read clamped inside both sequences, STRIDE bytes apart. Also returns the
// tagged H in `h`.
template <bool RECORD, bool SH, int STRIDE = 1>
__device__ __forceinline__ void careful_cell(const GroupCosts &g, int r, int c, const uint8_t *ref, const uint8_t *qry,
                                             int last_ref, int last_qry, int hd, int dx, int ix, int &h, int &hc,
                                             int &dn, int &in, uint32_t &acc) {
    const int ri = min(max(r - 1, 0), last_ref), ci = min(max(c - 1, 0), last_qry);
    const bool eq = ref[STRIDE * ri] == qry[STRIDE * ci];
    const int sub = hd + (eq ? g.sub_eq : g.sub_ne);
    const int dc = RECORD ? (dx & ~2) : dx;
    const int ic = RECORD ? (ix & ~1) : ix;
    h = min(sub, min(dc, ic));
    if (r <= 0 || c <= 0) {
        if (r < 0 || c < 0) h = TAG_INF;
        else if (r == 0) h = c == 0 ? 0 : g.ins_open + c * g.row0;
        else h = g.del_open + r * g.col0;
    }
    hc = RECORD ? (h & ~15) : h;
    group_tail<RECORD, SH>(g, h, hc, dc, ic, dn, in, acc);
    if (r < 0 || c < 0) dn = in = TAG_INF;
}

// A lane's state: per slot the clean H of its last cell (He/Ho), the D it handed the cell below (De/Do) and the
// I it handed the cell to the right (Ie/Io), for the even and the odd diagonal of each pair; the DPP receivers
// of the edge slots; the decisions of the current block of 8 anti-diagonals.
template <int S>
struct GroupState {
    int He[S], De[S], Ie[S], Ho[S], Do[S], Io[S];
    int from_lower_i, from_upper_d;
    uint32_t acc[S];      // the current block's decisions
    uint32_t done[S][3];  // the tile's earlier blocks
};

// Even anti-diagonal: even slot s's cell takes D from odd slot s (one diagonal up) and I from odd slot s - 1
// (the lane below's last for slot 0, in from_lower_i, which the caller has shifted in). The careful form captures
// the corner's tagged H when it is this cell.
template <int S, bool RECORD, bool SH, bool CAREFUL, int STRIDE = 1>
__device__ __forceinline__ void even_half(GroupState<S> &st, const GroupCosts &g, int t, int base_k,
                                          const uint8_t *ref, const uint8_t *qry, int last_ref, int last_qry,
                                          const uint8_t *rp, const uint8_t *qp, int first_q, int q_end, int tend,
                                          int &corner) {
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int ix = s > 0 ? st.Io[s - 1] : st.from_lower_i;
        int hc, dn, in;
        if (CAREFUL) {
            const int kk = base_k + 2 * s, r = (t - kk) >> 1;
            int h;
            careful_cell<RECORD, SH, STRIDE>(g, r, r + kk, ref, qry, last_ref, last_qry, st.He[s], st.Do[s], ix, h,
                                             hc, dn, in, st.acc[s]);
            if (t == tend && first_q + s == q_end) corner = h;
        } else {
            group_cell<RECORD, SH>(g, rp[-s] == qp[s], st.He[s], st.Do[s], ix, hc, dn, in, st.acc[s]);
        }
        st.He[s] = hc;
        st.De[s] = dn;
        st.Ie[s] = in;
    }
}

// Odd anti-diagonal: odd slot s's cell takes D from even slot s + 1 (the lane above's first for the last slot, in
// from_upper_d, which the caller has shifted in) and I from even slot s.
template <int S, bool RECORD, bool SH, bool CAREFUL, int STRIDE = 1>
__device__ __forceinline__ void odd_half(GroupState<S> &st, const GroupCosts &g, int t, int base_k,
                                         const uint8_t *ref, const uint8_t *qry, int last_ref, int last_qry,
                                         const uint8_t *rp, const uint8_t *qp, int first_q, int q_end, int tend,
                                         int &corner) {
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int dx = s + 1 < S ? st.De[s + 1] : st.from_upper_d;
        int hc, dn, in;
        if (CAREFUL) {
            const int kk = base_k + 2 * s + 1, r = (t - kk) >> 1;
            int h;
            careful_cell<RECORD, SH, STRIDE>(g, r, r + kk, ref, qry, last_ref, last_qry, st.Ho[s], dx, st.Ie[s], h,
                                             hc, dn, in, st.acc[s]);
            if (t == tend && first_q + s == q_end) corner = h;
        } else {
            group_cell<RECORD, SH>(g, rp[-s] == qp[s + 1], st.Ho[s], dx, st.Ie[s], hc, dn, in, st.acc[s]);
        }
        st.Ho[s] = hc;
        st.Do[s] = dn;
        st.Io[s] = in;
    }
}

// Packed cells: two slots of a lane in the 16-bit halves of one register, slot q low and slot q + S/2 high, for
// the steady anti-diagonals of a shifted sweep (both extensions equal). There a tagged value is 16 (H - e(r + c))
// plus its tag, between -16 e (r + c) and 32 max(opening) + 15, and a cell no path inside the band reaches holds
// `top` less at most 16 e (r + c): the host takes this form only where both fit 16 bits apart (group16_top), so
// every min picks what the 32-bit form picks and the low nibble of every sum, the decision, is the same. The adds
// wrap as 16-bit integers and never reach the wrap; the mins are signed.
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
typedef short i16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t pk_add(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2, a) + __builtin_bit_cast(u16x2, b));
}
__device__ __forceinline__ uint32_t pk_min(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t,
                              __builtin_elementwise_min(__builtin_bit_cast(i16x2, a), __builtin_bit_cast(i16x2, b)));
}
__device__ __forceinline__ uint32_t pk_mad(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2, a) * __builtin_bit_cast(u16x2, b) +
                                            __builtin_bit_cast(u16x2, c));
}
__device__ __forceinline__ uint32_t both_halves(int v) { return ((uint32_t)v & 0xFFFFu) * 0x10001u; }
__device__ __forceinline__ uint32_t pack_pair(int lo, int hi, int top) {
    return ((uint32_t)min(lo, top) & 0xFFFFu) | ((uint32_t)min(hi, top) << 16);
}

struct PackedCosts {
    uint32_t sub_eq, sub_gap, open_d, open_i;  // both halves: two equal letters, unequal less equal, the openings
};

template <int S>
struct PackedState {
    uint32_t He[S / 2], De[S / 2], Ie[S / 2], Ho[S / 2], Do[S / 2], Io[S / 2];
    uint32_t lower, upper;  // the DPP receivers; an edge lane's keeps `top`
};

// One register of cells: `apart` the two slots' letter differences in bytes 0 and 1 (zero: equal letters).
template <bool RECORD>
__device__ __forceinline__ void packed_cell(const PackedCosts &k, uint32_t apart, uint32_t hd, uint32_t dx,
                                            uint32_t ix, uint32_t &hc, uint32_t &dn, uint32_t &in, uint32_t &acc_lo,
                                            uint32_t &acc_hi) {
    const uint32_t unequal = pk_min(__builtin_amdgcn_perm(0u, apart, 0x0C010C00u), 0x00010001u);  // bytes to halves
    const uint32_t sub = pk_add(hd, pk_mad(unequal, k.sub_gap, k.sub_eq));
    const uint32_t dc = RECORD ? (dx & 0xFFFDFFFDu) : dx;
    const uint32_t ic = RECORD ? (ix & 0xFFFEFFFEu) : ix;
    const uint32_t h = pk_min(sub, pk_min(dc, ic));
    hc = RECORD ? (h & 0xFFF0FFF0u) : h;
    dn = pk_min(pk_add(hc, k.open_d), dc);
    in = pk_min(pk_add(hc, k.open_i), ic);
    if (RECORD) {
        const uint32_t y = pk_add(dn ^ in, h);  // each half's low nibble: the 32-bit form's decision
        acc_lo = __builtin_amdgcn_alignbit(y, acc_lo, 4);
        acc_hi = __builtin_amdgcn_alignbit(y >> 16, acc_hi, 4);
    }
}

// The packed half-steps. Letters: `rp` the reference's pairs at r0 - 1 (byte 0 letter i, byte 1 letter i - S/2),
// `qp` the query's at c0 - 1 (byte 0 letter i, byte 1 letter i + S/2). Even slot pair q takes I from odd pair q - 1
// (pair 0: the lane below's last slot and its own slot S/2 - 1); odd pair q takes D from even pair q + 1 (the last:
// its own slot S/2 and the lane above's slot 0).
template <int S, bool RECORD>
__device__ __forceinline__ void packed_even(PackedState<S> &ps, GroupState<S> &st, const PackedCosts &k,
                                            const uint16_t *rp, const uint16_t *qp) {
    constexpr int P = S / 2;
    ps.lower = (uint32_t)shift_from_lower<false>((int)ps.lower, (int)ps.Io[P - 1]);
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const uint32_t ix = q > 0 ? ps.Io[q > 0 ? q - 1 : 0] : __builtin_amdgcn_alignbit(ps.Io[P - 1], ps.lower, 16);
        uint32_t hc, dn, in;
        packed_cell<RECORD>(k, (uint32_t)rp[-q] ^ (uint32_t)qp[q], ps.He[q], ps.Do[q], ix, hc, dn, in, st.acc[q],
                            st.acc[q + P]);
        ps.He[q] = hc;
        ps.De[q] = dn;
        ps.Ie[q] = in;
    }
}

template <int S, bool RECORD>
__device__ __forceinline__ void packed_odd(PackedState<S> &ps, GroupState<S> &st, const PackedCosts &k,
                                           const uint16_t *rp, const uint16_t *qp) {
    constexpr int P = S / 2;
    ps.upper = (uint32_t)shift_from_upper<false>((int)ps.upper, (int)ps.De[0]);
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const uint32_t dx = q + 1 < P ? ps.De[q + 1 < P ? q + 1 : 0] : __builtin_amdgcn_alignbit(ps.upper, ps.De[0], 16);
        uint32_t hc, dn, in;
        packed_cell<RECORD>(k, (uint32_t)rp[-q] ^ (uint32_t)qp[q + 1], ps.Ho[q], dx, ps.Ie[q], hc, dn, in, st.acc[q],
                            st.acc[q + P]);
        ps.Ho[q] = hc;
        ps.Do[q] = dn;
        ps.Io[q] = in;
    }
}

template <int S>
__device__ __forceinline__ void pack_state(PackedState<S> &ps, const GroupState<S> &st, int top) {
    constexpr int P = S / 2;
#pragma unroll
    for (int q = 0; q < P; ++q) {
        ps.He[q] = pack_pair(st.He[q], st.He[q + P], top);
        ps.De[q] = pack_pair(st.De[q], st.De[q + P], top);
        ps.Ie[q] = pack_pair(st.Ie[q], st.Ie[q + P], top);
        ps.Ho[q] = pack_pair(st.Ho[q], st.Ho[q + P], top);
        ps.Do[q] = pack_pair(st.Do[q], st.Do[q + P], top);
        ps.Io[q] = pack_pair(st.Io[q], st.Io[q + P], top);
    }
    ps.lower = ps.upper = pack_pair(top, top, top);
}

template <int S>
__device__ __forceinline__ void unpack_state(GroupState<S> &st, const PackedState<S> &ps) {
    constexpr int P = S / 2;
#pragma unroll
    for (int q = 0; q < P; ++q) {
        st.He[q] = __builtin_amdgcn_sbfe((int)ps.He[q], 0, 16), st.He[q + P] = (int)ps.He[q] >> 16;
        st.De[q] = __builtin_amdgcn_sbfe((int)ps.De[q], 0, 16), st.De[q + P] = (int)ps.De[q] >> 16;
        st.Ie[q] = __builtin_amdgcn_sbfe((int)ps.Ie[q], 0, 16), st.Ie[q + P] = (int)ps.Ie[q] >> 16;
        st.Ho[q] = __builtin_amdgcn_sbfe((int)ps.Ho[q], 0, 16), st.Ho[q + P] = (int)ps.Ho[q] >> 16;
        st.Do[q] = __builtin_amdgcn_sbfe((int)ps.Do[q], 0, 16), st.Do[q + P] = (int)ps.Do[q] >> 16;
        st.Io[q] = __builtin_amdgcn_sbfe((int)ps.Io[q], 0, 16), st.Io[q + P] = (int)ps.Io[q] >> 16;
    }
}

// PACK: the steady anti-diagonals in packed cells (S >= 2, shifted). A pair's letters are then staged as 16-bit
// positions, the reference's position i holding letters i and i - S/2, the query's i and i + S/2 (the careful
// cells read the low bytes).
template <int S, bool RECORD, bool SH, bool PACK>
__global__ __launch_bounds__(64) void sweep_group(SweepArgs a) {
    static_assert(!PACK || (SH && S >= 2 && S % 2 == 0), "packed cells: shifted sweeps of two slots or more");
    extern __shared__ uint8_t letters[];
    constexpr int W = 32 * S;
    constexpr int STRIDE = PACK ? 2 : 1;  // bytes between a sequence's letters in LDS
    constexpr int P = S / 2;
    const int lane = threadIdx.x & 15;
    const int local = threadIdx.x >> 4;  // the pair within the block (four a block)
    const int index = blockIdx.x * 4 + local;
    const bool live = index < a.count;
    // Lanes past the list sweep a copy of the first pair, store nothing, and keep the DPP rows whole.
    const int slot = a.list[live ? index : 0];
    const int n = a.tape.rows[slot], m = a.tape.cols[slot];
    const int need_low = a.want_low[slot], need_high = a.want_high[slot];
    int k0 = need_low - (W - (need_high - need_low + 1)) / 2;
    k0 -= k0 & 1;  // even: even diagonals on even anti-diagonals
    const int end = m - n;
    const GroupCosts g = group_costs<RECORD, SH>(a.costs);

    // A pair's region: its reference, its query, and room past the query for the steady state's reads, which run
    // up to half a band past either sequence's end by the pair's last anti-diagonal. Packed: the reference's S/2
    // positions past its end hold its last letters' second bytes.
    uint8_t *ref = letters + (size_t)local * a.group_letters * STRIDE;
    uint8_t *qry = ref + STRIDE * (n + (PACK ? P : 0));
    {
        const int group = slot >> 6, column = slot & 63;
        const uint32_t *ref_words = a.tape.words + a.tape.group_base[group] + column;
        const uint32_t *qry_words = ref_words + (size_t)a.tape.group_ref_words[group] * 64;
        const int bits = a.tape.bits, per = 32 / bits;
        const uint32_t mask = (1u << bits) - 1;
        for (int i = lane; i < n; i += 16) {
            const uint8_t letter = (uint8_t)((ref_words[(size_t)(i / per) * 64] >> (bits * (i % per))) & mask);
            ref[STRIDE * i] = letter;
            if (PACK) ref[STRIDE * (i + P) + 1] = letter;
        }
        for (int i = lane; i < m; i += 16) {
            const uint8_t letter = (uint8_t)((qry_words[(size_t)(i / per) * 64] >> (bits * (i % per))) & mask);
            qry[STRIDE * i] = letter;
            if (PACK && i >= P) qry[STRIDE * (i - P) + 1] = letter;
        }
    }
    __syncthreads();

    // Steps for the whole wave: to its longest pair's corner, the borders handled until every pair's band has
    // left row 0 and column 0.
    const int tend = n + m;
    int tmax = tend, tmin = tend;
    int careful_until = (k0 + W + 1) > (2 - k0) ? (k0 + W + 1) : (2 - k0);
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        tmax = max(tmax, __shfl_xor(tmax, d));
        tmin = min(tmin, __shfl_xor(tmin, d));
        careful_until = max(careful_until, __shfl_xor(careful_until, d));
    }
    // The steady loop starts on a block of 8 anti-diagonals once every pair's band has left the borders.
    careful_until = (careful_until + 7) & ~7;
    const int last_ref = n > 0 ? n - 1 : 0, last_qry = m > 0 ? m - 1 : 0;
    const int base_k = k0 + 2 * S * lane;  // this lane's lowest diagonal
    const int q_end = (end - k0) >> 1;     // the diagonal pair holding the corner
    const int first_q = S * lane;

    GroupState<S> st;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        st.He[s] = st.De[s] = st.Ie[s] = st.Ho[s] = st.Do[s] = st.Io[s] = TAG_INF;
        st.acc[s] = 0;
    }
    st.from_lower_i = TAG_INF;
    st.from_upper_d = TAG_INF;
#pragma unroll
    for (int s = 0; s < S; ++s) st.done[s][0] = st.done[s][1] = st.done[s][2] = 0;
    int corner = TAG_INF;
    uint4 *tiles = nullptr;  // ((t / 32) * (W / 2) + q): a tile of one diagonal pair
    if (RECORD) tiles = reinterpret_cast<uint4 *>(a.arena + (size_t)index * a.arena_stride);
    const int last_tile = tend >> 5;

    auto store_tile = [&](int tile) __attribute__((always_inline)) {
        if (RECORD && live && tile <= last_tile) {
#pragma unroll
            for (int s = 0; s < S; ++s)
                tiles[(size_t)tile * (W / 2) + first_q + s] =
                    make_uint4(st.done[s][0], st.done[s][1], st.done[s][2], st.acc[s]);
        }
    };
    // After odd anti-diagonal t = 8b + 7 block b is whole: kept, or the tile stored with it.
    auto finish_block = [&](int t) __attribute__((always_inline)) {
        const int block = t >> 3, place = block & 3;
        if (place == 3) {
            store_tile(block >> 2);
        } else {
#pragma unroll
            for (int s = 0; s < S; ++s) {
                st.done[s][0] = place == 0 ? st.acc[s] : st.done[s][0];
                st.done[s][1] = place == 1 ? st.acc[s] : st.done[s][1];
                st.done[s][2] = place == 2 ? st.acc[s] : st.done[s][2];
            }
        }
    };
    auto careful_pair = [&](int t) __attribute__((always_inline)) {
        st.from_lower_i = shift_from_lower<false>(st.from_lower_i, st.Io[S - 1]);
        even_half<S, RECORD, SH, true, STRIDE>(st, g, t, base_k, ref, qry, last_ref, last_qry, nullptr, nullptr,
                                               first_q, q_end, tend, corner);
        st.from_upper_d = shift_from_upper<false>(st.from_upper_d, st.De[0]);
        odd_half<S, RECORD, SH, true, STRIDE>(st, g, t + 1, base_k, ref, qry, last_ref, last_qry, nullptr, nullptr,
                                              first_q, q_end, tend, corner);
        if (((t + 1) & 7) == 7) finish_block(t + 1);
    };
    int t = 0;
    for (; t <= tmax && t < careful_until; t += 2) careful_pair(t);
    // Steady state: until the wave's first corner, blocks of 8 anti-diagonals, the letters from two cursors. On
    // even anti-diagonal t slot s reads reference letter r0 - 1 - s and query letter c0 - 1 + s (odd: c0 + s),
    // r0 and c0 the row and column of slot 0's even cell, both one more each anti-diagonal pair. Whole tiles run
    // four blocks at a time, their blocks' places known.
    if (t <= tmax) {
        int r0 = (t - base_k) >> 1, c0 = r0 + base_k;
        PackedState<S> ps;
        PackedCosts pk{};
        if constexpr (PACK) {
            pk.sub_eq = both_halves(g.sub_eq);
            pk.sub_gap = both_halves(g.sub_ne - g.sub_eq);
            pk.open_d = both_halves(g.open_d);
            pk.open_i = both_halves(g.open_i);
            pack_state<S>(ps, st, a.group16_top);
        }
        auto steady_block = [&]() __attribute__((always_inline)) {
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                if constexpr (PACK) {
                    const uint16_t *rp = reinterpret_cast<const uint16_t *>(ref) + r0 - 1;
                    const uint16_t *qp = reinterpret_cast<const uint16_t *>(qry) + c0 - 1;
                    packed_even<S, RECORD>(ps, st, pk, rp, qp);
                    packed_odd<S, RECORD>(ps, st, pk, rp, qp);
                } else {
                    const uint8_t *rp = ref + r0 - 1, *qp = qry + c0 - 1;
                    st.from_lower_i = shift_from_lower<false>(st.from_lower_i, st.Io[S - 1]);
                    even_half<S, RECORD, SH, false>(st, g, t, base_k, ref, qry, last_ref, last_qry, rp, qp, first_q,
                                                    q_end, tend, corner);
                    st.from_upper_d = shift_from_upper<false>(st.from_upper_d, st.De[0]);
                    odd_half<S, RECORD, SH, false>(st, g, t, base_k, ref, qry, last_ref, last_qry, rp, qp, first_q,
                                                   q_end, tend, corner);
                }
                ++r0;
                ++c0;
            }
        };
        for (; t + 7 < tmin && (t & 31) != 0; t += 8) {
            steady_block();
            finish_block(t + 7);
        }
        for (; t + 31 < tmin; t += 32) {
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                steady_block();
#pragma unroll
                for (int s = 0; s < S; ++s) st.done[s][b] = st.acc[s];
            }
            steady_block();
            store_tile(t >> 5);
        }
        for (; t + 7 < tmin; t += 8) {
            steady_block();
            finish_block(t + 7);
        }
        if constexpr (PACK) unpack_state<S>(st, ps);
    }
    // The tail: every pair's corner, with the borders' care.
    for (; t <= tmax; t += 2) careful_pair(t);
    // The corner's tile, unless the wave went past it (then it was stored whole): a part-filled last block's
    // nibbles sit in the word's top, shifted down into place.
    const int last = t - 1, filled = (last & 7) + 1, place = (last >> 3) & 3;
    if (RECORD && last_tile == (last >> 5) && !(filled == 8 && place == 3)) {
        if (filled != 8) {
#pragma unroll
            for (int s = 0; s < S; ++s) st.acc[s] >>= 4 * (8 - filled);
            finish_block(last | 7);  // keeps the part-filled block in its place, or stores the tile
        }
        // A tile the corner's block left open: stored with what it holds (its blocks past the corner's are never
        // read).
        if (place != 3) store_tile(last_tile);
    }

    const int holder = q_end / S;
    const int tagged = __shfl(corner, (threadIdx.x & 48) + holder);
    const int found = tagged >= TAG_INF / 2 ? INF : (tagged >> 4) + (SH ? a.costs.del_ext * tend : 0);
    if (!live || lane != 0) return;
    int want_low, want_high;
    if (!certify(a, n, m, found, k0, k0 + W - 1, &want_low, &want_high)) {
        file_retry(a, slot, n + m, want_low, want_high);
        return;
    }
    a.cost[slot] = found;
    a.status[slot] = 0;
    a.walk_low[slot] = k0;
    a.walk_class[slot] = a.cls;
    a.walk_flags[slot] = reinterpret_cast<uint64_t>(tiles);
}

// The same sweep over a wide band, one pair a block of G <= 8 waves: the band's W = 128SG diagonals dealt out as above
// to 64G lanes, lane j of the block holding diagonal pairs k0 + 2(Sj + s). Inside a wave the neighbours come over
// with wave_shr:1 / wave_shl:1; across a wave boundary through LDS, each half-step behind a barrier (the edge lane
// writes its value, the neighbour wave's edge lane reads it after the barrier; the next write of that slot comes
// after the other half-step's barrier, which the reader passes only once it has read).
template <int S, bool RECORD, bool SH>
__global__ __launch_bounds__(512) void sweep_block(SweepArgs a) {
    extern __shared__ uint8_t letters[];
    __shared__ int edge_i[8], edge_d[8], found_at;
    const int waves = (int)(blockDim.x >> 6), wave = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
    const int W = 128 * S * waves;
    const int index = blockIdx.x;
    const int slot = a.list[index];
    const int n = a.tape.rows[slot], m = a.tape.cols[slot];
    const int need_low = a.want_low[slot], need_high = a.want_high[slot];
    int k0 = need_low - (W - (need_high - need_low + 1)) / 2;
    k0 -= k0 & 1;
    const int end = m - n;
    const GroupCosts g = group_costs<RECORD, SH>(a.costs);

    uint8_t *ref = letters;
    uint8_t *qry = ref + n;
    {
        const int group = slot >> 6, column = slot & 63;
        const uint32_t *ref_words = a.tape.words + a.tape.group_base[group] + column;
        const uint32_t *qry_words = ref_words + (size_t)a.tape.group_ref_words[group] * 64;
        const int bits = a.tape.bits, per = 32 / bits;
        const uint32_t mask = (1u << bits) - 1;
        for (int i = threadIdx.x; i < n; i += blockDim.x)
            ref[i] = (uint8_t)((ref_words[(size_t)(i / per) * 64] >> (bits * (i % per))) & mask);
        for (int i = threadIdx.x; i < m; i += blockDim.x)
            qry[i] = (uint8_t)((qry_words[(size_t)(i / per) * 64] >> (bits * (i % per))) & mask);
    }
    __syncthreads();

    const int tend = n + m, tmax = tend, tmin = tend;
    int careful_until = (k0 + W + 1) > (2 - k0) ? (k0 + W + 1) : (2 - k0);
    careful_until = (careful_until + 7) & ~7;
    const int last_ref = n > 0 ? n - 1 : 0, last_qry = m > 0 ? m - 1 : 0;
    const int gl = (int)threadIdx.x;       // the lane within the pair's band
    const int base_k = k0 + 2 * S * gl;
    const int q_end = (end - k0) >> 1;
    const int first_q = S * gl;

    GroupState<S> st;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        st.He[s] = st.De[s] = st.Ie[s] = st.Ho[s] = st.Do[s] = st.Io[s] = TAG_INF;
        st.acc[s] = 0;
        st.done[s][0] = st.done[s][1] = st.done[s][2] = 0;
    }
    st.from_lower_i = TAG_INF;
    st.from_upper_d = TAG_INF;
    int corner = TAG_INF;
    uint4 *tiles = nullptr;
    if (RECORD) tiles = reinterpret_cast<uint4 *>(a.arena + (size_t)index * a.arena_stride);
    const int last_tile = tend >> 5;

    auto store_tile = [&](int tile) __attribute__((always_inline)) {
        if (RECORD && tile <= last_tile) {
#pragma unroll
            for (int s = 0; s < S; ++s)
                tiles[(size_t)tile * (W / 2) + first_q + s] =
                    make_uint4(st.done[s][0], st.done[s][1], st.done[s][2], st.acc[s]);
        }
    };
    auto finish_block = [&](int t) __attribute__((always_inline)) {
        const int block = t >> 3, place = block & 3;
        if (place == 3) {
            store_tile(block >> 2);
        } else {
#pragma unroll
            for (int s = 0; s < S; ++s) {
                st.done[s][0] = place == 0 ? st.acc[s] : st.done[s][0];
                st.done[s][1] = place == 1 ? st.acc[s] : st.done[s][1];
                st.done[s][2] = place == 2 ? st.acc[s] : st.done[s][2];
            }
        }
    };
    // The neighbours for the next half-steps: in-wave by DPP, across a wave boundary through LDS.
    auto take_lower = [&]() __attribute__((always_inline)) {
        if (waves > 1) {
            if (lane == 63) edge_i[wave] = st.Io[S - 1];
            __syncthreads();
        }
        st.from_lower_i = shift_from_lower<true>(st.from_lower_i, st.Io[S - 1]);
        if (waves > 1 && lane == 0 && wave > 0) st.from_lower_i = edge_i[wave - 1];
    };
    auto take_upper = [&]() __attribute__((always_inline)) {
        if (waves > 1) {
            if (lane == 0) edge_d[wave] = st.De[0];
            __syncthreads();
        }
        st.from_upper_d = shift_from_upper<true>(st.from_upper_d, st.De[0]);
        if (waves > 1 && lane == 63 && wave < waves - 1) st.from_upper_d = edge_d[wave + 1];
    };
    auto careful_pair = [&](int t) __attribute__((always_inline)) {
        take_lower();
        even_half<S, RECORD, SH, true>(st, g, t, base_k, ref, qry, last_ref, last_qry, nullptr, nullptr, first_q,
                                       q_end, tend, corner);
        take_upper();
        odd_half<S, RECORD, SH, true>(st, g, t + 1, base_k, ref, qry, last_ref, last_qry, nullptr, nullptr, first_q,
                                      q_end, tend, corner);
        if (((t + 1) & 7) == 7) finish_block(t + 1);
    };
    int t = 0;
    for (; t <= tmax && t < careful_until; t += 2) careful_pair(t);
    if (t <= tmax) {
        int r0 = (t - base_k) >> 1, c0 = r0 + base_k;
        auto steady_block = [&]() __attribute__((always_inline)) {
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                const uint8_t *rp = ref + r0 - 1, *qp = qry + c0 - 1;
                take_lower();
                even_half<S, RECORD, SH, false>(st, g, t, base_k, ref, qry, last_ref, last_qry, rp, qp, first_q,
                                                q_end, tend, corner);
                take_upper();
                odd_half<S, RECORD, SH, false>(st, g, t, base_k, ref, qry, last_ref, last_qry, rp, qp, first_q,
                                               q_end, tend, corner);
                ++r0;
                ++c0;
            }
        };
        // Blocks of 8 anti-diagonals (the barriers already order the half-steps; a longer unroll only keeps more
        // registers live).
        for (; t + 7 < tmin; t += 8) {
            steady_block();
            finish_block(t + 7);
        }
    }
    for (; t <= tmax; t += 2) careful_pair(t);
    const int last = t - 1, filled = (last & 7) + 1, place = (last >> 3) & 3;
    if (RECORD && last_tile == (last >> 5) && !(filled == 8 && place == 3)) {
        if (filled != 8) {
#pragma unroll
            for (int s = 0; s < S; ++s) st.acc[s] >>= 4 * (8 - filled);
            finish_block(last | 7);
        }
        if (place != 3) store_tile(last_tile);
    }

    // The corner's lane hands its cost to lane 0, which certifies the band.
    if (gl == q_end / S) found_at = corner;
    __syncthreads();
    if (gl != 0) return;
    const int tagged = found_at;
    const int found = tagged >= TAG_INF / 2 ? INF : (tagged >> 4) + (SH ? a.costs.del_ext * tend : 0);
    int want_low, want_high;
    if (!certify(a, n, m, found, k0, k0 + W - 1, &want_low, &want_high)) {
        file_retry(a, slot, n + m, want_low, want_high);
        return;
    }
    a.cost[slot] = found;
    a.status[slot] = 0;
    a.walk_low[slot] = k0;
    a.walk_class[slot] = a.cls;
    a.walk_flags[slot] = reinterpret_cast<uint64_t>(tiles);
}

// Launches a block class: S diagonal pairs a lane, `waves` waves a pair; `lds` one pair's letters and reads past.
void launch_block(int s, int waves, const SweepArgs &a, size_t lds, hipStream_t stream) {
    const dim3 grid(a.count), block(64 * waves);
    const bool shifted = a.costs.del_ext == a.costs.ins_ext;
#define DINARA_BLOCK(SV)                                                                                       \
    if (a.record) {                                                                                          \
        if (shifted) hipLaunchKernelGGL((sweep_block<SV, true, true>), grid, block, lds, stream, a);           \
        else hipLaunchKernelGGL((sweep_block<SV, true, false>), grid, block, lds, stream, a);                 \
    } else {                                                                                                 \
        if (shifted) hipLaunchKernelGGL((sweep_block<SV, false, true>), grid, block, lds, stream, a);          \
        else hipLaunchKernelGGL((sweep_block<SV, false, false>), grid, block, lds, stream, a);                \
    }
    switch (s) {
        case 4: DINARA_BLOCK(4) break;
        case 8: DINARA_BLOCK(8) break;
        // Distances only past S = 8: a band this wide keeps no decisions (its record would spill registers).
        case 12:
            if (!a.record) {
                if (shifted) hipLaunchKernelGGL((sweep_block<12, false, true>), grid, block, lds, stream, a);
                else hipLaunchKernelGGL((sweep_block<12, false, false>), grid, block, lds, stream, a);
            }
            break;
        case 16:
            if (!a.record) {
                if (shifted) hipLaunchKernelGGL((sweep_block<16, false, true>), grid, block, lds, stream, a);
                else hipLaunchKernelGGL((sweep_block<16, false, false>), grid, block, lds, stream, a);
            }
            break;
        default: break;
    }
#undef DINARA_BLOCK
}

template <int S>
static void launch_group_s(const SweepArgs &a, size_t lds, hipStream_t stream) {
    const dim3 grid((a.count + 3) / 4), block(64);
    const bool shifted = a.costs.del_ext == a.costs.ins_ext;
    const bool packed = S >= 2 && shifted && a.group16_top > 0;
    if (a.record) {
        if (packed) hipLaunchKernelGGL((sweep_group<S, true, true, S >= 2>), grid, block, lds, stream, a);
        else if (shifted) hipLaunchKernelGGL((sweep_group<S, true, true, false>), grid, block, lds, stream, a);
        else hipLaunchKernelGGL((sweep_group<S, true, false, false>), grid, block, lds, stream, a);
    } else {
        if (packed) hipLaunchKernelGGL((sweep_group<S, false, true, S >= 2>), grid, block, lds, stream, a);
        else if (shifted) hipLaunchKernelGGL((sweep_group<S, false, true, false>), grid, block, lds, stream, a);
        else hipLaunchKernelGGL((sweep_group<S, false, false, false>), grid, block, lds, stream, a);
    }
}

// Launches a group class: S = 1, 2, 4 or 8 (32 to 256 diagonals); `lds` four pairs' letters (16-bit positions when
// the batch's group16_top is set and S >= 2).
void launch_group(int s, const SweepArgs &a, size_t lds, hipStream_t stream) {
    switch (s) {
        case 1: launch_group_s<1>(a, lds, stream); break;
        case 2: launch_group_s<2>(a, lds, stream); break;
        case 4: launch_group_s<4>(a, lds, stream); break;
        case 8: launch_group_s<8>(a, lds, stream); break;
        default: break;
    }
}

}  // namespace dinara
