// Gotoh's recurrence under a substitution table (Scoring), one wave a pair, on MI355X (gfx950).
//
// Scores are maximized (reference gotoh.mojo:265-352): D(r, c) = max(H(r-1, c) + open, D(r-1, c) + extend),
// I(r, c) = max(H(r, c-1) + open, I(r, c-1) + extend), H(r, c) = max(H(r-1, c-1) + table[a][b], D, I), a local
// alignment floored at zero. `open` scores a gap's first letter, `extend` each later one.
//
// Layout. A wave sweeps a strip of 64C columns, lane j holding columns jC + 1 .. jC + C, one row behind the lane
// to its left: at step s lane j takes row s - j. The cell left of a lane's first column is the left lane's last,
// of the same row, which that lane made the step before; one DPP wave_shr:1 hands it over, and the cell above-left
// is what came over the step before. A strip of more than 64C columns is swept strip after strip, the last column
// of each kept for the next. The table is staged in LDS as a query profile: for every letter a and every lane its C
// columns' scores table[a][query letter], 8 bytes, so a lane reads a row's scores with one 8-byte LDS load.
//
// Modes. SCORE_GLOBAL: H(rows, cols). SCORE_LOCAL: the best cell, above zero, by the host's rule (reference
// scored.mojo:81-84): the highest score, then the furthest along both sequences, then along the reference.
// SCORE_REACH (reference vector_score.mojo:37-162): over both prefixes reversed from a local end, the global score
// anchored there; the first cell, by anti-diagonal and then by row, that reaches the target, with the sweep cut
// off once no cell left can come earlier. SCORE_TRACE: H(rows, cols) and every cell's decision (reference
// gotoh.mojo:129-192): bits 0-1 the source (0 aligned, 1 the deletion layer, 2 the insertion layer; on a tie the
// aligned first, then the deletion), bit 2 the deletion extends into the cell, bit 3 the insertion does (each only
// when strictly better than opening). score_walk then walks them back as the reference's reconstruct does
// (gotoh.mojo:355-442).

#include "cigar.h"
#include "score.h"

namespace dinara {

// Lane j reads lane j - 1's `v`; lane 0 keeps `old`.
__device__ __forceinline__ int from_left_lane(int old, int v) {
    return __builtin_amdgcn_update_dpp(old, v, 0x138, 0xF, 0xF, false);
}

__device__ __forceinline__ int gap_score(const ScoreArgs &a, int letters) {
    return letters <= 0 ? 0 : a.open + (letters - 1) * a.extend;
}

// Whether (h, r, c) beats (bh, br, bc) as a local end: higher, then further along both, then along the reference.
__device__ __forceinline__ bool better_end(int h, int r, int c, int bh, int br, int bc) {
    if (h != bh) return h > bh;
    if (r + c != br + bc) return r + c > br + bc;
    return r > br;
}

// Whether (r, c) comes before (br, bc) as a reach's start: an earlier anti-diagonal, then an earlier row.
__device__ __forceinline__ bool earlier_start(int r, int c, int br, int bc) {
    if (r + c != br + bc) return r + c < br + bc;
    return r < br;
}

// KEYED (SCORE_LOCAL): a column's best is one unsigned key, score << end_bits | row, so one max a cell keeps the
// highest score and, among equal scores, the latest row, as the compare-and-select form does (the host picks it
// where every score and row fit the key).
template <int C, int MODE, bool KEYED = false>
__global__ __launch_bounds__(64) void score_wave(ScoreArgs a) {
    constexpr bool LOCAL = MODE == SCORE_LOCAL;
    constexpr bool REACH = MODE == SCORE_REACH;
    constexpr bool TRACE = MODE == SCORE_TRACE;
    extern __shared__ uint8_t lds[];
    const int lane = threadIdx.x;
    const int index = blockIdx.x;
    const int slot = a.list[index];
    const int32_t *view = a.view + 4 * (size_t)slot;
    const int rows = view[1], cols = view[3];
    const uint8_t *ref_codes = a.codes + a.ref_at[slot] + view[0];
    const uint8_t *qry_codes = a.codes + a.qry_at[slot] + view[2];
    uint2 *profile = reinterpret_cast<uint2 *>(lds);
    uint8_t *ref = lds + (size_t)a.letters * 64 * 8;
    // Four bytes of the reference's letters from `at` on, as the views read them (zero past the end).
    auto letters4 = [&](int at) {
        uint32_t w = 0;
#pragma unroll
        for (int t = 0; t < 4; ++t)
            if (at + t < rows) w |= (uint32_t)(a.reversed ? ref_codes[rows - 1 - at - t] : ref_codes[at + t]) << (8 * t);
        return w;
    };
    // A reference too long for LDS goes through a ring of SCORE_RING_CHUNKS chunks instead of being staged whole.
    const bool ring = a.ring != 0;
    const int ref_mask = ring ? SCORE_RING_CHUNKS * SCORE_RING_CHUNK - 1 : 0x7fffffff;
    if (!ring)
        for (int i = lane; i < rows; i += 64) ref[i] = a.reversed ? ref_codes[rows - 1 - i] : ref_codes[i];

    const int target = REACH ? a.target[slot] : 0;
    const int strip_cols = 64 * C;
    const int strips = (cols + strip_cols - 1) / strip_cols;
    int32_t *carry = a.carry ? a.carry + (size_t)index * a.carry_stride : nullptr;
    uint32_t *decisions = TRACE ? a.arena + (size_t)index * a.arena_stride : nullptr;

    int corner = SCORE_NEG;
    int best_h = 0, best_r = 0, best_c = 0;          // SCORE_LOCAL: this lane's best end so far
    int first_r = 0, first_c = 0;                    // SCORE_REACH: this lane's earliest start so far (0: none)
    int cutoff = 0x7fffffff;                         // SCORE_REACH: the wave's earliest start's anti-diagonal

    for (int strip = 0; strip < strips; ++strip) {
        // A later strip's columns all lie past the earliest start found: none of them can come before it.
        if (REACH && cutoff < strip * strip_cols + 2) break;
        const int c0 = strip * strip_cols + lane * C;  // the column left of the lane's first
        __syncthreads();
        for (int letter = 0; letter < a.letters; ++letter) {
            uint32_t lo = 0, hi = 0;
#pragma unroll
            for (int i = 0; i < C; ++i) {
                const int col = c0 + i + 1;
                const int q = col <= cols ? (a.reversed ? qry_codes[cols - col] : qry_codes[col - 1]) : 0;
                const uint32_t s = (uint32_t)(uint8_t)a.table[letter * 32 + q];
                if (i < 4) lo |= s << (8 * i);
                else hi |= s << (8 * (i - 4));
            }
            profile[letter * 64 + lane] = make_uint2(lo, hi);
        }
        // Ring: chunk k of the letters (SCORE_RING_CHUNK of them from k * SCORE_RING_CHUNK) sits in ring place
        // k mod SCORE_RING_CHUNKS. Step s reads letters s - 63 .. s, so while steps 256k + 1 .. 256k + 256 run
        // chunks k - 1, k and k + 1 are held; chunk k + 1 is written at step 256k + 1, from a register loaded a
        // range of steps before, and chunk k + 2 loaded into that register.
        uint32_t pending = 0;
        if (ring) {
            static_assert(SCORE_RING_CHUNK == 4 * 64, "a chunk is four letters a lane");
            reinterpret_cast<uint32_t *>(ref)[lane] = letters4(4 * lane);
            reinterpret_cast<uint32_t *>(ref)[64 + lane] = letters4(SCORE_RING_CHUNK + 4 * lane);
            pending = letters4(2 * SCORE_RING_CHUNK + 4 * lane);
        }
        __syncthreads();

        int Hp[C], Dp[C];  // the row above, the lane's columns
#pragma unroll
        for (int i = 0; i < C; ++i) {
            Hp[i] = LOCAL ? 0 : gap_score(a, c0 + i + 1);
            Dp[i] = SCORE_NEG;
        }
        int left_prev = LOCAL ? 0 : gap_score(a, c0);  // H(r - 1, c0), the cell above-left of the first column
        int out_h = SCORE_NEG, out_i = SCORE_NEG;      // H and I of the lane's last column, the row just made
        int col_h[C], col_r[C];                        // per column: SCORE_LOCAL's best, SCORE_REACH's first row
#pragma unroll
        for (int i = 0; i < C; ++i) col_h[i] = col_r[i] = 0;

        const int steps = rows + 63;
        uint2 prof = make_uint2(0, 0);
        {
            const int r = 1 - lane;
            if (r >= 1 && r <= rows) prof = profile[ref[(r - 1) & ref_mask] * 64 + lane];
        }
        for (int s = 1; s <= steps; ++s) {
            const int r = s - lane;
            if (ring && (s & (SCORE_RING_CHUNK - 1)) == 1 && s > 1) {
                const int k = s / SCORE_RING_CHUNK;
                reinterpret_cast<uint32_t *>(ref)[((k + 1) & (SCORE_RING_CHUNKS - 1)) * 64 + lane] = pending;
                pending = letters4((k + 2) * SCORE_RING_CHUNK + 4 * lane);
                __syncthreads();
            }
            int lh = from_left_lane(SCORE_NEG, out_h), li = from_left_lane(SCORE_NEG, out_i);
            if (lane == 0) {
                if (strip == 0 || r < 1 || r > rows) {
                    lh = LOCAL ? 0 : gap_score(a, r);
                    li = SCORE_NEG;
                } else {
                    lh = carry[2 * r];
                    li = carry[2 * r + 1];
                }
            }
            // The next step's row, its profile read ahead.
            uint2 next = make_uint2(0, 0);
            if (r + 1 >= 1 && r + 1 <= rows) next = profile[ref[r & ref_mask] * 64 + lane];
            if (r >= 1 && r <= rows) {
                uint32_t word = 0;
                int hl = lh, il = li, hd = left_prev;
#pragma unroll
                for (int i = 0; i < C; ++i) {
                    const int hu = Hp[i], du = Dp[i];
                    const int d = max(hu + a.open, du + a.extend);
                    const int ins = max(hl + a.open, il + a.extend);
                    const uint32_t bytes = i < 4 ? prof.x : prof.y;
                    const int sub = hd + __builtin_amdgcn_sbfe(bytes, 8 * (i & 3), 8);
                    int h = max(sub, max(d, ins));
                    if (LOCAL) h = max(h, 0);
                    if (TRACE) {
                        const uint32_t src = h == sub ? 0u : (h == d ? 1u : 2u);
                        const uint32_t nib = src | (du + a.extend > hu + a.open ? 4u : 0u) |
                                             (il + a.extend > hl + a.open ? 8u : 0u);
                        word |= nib << (4 * i);
                    }
                    if constexpr (LOCAL && KEYED) {
                        col_h[i] = (int)max((uint32_t)col_h[i], ((uint32_t)h << a.end_bits) | (uint32_t)r);
                    } else if (LOCAL && h >= col_h[i] && h > 0) {
                        col_h[i] = h;
                        col_r[i] = r;
                    }
                    if (REACH && col_r[i] == 0 && h >= target) col_r[i] = r;
                    if ((MODE == SCORE_GLOBAL || TRACE) && r == rows && c0 + i + 1 == cols) corner = h;
                    hd = hu;
                    hl = h;
                    il = ins;
                    Hp[i] = h;
                    Dp[i] = d;
                }
                out_h = hl;
                out_i = il;
                left_prev = lh;
                if (carry && lane == 63 && strip + 1 < strips) {
                    carry[2 * r] = out_h;
                    carry[2 * r + 1] = out_i;
                }
                if (TRACE) decisions[((size_t)strip * (rows + 1) + r) * 64 + lane] = word;
            }
            prof = next;
            if (REACH && (s & 31) == 0) {
                // The earliest start found in the wave so far; every cell still to come lies on anti-diagonal
                // s + 2 or later, so a start before that is final. Columns past the query's end are swept too (their
                // letters stand in) and may reach the target first: they are no start.
                int diag = 0x7fffffff;
#pragma unroll
                for (int i = 0; i < C; ++i)
                    if (col_r[i] != 0 && c0 + i + 1 <= cols) diag = min(diag, col_r[i] + c0 + i + 1);
                if (first_r != 0) diag = min(diag, first_r + first_c);
#pragma unroll
                for (int d = 1; d < 64; d <<= 1) diag = min(diag, __shfl_xor(diag, d));
                cutoff = diag;
                if (diag < s + 2) break;
            }
        }
        // Fold the strip's columns into the lane's own candidate.
#pragma unroll
        for (int i = 0; i < C; ++i) {
            const int col = c0 + i + 1;
            if (col > cols) continue;
            if (LOCAL && KEYED) {
                col_r[i] = (int)((uint32_t)col_h[i] & ((1u << a.end_bits) - 1u));
                col_h[i] = (int)((uint32_t)col_h[i] >> a.end_bits);
            }
            if (LOCAL && col_h[i] > 0 && better_end(col_h[i], col_r[i], col, best_h, best_r, best_c)) {
                best_h = col_h[i];
                best_r = col_r[i];
                best_c = col;
            }
            if (REACH && col_r[i] != 0 && (first_r == 0 || earlier_start(col_r[i], col, first_r, first_c))) {
                first_r = col_r[i];
                first_c = col;
            }
        }
        if (REACH) {
            int diag = first_r != 0 ? first_r + first_c : 0x7fffffff;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) diag = min(diag, __shfl_xor(diag, d));
            cutoff = diag;
        }
        if (carry) __threadfence_block();
    }

    if (MODE == SCORE_GLOBAL || TRACE) {
        const int owner = ((cols - 1) % strip_cols) / C;
        if (lane == owner) a.score[slot] = corner;
        if (TRACE && lane == 0) a.walk_base[slot] = reinterpret_cast<uint64_t>(decisions);
    }
    if (LOCAL) {
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int oh = __shfl_xor(best_h, d), orow = __shfl_xor(best_r, d), ocol = __shfl_xor(best_c, d);
            if (oh > 0 && (best_h == 0 || better_end(oh, orow, ocol, best_h, best_r, best_c))) {
                best_h = oh;
                best_r = orow;
                best_c = ocol;
            }
        }
        if (lane == 0) {
            a.score[slot] = best_h;
            a.end_row[slot] = best_r;
            a.end_col[slot] = best_c;
        }
    }
    if (REACH) {
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int orow = __shfl_xor(first_r, d), ocol = __shfl_xor(first_c, d);
            if (orow != 0 && (first_r == 0 || earlier_start(orow, ocol, first_r, first_c))) {
                first_r = orow;
                first_c = ocol;
            }
        }
        if (lane == 0) {
            a.end_row[slot] = first_r;
            a.end_col[slot] = first_c;
        }
    }
}

// The walk back through SCORE_TRACE's decisions, one thread a pair (reference gotoh.mojo:355-442): from the
// aligned layer follow the source; a gap layer keeps stepping while the cell says it extends; at row 0 or column 0
// the rest is gaps. Then every lane's CIGAR goes into the shared output.
__global__ __launch_bounds__(256) void score_walk(ScoreArgs a) {
    const int index = blockIdx.x * 256 + threadIdx.x;
    int64_t moves = -1;
    int slot = 0;
    if (index < a.count) {
        slot = a.list[index];
        const int32_t *view = a.view + 4 * (size_t)slot;
        const int rows = view[1], cols = view[3];
        const uint8_t *ref = a.codes + a.ref_at[slot] + view[0];
        const uint8_t *qry = a.codes + a.qry_at[slot] + view[2];
        const uint32_t *decisions = reinterpret_cast<const uint32_t *>(a.walk_base[slot]);
        const int strip_cols = 64 * a.c;
        OpTape ops{reinterpret_cast<uint64_t *>(a.op_scratch + a.op_base[slot]), 0, 0};
        int r = rows, c = cols, layer = 0;
        while (r > 0 && c > 0) {
            const int col = c - 1, strip = col / strip_cols, within = col - strip * strip_cols;
            const uint32_t word = decisions[((size_t)strip * (rows + 1) + r) * 64 + within / a.c];
            const uint32_t nib = (word >> (4 * (within % a.c))) & 15u;
            const uint32_t state = layer == 0 ? (nib & 3u) : (uint32_t)layer;
            if (state == 1) {
                ops.put(2u);
                --r;
                layer = (nib & 4u) ? 1 : 0;
            } else if (state == 2) {
                ops.put(3u);
                --c;
                layer = (nib & 8u) ? 2 : 0;
            } else {
                ops.put(ref[r - 1] == qry[c - 1] ? 0u : (a.eqx ? 1u : 0u));
                --r;
                --c;
                layer = 0;
            }
        }
        for (; r > 0; --r) ops.put(2u);
        for (; c > 0; --c) ops.put(3u);
        ops.close();
        moves = ops.count;
    }
    const uint64_t *words = reinterpret_cast<const uint64_t *>(a.op_scratch + (moves >= 0 ? a.op_base[slot] : 0));
    publish_wave(a.eqx, 0, words, moves, a.cigars, a.cigar_used, a.cigar_offset + slot, a.cigar_length + slot);
}

// ------------------------------------------------------------------ launches

size_t score_lds(int letters, int rows, bool ring) {
    return (size_t)letters * 64 * 8 + (ring ? (size_t)SCORE_RING_CHUNKS * SCORE_RING_CHUNK : (size_t)rows);
}

template <int MODE, bool KEYED = false>
static void launch_mode(const ScoreArgs &a, size_t lds, hipStream_t stream) {
    const dim3 grid(a.count), block(64);
    switch (a.c) {
        case 1: hipLaunchKernelGGL((score_wave<1, MODE, KEYED>), grid, block, lds, stream, a); break;
        case 2: hipLaunchKernelGGL((score_wave<2, MODE, KEYED>), grid, block, lds, stream, a); break;
        case 3: hipLaunchKernelGGL((score_wave<3, MODE, KEYED>), grid, block, lds, stream, a); break;
        case 4: hipLaunchKernelGGL((score_wave<4, MODE, KEYED>), grid, block, lds, stream, a); break;
        case 5: hipLaunchKernelGGL((score_wave<5, MODE, KEYED>), grid, block, lds, stream, a); break;
        case 6: hipLaunchKernelGGL((score_wave<6, MODE, KEYED>), grid, block, lds, stream, a); break;
        case 7: hipLaunchKernelGGL((score_wave<7, MODE, KEYED>), grid, block, lds, stream, a); break;
        default: hipLaunchKernelGGL((score_wave<8, MODE, KEYED>), grid, block, lds, stream, a); break;
    }
}

void launch_score(const ScoreArgs &a, size_t lds, hipStream_t stream) {
    if (a.count <= 0) return;
    switch (a.mode) {
        case SCORE_GLOBAL: launch_mode<SCORE_GLOBAL>(a, lds, stream); break;
        case SCORE_LOCAL:
            if (a.end_bits > 0) launch_mode<SCORE_LOCAL, true>(a, lds, stream);
            else launch_mode<SCORE_LOCAL>(a, lds, stream);
            break;
        case SCORE_REACH: launch_mode<SCORE_REACH>(a, lds, stream); break;
        default: launch_mode<SCORE_TRACE>(a, lds, stream); break;
    }
}

void launch_score_walk(const ScoreArgs &a, hipStream_t stream) {
    if (a.count <= 0) return;
    hipLaunchKernelGGL(score_walk, dim3((a.count + 255) / 256), dim3(256), 0, stream, a);
}

}  // namespace dinara
