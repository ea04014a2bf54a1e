// Banded global alignment under gap-affine costs on MI355X (gfx950): the band certificate, the wave-per-pair
// sweep, the walk back and the CIGAR text. The sixteen-lane group sweep, which takes most pairs, is group.hip.
//
// Every sweep runs Gotoh's recurrence over a band of diagonals, each cell keeping the decision the reference's
// tie rule reads back (reference lanes.mojo:795-947 recurrence, :905-923 flags, :1410-1460 walk), and proves its
// band by the cost it finds (reference lanes.mojo:1337-1392): a band no path off which costs as little (for
// alignments) or less (for distances) settles the pair; any other is filed, with the band its cost proves, for a
// wider class. Every settled alignment is walked back through its decisions by walk_bands, one thread a pair,
// and its CIGAR written as text, then compacted into one buffer.
//
// sweep_wave<Q>: one wave per pair, lane j holding diagonals lo + jQ .. lo + jQ + Q - 1 (a band of 64Q), a row
// at a time. The deletion layer reads the row above's next diagonal across one lane; the insertion layer, a run
// along the row, is a prefix minimum over the row's diagonals: in-lane, then across the wave. It takes the bands
// wider than a group's and every pair under a user band.
//
// Decisions are 4 bits a cell: bits 0-1 the source the aligned layer takes (0 two equal letters aligned,
// 3 two unequal, 1 the deletion layer, 2 the insertion layer), bit 2 the deletion extends into the cell,
// bit 3 the insertion does.

#include "bands.h"

namespace dinara {

// ------------------------------------------------------------------ one wave per pair

// P = 2: two gap pieces (reference lanes.mojo:795-947 with pieces == 2). Each piece's deletion layer comes from the
// row above as the first's does; each piece's insertion run is a prefix minimum of its own along the row. A run
// mixing the two pieces never costs less than the better piece alone (its openings include the cheaper-extending
// piece's, its letters cost no less), so the cells' values, and the decisions on every path the walk can take,
// are the recurrence's. Decisions are then a byte a cell: bits 0-2 the source (0 two equal letters, 5 two unequal,
// 1 and 2 the first piece's deletion and insertion layers, 3 and 4 the second's), bit 2 + L layer L extends.
// FREE: free ends. The slot sweeps a view of its pair (SweepArgs: view_*), cells (r, 0) with r up to its free
// leading reference letters and (0, c) with c up to its free leading query letters start an alignment for nothing,
// and the least cost is taken over the cells it may end on: column m in the last free reference letters' rows,
// row n in the last free query letters' columns (ties to the highest diagonal, reference letters less query
// letters, or with `lowest_end` the lowest); reference tests/test_alignment.mojo:1944-2020.
template <int Q, bool RECORD, int P, bool FREE>
__global__ __launch_bounds__(64) void sweep_wave(SweepArgs a) {
    extern __shared__ uint8_t letters[];
    const int t = blockIdx.x;
    const int slot = a.list[t];
    const int lane = threadIdx.x;
    const int n = FREE ? a.view_rows[slot] : a.tape.rows[slot], m = FREE ? a.view_cols[slot] : a.tape.cols[slot];
    constexpr int W = 64 * Q;
    const int need_low = a.want_low[slot], need_high = a.want_high[slot];
    const int lo = need_low - (W - (need_high - need_low + 1)) / 2;
    const int end = m - n;
    const CostParams k = a.costs;
    int fr[4] = {0, 0, 0, 0};  // free letters: reference start, end, query start, end
    if (FREE) {
#pragma unroll
        for (int f = 0; f < 4; ++f) fr[f] = a.free[4 * slot + f];
    }

    // Both sequences' letters into LDS, a byte each: the reference's first, then the query's.
    {
        const int group = slot >> 6, column = slot & 63;
        const uint32_t *ref_words = a.tape.words + a.tape.group_base[group] + column;
        const uint32_t *qry_words = ref_words + (size_t)a.tape.group_ref_words[group] * 64;
        const int bits = a.tape.bits, per = 32 / bits;
        const uint32_t mask = (1u << bits) - 1;
        const int ref_at = FREE ? a.view_ref_at[slot] : 0, qry_at = FREE ? a.view_qry_at[slot] : 0;
        const bool back = FREE && a.view_reversed;
        for (int i = lane; i < n; i += 64) {
            const int x = back ? ref_at + n - 1 - i : ref_at + i;
            letters[i] = (uint8_t)((ref_words[(size_t)(x / per) * 64] >> (bits * (x % per))) & mask);
        }
        for (int i = lane; i < m; i += 64) {
            const int x = back ? qry_at + m - 1 - i : qry_at + i;
            letters[n + i] = (uint8_t)((qry_words[(size_t)(x / per) * 64] >> (bits * (x % per))) & mask);
        }
    }
    __syncthreads();

    const int first = lo + lane * Q;  // this lane's lowest diagonal
    int H[Q], D[Q], D2[Q];
    // Free ends: this lane's best end so far, and its lane diagonal (query letters less reference letters).
    int best_end = INF, best_diagonal = 0;
    auto consider = [&](int value, int kk) {
        // A higher library diagonal (reference less query) is a lower lane diagonal.
        const bool nearer = a.lowest_end ? kk > best_diagonal : kk < best_diagonal;
        if (value < best_end || (value == best_end && value < INF && nearer)) {
            best_end = value;
            best_diagonal = kk;
        }
    };
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const int kk = first + i;
        const int from = FREE ? fr[2] : 0;  // row 0: the free leading query letters cost nothing
        int top = kk <= m ? k.ins_open + (kk - from) * k.ins_ext : INF;
        if (P == 2 && kk <= m) top = imin(top, k.ins_open2 + (kk - from) * k.ins_ext2);
        H[i] = kk < 0 ? INF : (kk <= from ? 0 : top);
        D[i] = INF;
        D2[i] = INF;
        if (FREE && kk >= 0 && kk <= m && ((n == 0 && kk >= m - fr[3]) || (kk == m && n <= fr[1])))
            consider(H[i], kk);  // row 0's cells an alignment may end on
    }
    const bool barred_band = a.user_band != 0;
    const int user_low = a.user_low, user_high = a.user_high;
    uint8_t *flags = nullptr;
    if (RECORD) flags = a.arena + (size_t)t * a.arena_stride;  // W bytes a row

    for (int r = 1; r <= n; ++r) {
        const uint8_t letter = letters[r - 1];
        // The row above's next diagonal, for the lane's last one: the next lane's first.
        int hu_next = __shfl_down(H[0], 1);
        int du_next = __shfl_down(D[0], 1);
        int du2_next = P == 2 ? __shfl_down(D2[0], 1) : INF;
        if (lane == 63) hu_next = du_next = du2_next = INF;
        int T[Q], SUB[Q], DN[Q], DN2[Q], X[Q], X2[Q];
        bool EQ[Q];
        bool ED_LE[Q], ED2_LE[Q];
        const int skipped = FREE ? r - fr[0] : r;  // column 0: the free leading reference letters cost nothing
        int border = k.del_open + skipped * k.del_ext;
        if (P == 2) border = imin(border, k.del_open2 + skipped * k.del_ext2);
        if (FREE && skipped <= 0) border = 0;
        int aggregate = INF, aggregate2 = INF;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const int kk = first + i;
            const int c = r + kk;
            const int hu = i + 1 < Q ? H[i + 1] : hu_next;
            const int du = i + 1 < Q ? D[i + 1] : du_next;
            const int od = hu + k.del_first, ed = du + k.del_ext;
            const int dn = imin(od, ed);
            int dn2 = INF;
            bool ed2_le = false;
            if (P == 2) {
                const int du2 = i + 1 < Q ? D2[i + 1] : du2_next;
                const int od2 = hu + k.del_first2, ed2 = du2 + k.del_ext2;
                dn2 = imin(od2, ed2);
                ed2_le = ed2 <= od2;
            }
            const uint8_t b = (c >= 1 && c <= m) ? letters[n + c - 1] : (uint8_t)0xFF;
            const bool eq = b == letter;
            const int sub = H[i] + (eq ? 0 : k.mismatch);
            int tt = imin(sub, imin(dn, dn2));
            const bool barred = barred_band && (kk < user_low || kk > user_high);
            if (c <= 0) tt = c == 0 ? border : INF;
            if (barred) tt = INF;
            T[i] = tt;
            SUB[i] = sub;
            DN[i] = (c <= 0 || barred) ? INF : dn;
            DN2[i] = (c <= 0 || barred) ? INF : dn2;
            EQ[i] = eq;
            ED_LE[i] = ed <= od;
            ED2_LE[i] = ed2_le;
            // The insertion run from diagonal kk onward: its cost less kk extensions, so the run's
            // extensions cancel and the row's insertions are a prefix minimum (the row's own index drops out).
            X[i] = tt - kk * k.ins_ext;
            aggregate = imin(aggregate, X[i]);
            if (P == 2) {
                X2[i] = tt - kk * k.ins_ext2;
                aggregate2 = imin(aggregate2, X2[i]);
            }
        }
        // Inclusive minimum across the wave, then shifted by one lane: the best run start below this lane.
        int scan = aggregate, scan2 = aggregate2;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int other = __shfl_up(scan, d);
            if (lane >= d) scan = imin(scan, other);
            if (P == 2) {
                const int other2 = __shfl_up(scan2, d);
                if (lane >= d) scan2 = imin(scan2, other2);
            }
        }
        int before = __shfl_up(scan, 1);
        int before2 = P == 2 ? __shfl_up(scan2, 1) : INF;
        if (lane == 0) before = before2 = INF + INF / 2;
        int I[Q], I2[Q], HN[Q];
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const int kk = first + i;
            // A gap opened after the best start below: opening plus one extension a letter.
            int ii = before + (k.ins_first - k.ins_ext) + kk * k.ins_ext;
            ii = ii < INF ? ii : INF;
            before = imin(before, X[i]);
            const int c = r + kk;
            const bool off = c <= 0 || (barred_band && (kk < user_low || kk > user_high));
            if (off) ii = INF;
            I[i] = ii;
            int h = imin(T[i], ii);
            if (P == 2) {
                int i2 = before2 + (k.ins_first2 - k.ins_ext2) + kk * k.ins_ext2;
                i2 = i2 < INF ? i2 : INF;
                before2 = imin(before2, X2[i]);
                if (off) i2 = INF;
                I2[i] = i2;
                h = imin(h, i2);
            }
            if (barred_band && (kk < user_low || kk > user_high)) h = INF;
            HN[i] = h;
            if (FREE && c >= 0 && c <= m && ((c == m && r >= n - fr[1]) || (r == n && c >= m - fr[3])))
                consider(c == 0 ? T[i] : h, kk);
        }
        if (RECORD) {
            int hl_prev = __shfl_up(HN[Q - 1], 1);
            int il_prev = __shfl_up(I[Q - 1], 1);
            int il2_prev = P == 2 ? __shfl_up(I2[Q - 1], 1) : INF;
            if (lane == 0) hl_prev = il_prev = il2_prev = INF;
            uint32_t packed = 0;
#pragma unroll
            for (int i = 0; i < Q; ++i) {
                const int hl = i > 0 ? HN[i - 1] : hl_prev;
                const int il = i > 0 ? I[i - 1] : il_prev;
                const int oi = hl + k.ins_first, ei = il + k.ins_ext;
                const int h = HN[i];
                uint32_t flag;
                if (P == 1) {
                    // The lowest priority first, each later source as cheap as the cell taking over.
                    uint32_t src = EQ[i] ? 0u : 3u;
                    src = I[i] == h ? 2u : src;
                    src = DN[i] == h ? 1u : src;
                    src = (!EQ[i] && SUB[i] == h) ? 3u : src;
                    flag = src | (ED_LE[i] ? 4u : 0u) | (ei <= oi ? 8u : 0u);
                } else {
                    const int il2 = i > 0 ? I2[i - 1] : il2_prev;
                    const int oi2 = hl + k.ins_first2, ei2 = il2 + k.ins_ext2;
                    uint32_t src = EQ[i] ? 0u : 5u;
                    src = I[i] == h ? 2u : src;
                    src = I2[i] == h ? 4u : src;
                    src = DN[i] == h ? 1u : src;
                    src = DN2[i] == h ? 3u : src;
                    src = (!EQ[i] && SUB[i] == h) ? 5u : src;
                    flag = src | (ED_LE[i] ? 8u : 0u) | (ei <= oi ? 16u : 0u) | (ED2_LE[i] ? 32u : 0u) |
                           (ei2 <= oi2 ? 64u : 0u);
                }
                if (Q % 4 == 0) {
                    packed |= flag << (8 * (i % 4));
                    if (i % 4 == 3) {
                        *reinterpret_cast<uint32_t *>(flags + (size_t)(r - 1) * W + lane * Q + (i - 3)) = packed;
                        packed = 0;
                    }
                } else {
                    flags[(size_t)(r - 1) * W + lane * Q + i] = (uint8_t)flag;
                }
            }
        }
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            H[i] = HN[i];
            D[i] = DN[i];
            if (P == 2) D2[i] = DN2[i];
        }
    }

    int found;
    if (FREE) {
        // The wave's best end: the least cost, then the preferred diagonal.
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int other = __shfl_xor(best_end, d), other_diagonal = __shfl_xor(best_diagonal, d);
            const bool nearer = a.lowest_end ? other_diagonal > best_diagonal : other_diagonal < best_diagonal;
            if (other < best_end || (other == best_end && other < INF && nearer)) {
                best_end = other;
                best_diagonal = other_diagonal;
            }
        }
        found = best_end;
    } else {
        // The corner's cost, from the lane holding diagonal `end`.
        int mine = INF;
#pragma unroll
        for (int i = 0; i < Q; ++i)
            if (first + i == end) mine = H[i];
        const int holder = (end - lo) / Q;
        found = __shfl(mine, holder);
    }
    if (lane != 0) return;
    int want_low, want_high;
    const bool proven = FREE ? found < INF && certify_ends(a, n, m, found, lo, lo + W - 1, fr, &want_low, &want_high)
                             : certify(a, n, m, found, lo, lo + W - 1, &want_low, &want_high);
    if (!proven) {
        if (FREE && found >= INF) {  // no end in the band: the whole matrix's diagonals next
            want_low = -n;
            want_high = m;
        }
        file_retry(a, slot, n + m, want_low, want_high);
        return;
    }
    if (FREE) {
        // The end cell on that diagonal: column m (at or above the corner's diagonal) or row n (below it).
        a.end_row[slot] = best_diagonal >= end ? m - best_diagonal : n;
        a.end_col[slot] = best_diagonal >= end ? m : n + best_diagonal;
    }
    a.cost[slot] = found;
    a.status[slot] = 0;
    a.walk_low[slot] = lo;
    a.walk_class[slot] = a.cls;
    a.walk_flags[slot] = reinterpret_cast<uint64_t>(flags);
}

// ------------------------------------------------------------------ the walk

// Walks every settled alignment back through its band's decisions, one thread a pair: a light kernel, so many
// waves hide the walk's chain of dependent loads.
__global__ __launch_bounds__(256) void walk_bands(SweepArgs a) {
    const int slot = blockIdx.x * 256 + threadIdx.x;
    // Every lane stays to the end: the CIGARs are placed in the output a wave at a time.
    const bool settled = slot < a.slots && a.status[slot] == 0 && a.cost[slot] <= a.cap;
    int64_t moves = -1;
    if (settled) {
        const int cls = a.walk_class[slot];
        // Free ends: from the end the anchored phase found, back to the view's origin.
        const int n = a.free_ends ? a.end_row[slot] : a.tape.rows[slot];
        const int m = a.free_ends ? a.end_col[slot] : a.tape.cols[slot];
        const int lo = a.walk_low[slot];
        const uint8_t *base = reinterpret_cast<const uint8_t *>(a.walk_flags[slot]);
        if (!band_is_wave(cls)) {
            // A group band: tiles of 32 anti-diagonals, 16 bytes a diagonal pair q = (k - k0) / 2 at
            // (t / 32) * (W / 2) + q, word (t / 8) % 4, nibble t % 8 (group.hip).
            const int half = band_width(cls) / 2;
            const uint4 *tiles = reinterpret_cast<const uint4 *>(base);
            int cached = -1, ahead_at = -1;
            uint4 held = make_uint4(0, 0, 0, 0), ahead = held;
            moves = walk_tiles(a, slot, n, m, [&](int r, int c) -> uint32_t {
                // A proven band holds the path: the guard only keeps a wrong path inside the decisions.
                const int j = c - r - lo;
                if ((unsigned)j >= (unsigned)(2 * half)) return 0u;
                const int t = r + c, tile = t >> 5;
                const int at = tile * half + (j >> 1);
                if (at != cached) {
                    held = at == ahead_at ? ahead : tiles[at];
                    cached = at;
                    // The walk runs back, mostly along its diagonal: fetch the tile before on it ahead of need.
                    ahead_at = tile > 0 ? at - half : -1;
                    if (tile > 0) ahead = tiles[ahead_at];
                }
                const int w = (t >> 3) & 3;
                const uint32_t word = w == 0 ? held.x : (w == 1 ? held.y : (w == 2 ? held.z : held.w));
                return (word >> (4 * (t & 7))) & 15u;
            });
        } else {
            const int W = band_width(cls);
            auto flag_at = [&](int r, int j) -> uint32_t {
                return (unsigned)j < (unsigned)W ? base[(size_t)(r - 1) * W + j] : 0u;
            };
            moves = a.costs.pieces == 2 ? walk_rows2(a, slot, n, m, lo, flag_at) : walk_rows(a, slot, n, m, lo, flag_at);
        }
    }
    publish_cigars(a, slot, moves);
}

// ------------------------------------------------------------------ launches

template <int Q, bool FREE>
static void launch_wave_as(const SweepArgs &a, size_t lds, hipStream_t stream) {
    const dim3 grid(a.count), block(64);
    if (a.costs.pieces == 2) {
        if (a.record) hipLaunchKernelGGL((sweep_wave<Q, true, 2, FREE>), grid, block, lds, stream, a);
        else hipLaunchKernelGGL((sweep_wave<Q, false, 2, FREE>), grid, block, lds, stream, a);
    } else {
        if (a.record) hipLaunchKernelGGL((sweep_wave<Q, true, 1, FREE>), grid, block, lds, stream, a);
        else hipLaunchKernelGGL((sweep_wave<Q, false, 1, FREE>), grid, block, lds, stream, a);
    }
}

template <int Q>
static void launch_wave(const SweepArgs &a, size_t lds, hipStream_t stream) {
    if (a.free_ends) launch_wave_as<Q, true>(a, lds, stream);
    else launch_wave_as<Q, false>(a, lds, stream);
}

int class_width(int cls) { return band_width(cls); }

bool class_is_wave(int cls) { return band_is_wave(cls); }

// Launches one class's sweep over `a.list`; `lds` the letters' bytes a block stages.
void launch_sweep(const SweepArgs &a, size_t lds, hipStream_t stream) {
    if (a.count <= 0) return;
    switch (a.cls) {
        case 0: launch_group(1, a, lds, stream); break;
        case 1: launch_group(2, a, lds, stream); break;
        case 2: launch_group(4, a, lds, stream); break;
        case 3: launch_group(8, a, lds, stream); break;
        case CLASS_WAVE_FIRST + 0: launch_wave<1>(a, lds, stream); break;
        case CLASS_WAVE_FIRST + 1: launch_wave<2>(a, lds, stream); break;
        case CLASS_WAVE_FIRST + 2: launch_wave<4>(a, lds, stream); break;
        case CLASS_WAVE_FIRST + 3: launch_wave<8>(a, lds, stream); break;
        case CLASS_WAVE_FIRST + 4: launch_wave<16>(a, lds, stream); break;
        case CLASS_WAVE_FIRST + 5: launch_wave<32>(a, lds, stream); break;
        default:
            if (band_is_block(a.cls)) launch_block(block_s(a.cls), block_waves(a.cls), a, lds, stream);
            break;
    }
}

// Walks every settled alignment.
void launch_walk(const SweepArgs &a, hipStream_t stream) {
    if (a.slots <= 0) return;
    hipLaunchKernelGGL(walk_bands, dim3((a.slots + 255) / 256), dim3(256), 0, stream, a);
}

}  // namespace dinara
