// The whole-matrix path for pairs no band class takes, on MI355X (gfx950): Gotoh's recurrence in cost form over
// every cell (reference lanes.mojo:795-947, the whole matrix as the reference's CPU path fills it when no band is
// asked for), each cell's decision kept under the reference's flag rule (lanes.mojo:905-923), then the walk back
// (lanes.mojo:1410-1460).
//
// Layout. The query's columns are cut into strips of 512 (FULL_STRIP); a wave sweeps a strip top to bottom, lane j
// holding columns c0 + 1 .. c0 + 8 (c0 = strip * 512 + 8j) one row behind lane j - 1: at step s lane j takes row
// s - j. A cell's left neighbours (H and the insertion layers) come from the lane to the left, made the step before,
// by one DPP wave_shr:1; the cells above are the lane's own registers. A strip's lane 0 takes its left neighbours
// from the last column of the strip to its left, which that strip's lane 63 writes to a carry record a row, with
// a relaxed agent-scope atomic store: the record's own nonzero value says it is there, so no fence or flag is
// needed, and agent scope keeps it coherent across the XCDs' L2s. The reading wave fetches 64 rows' records at
// once (lane i row s + i) every 64 steps, and lane 0 takes row s's with a readlane. Strips run as a pipeline, each
// about 130 rows behind the one to its left. One launch takes every listed pair, their strips numbered pair after
// pair; waves claim strips in order from a counter, so the strip a wave waits on is always held by a running wave,
// and every wait gives up after FULL_POLL_LIMIT polls.
//
// The reference's letters pass through a ring of four 256-letter chunks in LDS, as the Scoring sweep's long
// references do; the lane's eight query letters sit in registers.
//
// Decisions are those the wave sweep keeps (sweep.hip), so the walk is its rule: one piece, 4 bits a cell (bits
// 0-1 the source: 0 two equal letters, 3 two unequal, 1 the deletion layer, 2 the insertion layer; bit 2 the
// deletion extends into the cell, bit 3 the insertion does), eight cells a 32-bit lane record; two pieces, a byte
// a cell (bits 0-2 the source, 0 equal, 5 unequal, 1-4 the gap layers; bit 2 + L layer L extends), eight a 64-bit
// record. A pair's record (strip, r, lane) is at its first record + ((strip * rows) + r - 1) * 64 + lane.

#include "bands.h"
#include "full.h"

namespace dinara {

namespace {

constexpr int RING_CHUNK = 256;  // letters a chunk: four a lane
constexpr int RING_CHUNKS = 4;
constexpr int RING_MASK = RING_CHUNK * RING_CHUNKS - 1;

__device__ __forceinline__ int from_left_lane(int old, int v) {
    return __builtin_amdgcn_update_dpp(old, v, 0x138, 0xF, 0xF, false);
}

// Letter x of a sequence whose first tape word is at `words` (the slot's column of its group).
__device__ __forceinline__ uint32_t tape_letter(const uint32_t *words, int bits, int per, uint32_t mask, int x) {
    return (words[(size_t)(x / per) * 64] >> (bits * (x % per))) & mask;
}

// A gap of `letters` letters at its cheaper piece (none: nothing).
template <int P>
__device__ __forceinline__ int gap_of(int open, int ext, int open2, int ext2, int letters) {
    if (letters <= 0) return 0;
    const int one = open + letters * ext;
    return P == 2 ? min(one, open2 + letters * ext2) : one;
}

__device__ __forceinline__ unsigned long long carry_record(int h, int i) {
    return ((unsigned long long)(uint32_t)(h + 1) << 32) | (uint32_t)(i + 1);
}

template <int P>
__global__ __launch_bounds__(64) void full_sweep(FullArgs a) {
    __shared__ uint32_t ring_words[RING_CHUNK * RING_CHUNKS / 4];
    const uint8_t *ring = reinterpret_cast<const uint8_t *>(ring_words);
    const int lane = threadIdx.x;
    const CostParams k = a.s.costs;
    const int bits = a.s.tape.bits, per = 32 / bits;
    const uint32_t mask = (1u << bits) - 1;
    const int total = a.strip_first[a.pairs];

    for (;;) {
        int claimed = 0;
        if (lane == 0) claimed = atomicAdd(a.claim, 1);
        claimed = __builtin_amdgcn_readfirstlane(__shfl(claimed, 0));
        if (claimed >= total) return;
        // The claimed strip's pair: the last whose first strip is at or before it.
        int lo = 0, hi = a.pairs - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (a.strip_first[mid] <= claimed) lo = mid; else hi = mid - 1;
        }
        const int pair = lo;
        const int slot = a.pair_slot[pair];
        const int strip = claimed - a.strip_first[pair];
        const int strips = a.strip_first[pair + 1] - a.strip_first[pair];
        const int rows = a.s.tape.rows[slot], cols = a.s.tape.cols[slot];
        if (rows == 0 || cols == 0) {  // one gap of every letter; the walk makes it
            if (lane == 0) {
                a.s.cost[slot] = rows == 0 ? gap_of<P>(k.ins_open, k.ins_ext, k.ins_open2, k.ins_ext2, cols)
                                           : gap_of<P>(k.del_open, k.del_ext, k.del_open2, k.del_ext2, rows);
                a.s.status[slot] = 0;
            }
            continue;
        }
        const int group = slot >> 6, column = slot & 63;
        const uint32_t *ref_words = a.s.tape.words + a.s.tape.group_base[group] + column;
        const uint32_t *qry_words = ref_words + (size_t)a.s.tape.group_ref_words[group] * 64;
        // Four reference letters from `at` on, a byte each (zero past the end).
        auto ref4 = [&](int at) {
            uint32_t w = 0;
#pragma unroll
            for (int t = 0; t < 4; ++t)
                if (at + t < rows) w |= tape_letter(ref_words, bits, per, mask, at + t) << (8 * t);
            return w;
        };
        const int c0 = strip * FULL_STRIP + lane * FULL_C;  // the column left of the lane's first
        uint32_t q[FULL_C];
#pragma unroll
        for (int i = 0; i < FULL_C; ++i) {
            const int col = c0 + i + 1;
            q[i] = col <= cols ? tape_letter(qry_words, bits, per, mask, col - 1) : 0xFFFFu;
        }
        int H[FULL_C], D[FULL_C], D2[FULL_C];  // the row above, the lane's columns
#pragma unroll
        for (int i = 0; i < FULL_C; ++i) {
            H[i] = gap_of<P>(k.ins_open, k.ins_ext, k.ins_open2, k.ins_ext2, c0 + i + 1);
            D[i] = D2[i] = INF;
        }
        int left_prev = gap_of<P>(k.ins_open, k.ins_ext, k.ins_open2, k.ins_ext2, c0);  // H(r - 1, c0)
        int out_h = 0, out_i = INF, out_i2 = INF;  // the lane's last column, the row it just made
        unsigned long long *carry = a.carry + a.carry_at[pair];
        const unsigned long long *left_carry = strip > 0 ? carry + (size_t)(strip - 1) * (rows + 1) * P : nullptr;
        unsigned long long *own_carry = strip + 1 < strips ? carry + (size_t)strip * (rows + 1) * P : nullptr;
        unsigned long long fetched = 0, fetched2 = 0;  // lane i: the left strip's record of row (s - 1 & ~63) + 1 + i

        // The ring's first two chunks, the third held in a register (the previous strip's reads are done).
        __syncthreads();
        ring_words[lane] = ref4(4 * lane);
        ring_words[64 + lane] = ref4(RING_CHUNK + 4 * lane);
        uint32_t pending = ref4(2 * RING_CHUNK + 4 * lane);
        __syncthreads();

        bool gave_up = false;
        const int steps = rows + 63;
        for (int s = 1; s <= steps; ++s) {
            // Steps 256c + 1 .. 256c + 256 read letters 256c - 63 .. 256c + 255: chunk c + 1 goes in at the range's
            // start, from the register loaded a range before.
            if ((s & (RING_CHUNK - 1)) == 1 && s > 1) {
                const int c = s / RING_CHUNK;
                ring_words[((c + 1) & (RING_CHUNKS - 1)) * 64 + lane] = pending;
                pending = ref4((c + 2) * RING_CHUNK + 4 * lane);
                __syncthreads();
            }
            if (left_carry && (s & 63) == 1 && s <= rows) {
                const int row = s + lane;
                const bool want = row <= rows;
                int polls = 0;
                for (;;) {
                    if (want) {
                        fetched = __hip_atomic_load(left_carry + (size_t)row * P, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
                        if (P == 2)
                            fetched2 = __hip_atomic_load(left_carry + (size_t)row * P + 1, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT);
                    }
                    const bool missing = want && (fetched == 0 || (P == 2 && fetched2 == 0));
                    if (__ballot(missing) == 0) break;
                    if (++polls > FULL_POLL_LIMIT) {
                        gave_up = true;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);
                }
                if (gave_up) break;
            }
            const int r = s - lane;
            int lh = from_left_lane(0, out_h), li = from_left_lane(INF, out_i);
            int li2 = P == 2 ? from_left_lane(INF, out_i2) : INF;
            if (left_carry) {
                // Row s of the left strip's last column, for lane 0 (every lane reads it; lane 0 keeps it).
                const int from = (s - 1) & 63;
                const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(fetched >> 32), from);
                const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)fetched, from);
                if (lane == 0) {
                    lh = (int)hi - 1;
                    li = (int)lo - 1;
                    if (P == 2) li2 = (int)__builtin_amdgcn_readlane((uint32_t)fetched2, from) - 1;
                }
            } else if (lane == 0) {
                lh = gap_of<P>(k.del_open, k.del_ext, k.del_open2, k.del_ext2, r);  // column 0
                li = li2 = INF;
            }
            if (r >= 1 && r <= rows) {
                const uint32_t letter = ring[(r - 1) & RING_MASK];
                int hl = lh, il = li, il2 = li2, hd = left_prev;
                uint32_t record = 0;
                unsigned long long record2 = 0;
#pragma unroll
                for (int i = 0; i < FULL_C; ++i) {
                    const int hu = H[i], du = D[i];
                    const int od = hu + k.del_first, ed = du + k.del_ext;
                    const int d = min(od, ed);
                    const int oi = hl + k.ins_first, ei = il + k.ins_ext;
                    const int ins = min(oi, ei);
                    const bool eq = q[i] == letter;
                    const int sub = hd + (eq ? 0 : k.mismatch);
                    int h = min(sub, min(d, ins));
                    int d2 = INF, ins2 = INF;
                    bool ed2_le = false, ei2_le = false;
                    if (P == 2) {
                        const int od2 = hu + k.del_first2, ed2 = D2[i] + k.del_ext2;
                        const int oi2 = hl + k.ins_first2, ei2 = il2 + k.ins_ext2;
                        d2 = min(od2, ed2);
                        ins2 = min(oi2, ei2);
                        ed2_le = ed2 <= od2;
                        ei2_le = ei2 <= oi2;
                        h = min(h, min(d2, ins2));
                    }
                    if (a.record) {
                        // The lowest priority first, each later source as cheap as the cell taking over.
                        if (P == 1) {
                            uint32_t src = eq ? 0u : 3u;
                            src = ins == h ? 2u : src;
                            src = d == h ? 1u : src;
                            src = (!eq && sub == h) ? 3u : src;
                            record |= (src | (ed <= od ? 4u : 0u) | (ei <= oi ? 8u : 0u)) << (4 * i);
                        } else {
                            uint32_t src = eq ? 0u : 5u;
                            src = ins == h ? 2u : src;
                            src = ins2 == h ? 4u : src;
                            src = d == h ? 1u : src;
                            src = d2 == h ? 3u : src;
                            src = (!eq && sub == h) ? 5u : src;
                            const uint32_t flag = src | (ed <= od ? 8u : 0u) | (ei <= oi ? 16u : 0u) |
                                                  (ed2_le ? 32u : 0u) | (ei2_le ? 64u : 0u);
                            record2 |= (unsigned long long)flag << (8 * i);
                        }
                    }
                    hd = hu;
                    hl = h;
                    il = ins;
                    il2 = ins2;
                    H[i] = h;
                    D[i] = d;
                    D2[i] = d2;
                }
                out_h = hl;
                out_i = il;
                out_i2 = il2;
                left_prev = lh;
                const size_t at = (size_t)a.decision_at[pair] + ((size_t)strip * rows + (r - 1)) * 64 + lane;
                if (a.record) {
                    if (P == 1)
                        static_cast<uint32_t *>(a.decisions)[at] = record;
                    else
                        static_cast<unsigned long long *>(a.decisions)[at] = record2;
                }
                if (own_carry && lane == 63) {
                    __hip_atomic_store(own_carry + (size_t)r * P, carry_record(out_h, out_i), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                    if (P == 2)
                        __hip_atomic_store(own_carry + (size_t)r * P + 1, carry_record(0, out_i2), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                }
                if (r == rows) {
#pragma unroll
                    for (int i = 0; i < FULL_C; ++i)
                        if (c0 + i + 1 == cols) {
                            a.s.cost[slot] = H[i];
                            a.s.status[slot] = 0;
                        }
                }
            }
        }
        if (gave_up) {
            if (lane == 0) atomicExch(a.failed + pair, 1);
            return;
        }
    }
}

// The walk, by a whole wave: the walk's own rule (walk_rows / walk_rows2 in bands.h) made one move at a time by
// every lane alike, while the lanes load the decisions of the next 64 cells along the line the walk is on (the
// diagonal in the aligned layer, up the column in a deletion layer, along the row in an insertion layer), so a round
// costs one load's latency for up to 64 moves. A move off that line ends the round. Lane 0 keeps the moves.
template <int P>
__global__ __launch_bounds__(64) void full_walk(FullArgs a) {
    const int lane = threadIdx.x;
    const SweepArgs &s = a.s;
    const int pair = blockIdx.x;
    const int slot = a.pair_slot[pair];
    const int rows = s.tape.rows[slot], cols = s.tape.cols[slot];
    const int64_t first = a.decision_at[pair];
    int64_t moves = -1;
    if (a.record && a.failed[pair] == 0 && s.status[slot] == 0 && s.cost[slot] <= s.cap) {
        OpTape ops = op_tape(s, slot);
        const uint32_t unequal = s.eqx ? 1u : 0u;
        int r = rows, c = cols, layer = 0;
        while (r > 0 && c > 0) {
            const int dr = (layer == 0 || (layer & 1)) ? 1 : 0;
            const int dc = (layer == 0 || !(layer & 1)) ? 1 : 0;
            const int lr = r - lane * dr, lc = c - lane * dc;
            uint32_t flag = 0;
            if (lr > 0 && lc > 0) {
                const int strip = (lc - 1) / FULL_STRIP, within = (lc - 1) % FULL_STRIP;
                const size_t at = (size_t)first + ((size_t)strip * rows + (lr - 1)) * 64 + within / FULL_C;
                const int cell = within % FULL_C;
                if (P == 1)
                    flag = (static_cast<const uint32_t *>(a.decisions)[at] >> (4 * cell)) & 15u;
                else
                    flag = (uint32_t)(static_cast<const unsigned long long *>(a.decisions)[at] >> (8 * cell)) & 255u;
            }
            for (int step = 0; step < 64 && r > 0 && c > 0; ++step) {
                const uint32_t f = __builtin_amdgcn_readlane(flag, step);
                int mr = 1, mc = 1;
                uint32_t code = 0;
                bool aligned = false;
                if (layer == 0) {
                    const uint32_t src = P == 1 ? (f & 3u) : (f & 7u);
                    aligned = src == 0 || src == (P == 1 ? 3u : 5u);
                    if (aligned)
                        code = src == 0 ? 0u : unequal;
                    else
                        layer = (int)src;
                }
                if (!aligned) {
                    const uint32_t extended = (f >> (P == 1 ? layer + 1 : layer + 2)) & 1u;
                    if (layer & 1) {
                        code = 2u;
                        mc = 0;
                    } else {
                        code = 3u;
                        mr = 0;
                    }
                    if (!extended) layer = 0;
                }
                if (lane == 0) ops.put(code);
                r -= mr;
                c -= mc;
                if (mr != dr || mc != dc) break;  // off the loaded line
            }
        }
        if (lane == 0) {
            for (; r > 0; --r) ops.put(2u);
            for (; c > 0; --c) ops.put(3u);
            ops.close();
            moves = ops.count;
        }
    }
    publish_cigars(s, slot, moves);
}

__global__ __launch_bounds__(256) void full_collect(const int32_t *status, int slots, int32_t *list, int32_t *count) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t < slots && status[t] != 0) list[atomicAdd(count, 1)] = t;
}

}  // namespace

void launch_full_sweep(const FullArgs &a, int strips, hipStream_t stream) {
    if (a.pairs <= 0) return;
    const int grid = strips < FULL_GRID ? strips : FULL_GRID;
    if (a.s.costs.pieces == 2)
        hipLaunchKernelGGL(full_sweep<2>, dim3(grid), dim3(64), 0, stream, a);
    else
        hipLaunchKernelGGL(full_sweep<1>, dim3(grid), dim3(64), 0, stream, a);
}

void launch_full_walk(const FullArgs &a, hipStream_t stream) {
    if (a.pairs <= 0) return;
    if (a.s.costs.pieces == 2)
        hipLaunchKernelGGL(full_walk<2>, dim3(a.pairs), dim3(64), 0, stream, a);
    else
        hipLaunchKernelGGL(full_walk<1>, dim3(a.pairs), dim3(64), 0, stream, a);
}

void launch_full_collect(const int32_t *status, int slots, int32_t *list, int32_t *count, hipStream_t stream) {
    if (slots <= 0) return;
    hipLaunchKernelGGL(full_collect, dim3((slots + 255) / 256), dim3(256), 0, stream, status, slots, list, count);
}

}  // namespace dinara
