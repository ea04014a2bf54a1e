// Banded global alignment under gap-affine costs on MI355X (gfx950).
//
// Two sweeps of Gotoh's recurrence over a band of diagonals, row by row, each cell keeping the decision the
// reference's tie rule reads back (reference lanes.mojo:795-947 recurrence, :905-923 flags, :1410-1460 walk):
//
//  - sweep_lane<W>: one thread per pair, its W diagonals (16, 32 or 64) in VGPRs. The letters come from the
//    transposed 2-bit tape, a word per 16 rows, coalesced across the wave; the query's band window is a shift
//    register of packed letters, and one XOR against the row's letter gives every cell's match bit at once.
//  - sweep_wave<Q>: one wave per pair, lane j holding diagonals lo + jQ .. lo + jQ + Q - 1 (a band of 64Q).
//    The deletion layer reads the row above's next diagonal across one lane; the insertion layer, a run
//    along the row, is a prefix minimum over the row's diagonals: in-lane, then across the wave.
//
// Both prove their band by the cost they find (reference lanes.mojo:1337-1392): a band no path off which costs
// as little (for alignments) or less (for distances) settles the pair; any other is filed, with the band its
// cost proves, for a wider launch. A settled alignment is walked back through its decisions on the device and
// its CIGAR written as text, then compacted into one buffer.
//
// Decisions are 4 bits a cell: bits 0-1 the source the aligned layer takes (0 two equal letters aligned,
// 3 two unequal, 1 the deletion layer, 2 the insertion layer), bit 2 the deletion extends into the cell,
// bit 3 the insertion does.

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
__device__ int64_t off_band(const CostParams &c, int64_t k, int64_t end) {
    if (k > (end > 0 ? end : 0)) return inserted_cost(c, k) + deleted_cost(c, k - end);
    if (k < (end < 0 ? end : 0)) return deleted_cost(c, -k) + inserted_cost(c, end - k);
    return 0;
}

// Whether `found`, the least cost over lane diagonals lo..hi, settles the pair (reference lanes.mojo:1360-1392);
// if not, the band every cheaper (strict: no dearer) path could visit, in *want_low..*want_high.
__device__ bool certify(const SweepArgs &a, int rows, int cols, int64_t found, int lo, int hi, int *want_low,
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

// The class of band a pair needing `need` diagonals takes next (the smallest that holds them).
__device__ int class_for(int need, int user_band) {
    if (!user_band) {
        if (need <= 16) return CLASS_LANE16;
        if (need <= 32) return CLASS_LANE32;
        if (need <= 64) return CLASS_LANE64;
    } else if (need <= 64) {
        return CLASS_WAVE_Q1;
    }
    int q = 2, cls = CLASS_WAVE_FIRST;
    while (q < 32 && 64 * q < need) {
        q *= 2;
        ++cls;
    }
    return 64 * q >= need ? cls : CLASS_TOO_WIDE;
}

__device__ void file_retry(const SweepArgs &a, int slot, int letters, int want_low, int want_high) {
    a.want_low[slot] = want_low;
    a.want_high[slot] = want_high;
    const int cls = class_for(want_high - want_low + 1, a.user_band);
    // A wave sweep stages both sequences in LDS: a pair they overflow has no class to go to.
    if (cls == CLASS_TOO_WIDE || (cls >= CLASS_WAVE_FIRST && letters > a.lds_limit)) {
        a.status[slot] = 2;
        return;
    }
    a.status[slot] = 1;
    const int at = atomicAdd(&a.retry_counts[cls], 1);
    a.retry_lists[(size_t)cls * a.retry_capacity + at] = slot;
}

// ------------------------------------------------------------------ CIGAR text

// Runs of one operation, each its length then its letter, written forward from the start of a slot's scratch
// or, for a walk that comes from the end, backward from its end.
struct CigarText {
    char *buf;
    int64_t pos;
    int forward;
    int length;
    char op;

    __device__ void put(char ch) {
        if (forward) buf[pos++] = ch;
        else buf[--pos] = ch;
    }
    __device__ void flush() {
        if (length == 0) return;
        if (forward) {
            int p = 1;
            while (p * 10 <= length) p *= 10;
            for (; p > 0; p /= 10) put((char)('0' + (length / p) % 10));
            put(op);
        } else {
            put(op);
            int v = length;
            do {
                put((char)('0' + v % 10));
                v /= 10;
            } while (v > 0);
        }
        length = 0;
    }
    __device__ void add(char o) {
        if (length > 0 && o != op) flush();
        op = o;
        ++length;
    }
};

// Compacts a slot's finished CIGAR from its scratch into the shared output.
__device__ void publish_cigar(const SweepArgs &a, int slot, const char *start, int64_t length) {
    const unsigned long long at = atomicAdd(a.cigar_used, (unsigned long long)length);
    char *out = a.cigars + at;
    for (int64_t i = 0; i < length; ++i) out[i] = start[i];
    a.cigar_offset[slot] = (int64_t)at;
    a.cigar_length[slot] = (int32_t)length;
}

// The reference's walk (lanes.mojo:1410-1460) over one pair's decisions, `flag_at(r, j)` reading the decision
// of row r's diagonal lo + j; writes the CIGAR into the slot's scratch and publishes it.
template <typename FlagAt>
__device__ void walk_and_publish(const SweepArgs &a, int slot, int n, int m, int lo, FlagAt flag_at) {
    char *scratch = a.cigar_scratch + a.cigar_scratch_base[slot];
    const int64_t capacity = 2 * ((int64_t)n + m) + 2;
    CigarText out{scratch, a.forward_cigar ? 0 : capacity, a.forward_cigar, 0, 0};
    int r = n, c = m, layer = 0;
    while (r > 0 && c > 0) {
        const uint32_t f = flag_at(r, c - r - lo);
        if (layer == 0) {
            const uint32_t src = f & 3;
            if (src == 0 || src == 3) {
                out.add(a.eqx ? (src == 0 ? '=' : 'X') : 'M');
                --r;
                --c;
                continue;
            }
            layer = (int)src;
        }
        const bool extended = layer == 1 ? ((f >> 2) & 1) : ((f >> 3) & 1);
        if (layer == 1) {
            out.add('D');
            --r;
        } else {
            out.add('I');
            --c;
        }
        if (!extended) layer = 0;
    }
    for (; r > 0; --r) out.add('D');
    for (; c > 0; --c) out.add('I');
    out.flush();
    if (a.forward_cigar) publish_cigar(a, slot, scratch, out.pos);
    else publish_cigar(a, slot, scratch + out.pos, capacity - out.pos);
}

// ------------------------------------------------------------------ one thread per pair

template <int BITS>
struct LetterFetch {
    static constexpr int PER = 32 / BITS;
    static constexpr uint32_t MASK = (1u << BITS) - 1;
    const uint32_t *p;
    int cached;
    uint32_t word;
    __device__ uint32_t get(int i) {
        const int w = i / PER;
        if (w != cached) {
            word = p[(size_t)w * 64];
            cached = w;
        }
        return (word >> (BITS * (i % PER))) & MASK;
    }
};

// One cell of the recurrence in cost form with its decision. `hd` is the cell above-left (same diagonal, row
// before), `hu`/`du` the cell above's score and deletion layer (next diagonal, row before), `hl`/`il` the cell
// to the left (this row); `il` becomes this cell's insertion layer.
struct CellOut {
    int h, d;
    uint32_t flag;
};

template <bool RECORD>
__device__ __forceinline__ CellOut cell(const CostParams &k, bool eq, int hd, int hu, int du, int hl, int &il) {
    const int sub = hd + (eq ? 0 : k.mismatch);
    const int od = hu + k.del_first, ed = du + k.del_ext;
    const int dn = imin(od, ed);
    const int oi = hl + k.ins_first, ei = il + k.ins_ext;
    il = imin(oi, ei);
    const int h = imin(sub, imin(dn, il));
    uint32_t flag = 0;
    if (RECORD) {
        // The lowest priority first, each source as cheap as the cell taking over (lanes.mojo:906-913):
        // a match, then an insertion, a deletion, and a substitution above all.
        uint32_t src = eq ? 0u : 3u;
        src = il == h ? 2u : src;
        src = dn == h ? 1u : src;
        src = (!eq && sub == h) ? 3u : src;
        flag = src | (ed <= od ? 4u : 0u) | (ei <= oi ? 8u : 0u);
    }
    return CellOut{h, dn, flag};
}

template <int W, int BITS, bool RECORD>
__global__ __launch_bounds__(64) void sweep_lane(SweepArgs a) {
    const int t = blockIdx.x * 64 + threadIdx.x;
    if (t >= a.count) return;
    const int slot = a.list[t];
    const int n = a.tape.rows[slot], m = a.tape.cols[slot];
    const int need_low = a.want_low[slot], need_high = a.want_high[slot];
    const int lo = need_low - (W - (need_high - need_low + 1)) / 2;
    const int end = m - n;
    const CostParams k = a.costs;

    const int group = slot >> 6, lane = slot & 63;
    const uint32_t *ref_words = a.tape.words + a.tape.group_base[group] + lane;
    const uint32_t *qry_words = ref_words + (size_t)a.tape.group_ref_words[group] * 64;
    LetterFetch<BITS> ref_fetch{ref_words, -1, 0};
    LetterFetch<BITS> qry_fetch{qry_words, -1, 0};

    constexpr int WW = W * BITS / 32;  // words of the query window
    constexpr int PER = 32 / BITS;
    uint32_t win[WW];
#pragma unroll
    for (int i = 0; i < WW; ++i) win[i] = 0;
    // The window for "row 0": letter j is query letter lo - 1 + j; each row shifts one more in at the top.
#pragma unroll
    for (int j = 0; j < W; ++j) {
        const int q = lo - 1 + j;
        const uint32_t letter = (q >= 0 && q < m) ? qry_fetch.get(q) : 0u;
        win[j / PER] |= letter << (BITS * (j % PER));
    }

    int H[W], D[W];
#pragma unroll
    for (int j = 0; j < W; ++j) {
        const int kk = lo + j;
        H[j] = kk < 0 ? INF : (kk == 0 ? 0 : (kk <= m ? k.ins_open + kk * k.ins_ext : INF));
        D[j] = INF;
    }

    // This pair's decisions, W/2 bytes a row, rows one after another.
    uint8_t *flags = nullptr;
    if (RECORD) flags = a.lane_flags + a.lane_base[(size_t)a.cls * a.slots + slot];
    constexpr int WORDS = W / 8;
    uint32_t row_words[WORDS];

    // Rows whose band still holds column zero, or columns before it, take the border.
    const int border_rows = imin(n, -lo);
    for (int r = 1; r <= n; ++r) {
        const uint32_t letter = ref_fetch.get(r - 1);
        {
            const int q = r + lo + W - 2;
            const uint32_t incoming = (q >= 0 && q < m) ? qry_fetch.get(q) : 0u;
#pragma unroll
            for (int i = 0; i < WW; ++i) {
                const uint32_t above = i + 1 < WW ? win[i + 1] : incoming;
                win[i] = (win[i] >> BITS) | (above << (32 - BITS));
            }
        }
        // Every cell's match bit at once.
        uint32_t match[WW];
#pragma unroll
        for (int i = 0; i < WW; ++i) {
            if (BITS == 2) {
                const uint32_t x = win[i] ^ (letter * 0x55555555u);
                match[i] = ~(x | (x >> 1)) & 0x55555555u;
            } else {
                uint32_t bits = 0;
#pragma unroll
                for (int b = 0; b < 4; ++b) bits |= (((win[i] >> (8 * b)) & 0xFFu) == letter ? 1u : 0u) << (8 * b);
                match[i] = bits;
            }
        }
        int hl = INF, il = INF;
        uint32_t word = 0;
        const bool border = r <= border_rows;
        const int rlo = r + lo;
#pragma unroll
        for (int j = 0; j < W; ++j) {
            const bool eq = (match[j / PER] >> (BITS * (j % PER))) & 1u;
            const int hu = j + 1 < W ? H[j + 1] : INF;
            const int du = j + 1 < W ? D[j + 1] : INF;
            CellOut o = cell<RECORD>(k, eq, H[j], hu, du, hl, il);
            if (border) {
                const int c = rlo + j;
                if (c <= 0) {
                    o.h = c == 0 ? k.del_open + r * k.del_ext : INF;
                    o.d = INF;
                    il = INF;
                }
            }
            H[j] = o.h;
            D[j] = o.d;
            hl = o.h;
            if (RECORD) {
                word |= o.flag << (4 * (j % 8));
                if (j % 8 == 7) {
                    row_words[j / 8] = word;
                    word = 0;
                }
            }
        }
        if (RECORD) {
            uint8_t *row = flags + (size_t)(r - 1) * (W / 2);
            if constexpr (WORDS == 2) {
                *reinterpret_cast<uint2 *>(row) = make_uint2(row_words[0], row_words[1]);
            } else {
#pragma unroll
                for (int w = 0; w < WORDS; w += 4)
                    *reinterpret_cast<uint4 *>(row + 4 * w) =
                        make_uint4(row_words[w], row_words[w + 1], row_words[w + 2], row_words[w + 3]);
            }
        }
    }

    int found = INF;
#pragma unroll
    for (int j = 0; j < W; ++j)
        if (j == end - lo) found = H[j];
    int want_low, want_high;
    if (!certify(a, n, m, found, lo, lo + W - 1, &want_low, &want_high)) {
        file_retry(a, slot, n + m, want_low, want_high);
        return;
    }
    a.cost[slot] = found;
    a.status[slot] = 0;
    a.walk_low[slot] = lo;
    a.walk_class[slot] = a.cls;
}

// The walk back for every pair a lane sweep settled, one thread a pair: a light kernel, so many waves hide the
// walk's chain of dependent loads. A pair's decisions lie row after row, so a cache line holds several rows.
__global__ __launch_bounds__(256) void walk_lanes(SweepArgs a) {
    const int slot = blockIdx.x * 256 + threadIdx.x;
    if (slot >= a.slots) return;
    if (a.status[slot] != 0) return;
    const int cls = a.walk_class[slot];
    if (cls > CLASS_LANE64 || a.cost[slot] > a.cap) return;
    const int W = 16 << cls;
    const int row_bytes = W / 2;
    const int n = a.tape.rows[slot], m = a.tape.cols[slot];
    const int lo = a.walk_low[slot];
    const uint8_t *fl = a.lane_flags + a.lane_base[(size_t)cls * a.slots + slot];
    int cached_row = -1, cached_word = -1;
    uint32_t held = 0;
    walk_and_publish(a, slot, n, m, lo, [&](int r, int j) -> uint32_t {
        // A proven band holds the path: the guard only keeps a wrong path inside the decisions.
        if ((unsigned)j >= (unsigned)W) return 0u;
        const int w = j >> 3;
        if (r != cached_row || w != cached_word) {
            held = *reinterpret_cast<const uint32_t *>(fl + (size_t)(r - 1) * row_bytes + 4 * w);
            cached_row = r;
            cached_word = w;
        }
        return (held >> (4 * (j & 7))) & 15u;
    });
}

// ------------------------------------------------------------------ one wave per pair

template <int Q, bool RECORD>
__global__ __launch_bounds__(64) void sweep_wave(SweepArgs a) {
    extern __shared__ uint8_t letters[];
    const int t = blockIdx.x;
    const int slot = a.list[t];
    const int lane = threadIdx.x;
    const int n = a.tape.rows[slot], m = a.tape.cols[slot];
    constexpr int W = 64 * Q;
    const int need_low = a.want_low[slot], need_high = a.want_high[slot];
    const int lo = need_low - (W - (need_high - need_low + 1)) / 2;
    const int end = m - n;
    const CostParams k = a.costs;

    // Both sequences' letters into LDS, a byte each: the reference's first, then the query's.
    {
        const int group = slot >> 6, column = slot & 63;
        const uint32_t *ref_words = a.tape.words + a.tape.group_base[group] + column;
        const uint32_t *qry_words = ref_words + (size_t)a.tape.group_ref_words[group] * 64;
        const int bits = a.tape.bits, per = 32 / bits;
        const uint32_t mask = (1u << bits) - 1;
        for (int i = lane; i < n; i += 64)
            letters[i] = (uint8_t)((ref_words[(size_t)(i / per) * 64] >> (bits * (i % per))) & mask);
        for (int i = lane; i < m; i += 64)
            letters[n + i] = (uint8_t)((qry_words[(size_t)(i / per) * 64] >> (bits * (i % per))) & mask);
    }
    __syncthreads();

    const int first = lo + lane * Q;  // this lane's lowest diagonal
    int H[Q], D[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const int kk = first + i;
        H[i] = kk < 0 ? INF : (kk == 0 ? 0 : (kk <= m ? k.ins_open + kk * k.ins_ext : INF));
        D[i] = INF;
    }
    const bool barred_band = a.user_band != 0;
    const int user_low = a.user_low, user_high = a.user_high;
    uint8_t *flags = nullptr;
    if (RECORD) {
        // The launch's arena holds this pair's decisions, W bytes a row.
        unsigned long long at = 0;
        if (lane == 0) at = atomicAdd(a.arena_used, ((unsigned long long)n * W + 15) & ~15ull);
        at = __shfl(at, 0);
        flags = a.arena + at;
    }

    for (int r = 1; r <= n; ++r) {
        const uint8_t letter = letters[r - 1];
        // The row above's next diagonal, for the lane's last one: the next lane's first.
        int hu_next = __shfl_down(H[0], 1);
        int du_next = __shfl_down(D[0], 1);
        if (lane == 63) hu_next = du_next = INF;
        int T[Q], SUB[Q], DN[Q], X[Q];
        bool EQ[Q];
        bool ED_LE[Q];
        const int border = k.del_open + r * k.del_ext;
        int aggregate = INF;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const int kk = first + i;
            const int c = r + kk;
            const int hu = i + 1 < Q ? H[i + 1] : hu_next;
            const int du = i + 1 < Q ? D[i + 1] : du_next;
            const int od = hu + k.del_first, ed = du + k.del_ext;
            const int dn = imin(od, ed);
            const uint8_t b = (c >= 1 && c <= m) ? letters[n + c - 1] : (uint8_t)0xFF;
            const bool eq = b == letter;
            const int sub = H[i] + (eq ? 0 : k.mismatch);
            int tt = imin(sub, dn);
            if (c <= 0) tt = c == 0 ? border : INF;
            if (barred_band && (kk < user_low || kk > user_high)) tt = INF;
            T[i] = tt;
            SUB[i] = sub;
            DN[i] = (c <= 0 || (barred_band && (kk < user_low || kk > user_high))) ? INF : dn;
            EQ[i] = eq;
            ED_LE[i] = ed <= od;
            // The insertion run from diagonal kk onward: its cost less kk extensions, so the run's
            // extensions cancel and the row's insertions are a prefix minimum (the row's own index drops out).
            X[i] = tt - kk * k.ins_ext;
            aggregate = imin(aggregate, X[i]);
        }
        // Inclusive minimum across the wave, then shifted by one lane: the best run start below this lane.
        int scan = aggregate;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int other = __shfl_up(scan, d);
            if (lane >= d) scan = imin(scan, other);
        }
        int before = __shfl_up(scan, 1);
        if (lane == 0) before = INF + INF / 2;
        int I[Q], HN[Q];
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const int kk = first + i;
            // A gap opened after the best start below: opening plus one extension a letter.
            int ii = before + (k.ins_first - k.ins_ext) + kk * k.ins_ext;
            ii = ii < INF ? ii : INF;
            before = imin(before, X[i]);
            const int c = r + kk;
            if (c <= 0 || (barred_band && (kk < user_low || kk > user_high))) ii = INF;
            I[i] = ii;
            int h = imin(T[i], ii);
            if (barred_band && (kk < user_low || kk > user_high)) h = INF;
            HN[i] = h;
        }
        if (RECORD) {
            int hl_prev = __shfl_up(HN[Q - 1], 1);
            int il_prev = __shfl_up(I[Q - 1], 1);
            if (lane == 0) hl_prev = il_prev = INF;
            uint32_t packed = 0;
#pragma unroll
            for (int i = 0; i < Q; ++i) {
                const int hl = i > 0 ? HN[i - 1] : hl_prev;
                const int il = i > 0 ? I[i - 1] : il_prev;
                const int oi = hl + k.ins_first, ei = il + k.ins_ext;
                const int h = HN[i];
                uint32_t src = EQ[i] ? 0u : 3u;
                src = I[i] == h ? 2u : src;
                src = DN[i] == h ? 1u : src;
                src = (!EQ[i] && SUB[i] == h) ? 3u : src;
                const uint32_t flag = src | (ED_LE[i] ? 4u : 0u) | (ei <= oi ? 8u : 0u);
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
        }
    }

    // The corner's cost, from the lane holding diagonal `end`.
    int mine = INF;
#pragma unroll
    for (int i = 0; i < Q; ++i)
        if (first + i == end) mine = H[i];
    const int holder = (end - lo) / Q;
    const int found = __shfl(mine, holder);
    __threadfence();
    __syncthreads();
    if (lane != 0) return;
    int want_low, want_high;
    if (!certify(a, n, m, found, lo, lo + W - 1, &want_low, &want_high)) {
        file_retry(a, slot, n + m, want_low, want_high);
        return;
    }
    a.cost[slot] = found;
    a.status[slot] = 0;
    a.walk_low[slot] = lo;
    a.walk_class[slot] = a.cls;
    if (RECORD && found <= a.cap) {
        const uint8_t *fl = flags;
        walk_and_publish(a, slot, n, m, lo,
                         [&](int r, int j) -> uint32_t {
                             return (unsigned)j < (unsigned)W ? fl[(size_t)(r - 1) * W + j] : 0u;
                         });
    }
}

// ------------------------------------------------------------------ launches

template <int W, int BITS>
static void launch_lane_bits(const SweepArgs &a, hipStream_t stream) {
    const dim3 grid((a.count + 63) / 64), block(64);
    if (a.record) hipLaunchKernelGGL((sweep_lane<W, BITS, true>), grid, block, 0, stream, a);
    else hipLaunchKernelGGL((sweep_lane<W, BITS, false>), grid, block, 0, stream, a);
}

template <int W>
static void launch_lane(const SweepArgs &a, hipStream_t stream) {
    if (a.tape.bits == 2) launch_lane_bits<W, 2>(a, stream);
    else launch_lane_bits<W, 8>(a, stream);
}

template <int Q>
static void launch_wave(const SweepArgs &a, size_t lds, hipStream_t stream) {
    const dim3 grid(a.count), block(64);
    if (a.record) hipLaunchKernelGGL((sweep_wave<Q, true>), grid, block, lds, stream, a);
    else hipLaunchKernelGGL((sweep_wave<Q, false>), grid, block, lds, stream, a);
}

int class_width(int cls) {
    switch (cls) {
        case CLASS_LANE16: return 16;
        case CLASS_LANE32: return 32;
        case CLASS_LANE64: return 64;
        case CLASS_WAVE_Q1: return 64;
        default: return 64 << (cls - CLASS_WAVE_FIRST + 1);
    }
}

bool class_is_wave(int cls) { return cls >= CLASS_WAVE_FIRST; }

// Launches one class's sweep over `a.list`; `lds` the letters' bytes the wave kernel stages (n + m at most).
void launch_sweep(const SweepArgs &a, size_t lds, hipStream_t stream) {
    if (a.count <= 0) return;
    switch (a.cls) {
        case CLASS_LANE16: launch_lane<16>(a, stream); break;
        case CLASS_LANE32: launch_lane<32>(a, stream); break;
        case CLASS_LANE64: launch_lane<64>(a, stream); break;
        case CLASS_WAVE_Q1: launch_wave<1>(a, lds, stream); break;
        case CLASS_WAVE_FIRST + 0: launch_wave<2>(a, lds, stream); break;
        case CLASS_WAVE_FIRST + 1: launch_wave<4>(a, lds, stream); break;
        case CLASS_WAVE_FIRST + 2: launch_wave<8>(a, lds, stream); break;
        case CLASS_WAVE_FIRST + 3: launch_wave<16>(a, lds, stream); break;
        case CLASS_WAVE_FIRST + 4: launch_wave<32>(a, lds, stream); break;
        default: break;
    }
}

// Walks every pair a lane sweep settled.
void launch_walk(const SweepArgs &a, hipStream_t stream) {
    if (a.slots <= 0) return;
    hipLaunchKernelGGL(walk_lanes, dim3((a.slots + 255) / 256), dim3(256), 0, stream, a);
}

}  // namespace dinara
