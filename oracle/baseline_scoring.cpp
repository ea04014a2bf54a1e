/*
 * The CPU baseline for local alignment under a substitution table: a C++ restatement of the reference's host local
 * path (scoring.mojo:544-560 local_by_span; scored.mojo:81-84 the end; vector_score.mojo:37-162 the reach back;
 * gotoh.mojo:265-442 the span's global alignment and walk; scoring.mojo:724-745 spans by rfind). It is timed beside
 * the GPU as the reported CPU baseline and cross-checked against the oracle. Test and measurement infrastructure
 * only; the product never links it.
 *
 * Phase 1, the best end, sweeps 32 pairs side by side in 16-bit lanes (the reference sweeps one pair at a time in
 * 16-bit lanes; the cells and the rule are the same), each column keeping its best score and the last row holding
 * it, folded by the end rule afterwards. A group whose scores could pass 16 bits is swept in 32-bit lanes. The reach
 * back and the span's alignment run per pair in 32-bit scalars. Threads take groups from an atomic counter.
 */
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <string>
#include <thread>
#include <vector>

namespace {

constexpr int LANES = 32;

template <typename T>
struct Vec {
    typedef T type __attribute__((vector_size(sizeof(T) * LANES)));
};

struct Table {
    const int8_t *table;  // [a * 32 + b]
    int open, extend;     // a gap's first letter, each later one
};

// Phase 1 over a group of up to 32 pairs: per lane the best score (above zero, else 0), its row and column.
template <typename T>
void local_ends(const Table &s, int lanes, const uint8_t *const *ref, const int *rows, const uint8_t *const *qry,
                const int *cols, int *best, int *end_r, int *end_c) {
    typedef typename Vec<T>::type V;
    int n = 0, m = 0;
    for (int l = 0; l < lanes; ++l) n = std::max(n, rows[l]), m = std::max(m, cols[l]);
    const T neg = (T)(sizeof(T) == 2 ? -16000 : -(1 << 28));
    V zero = {}, NEG = zero + neg, OPEN = zero + (T)s.open, EXT = zero + (T)s.extend;
    std::vector<V> H(m + 1, zero), D(m + 1, NEG), colbest(m + 1, zero), colrow(m + 1, zero);
    std::vector<int32_t> idx((size_t)(m + 1) * LANES, 0);
    for (int c = 1; c <= m; ++c)
        for (int l = 0; l < lanes; ++l) idx[(size_t)c * LANES + l] = c <= cols[l] ? qry[l][c - 1] : 0;
    std::vector<V> in_cols(m + 1);
    for (int c = 1; c <= m; ++c)
        for (int l = 0; l < LANES; ++l) in_cols[c][l] = (T)(l < lanes && c <= cols[l] ? 1 : 0);
    int32_t base[LANES];
    for (int r = 1; r <= n; ++r) {
        for (int l = 0; l < LANES; ++l) base[l] = l < lanes && r <= rows[l] ? ref[l][r - 1] * 32 : 0;
        V in_row;
        for (int l = 0; l < LANES; ++l) in_row[l] = (T)(l < lanes && r <= rows[l] ? 1 : 0);
        V diag = zero, left = zero, ileft = NEG;
        const V rv = zero + (T)r;
        for (int c = 1; c <= m; ++c) {
            V sub;
            const int32_t *q = &idx[(size_t)c * LANES];
            for (int l = 0; l < LANES; ++l) sub[l] = (T)s.table[base[l] + q[l]];
            const V up = H[c];
            const V opened = up + OPEN, extended = D[c] + EXT;
            const V d = opened > extended ? opened : extended;
            const V iopened = left + OPEN, iextended = ileft + EXT;
            const V ins = iopened > iextended ? iopened : iextended;
            V h = diag + sub;
            h = h > d ? h : d;
            h = h > ins ? h : ins;
            h = h > zero ? h : zero;
            const V live = in_row & in_cols[c];
            const auto better = (h >= colbest[c]) & (h > zero) & (live != zero);
            colbest[c] = better ? h : colbest[c];
            colrow[c] = better ? rv : colrow[c];
            diag = up;
            left = h;
            ileft = ins;
            H[c] = h;
            D[c] = d;
        }
    }
    for (int l = 0; l < lanes; ++l) {
        int bh = 0, br = 0, bc = 0;
        for (int c = 1; c <= cols[l]; ++c) {
            const int h = (int)colbest[c][l], r = (int)colrow[c][l];
            if (h <= 0) continue;
            if (h > bh || (h == bh && (r + c > br + bc || (r + c == br + bc && r > br)))) bh = h, br = r, bc = c;
        }
        best[l] = bh;
        end_r[l] = br;
        end_c[l] = bc;
    }
}

int gap_run(const Table &s, int letters) { return letters == 0 ? 0 : s.open + (letters - 1) * s.extend; }

// The reach back: over both prefixes reversed from the end, the global score anchored there; the first cell by
// anti-diagonal, then by row, that reaches `target` (rows in order; a row past the earliest anti-diagonal found
// ends it). Writes the start.
void reach_back(const Table &s, const uint8_t *ref, const uint8_t *qry, int er, int ec, int target, int *sr,
                int *sc) {
    const int neg = -(1 << 28);
    std::vector<int> H(ec + 1), D(ec + 1, neg);
    for (int c = 0; c <= ec; ++c) H[c] = gap_run(s, c);
    int best_d = 1 << 30, best_r = 0, best_c = 0;
    for (int r = 1; r <= er && r + 1 <= best_d; ++r) {
        const int8_t *row = s.table + ref[er - r] * 32;
        int diag = H[0], left = gap_run(s, r), ileft = neg;
        H[0] = left;
        for (int c = 1; c <= ec; ++c) {
            const int up = H[c];
            const int d = std::max(up + s.open, D[c] + s.extend);
            const int ins = std::max(left + s.open, ileft + s.extend);
            const int h = std::max(diag + row[qry[ec - c]], std::max(d, ins));
            if (h >= target && (r + c < best_d || (r + c == best_d && r < best_r))) best_d = r + c, best_r = r, best_c = c;
            diag = up;
            left = h;
            ileft = ins;
            H[c] = h;
            D[c] = d;
        }
    }
    *sr = best_r ? er - best_r : 0;
    *sc = best_r ? ec - best_c : 0;
}

// The span's global alignment and Gotoh's walk back (decide/advance): the score, the CIGAR.
int span_alignment(const Table &s, const uint8_t *ref, int n, const uint8_t *qry, int m, const char *ref_bytes,
                   const char *qry_bytes, bool eqx, std::string &cigar) {
    const int neg = -(1 << 28);
    std::vector<int> H(m + 1), D(m + 1, neg);
    std::vector<uint8_t> dec((size_t)(n + 1) * (m + 1), 0);
    for (int c = 0; c <= m; ++c) H[c] = gap_run(s, c);
    for (int r = 1; r <= n; ++r) {
        const int8_t *row = s.table + ref[r - 1] * 32;
        int diag = H[0], left = gap_run(s, r), ileft = neg;
        H[0] = left;
        uint8_t *out = &dec[(size_t)r * (m + 1)];
        for (int c = 1; c <= m; ++c) {
            const int up = H[c];
            const int d = std::max(up + s.open, D[c] + s.extend);
            const int ins = std::max(left + s.open, ileft + s.extend);
            const int sub = diag + row[qry[c - 1]];
            const int h = std::max(sub, std::max(d, ins));
            out[c] = (uint8_t)((h == sub ? 0 : (h == d ? 1 : 2)) | (D[c] + s.extend > up + s.open ? 4 : 0) |
                               (ileft + s.extend > left + s.open ? 8 : 0));
            diag = up;
            left = h;
            ileft = ins;
            H[c] = h;
            D[c] = d;
        }
    }
    const int score = H[m];
    std::string moves;
    int r = n, c = m, layer = 0;
    while (r > 0 && c > 0) {
        const uint8_t f = dec[(size_t)r * (m + 1) + c];
        const int state = layer == 0 ? (f & 3) : layer;
        if (state == 1) {
            moves.push_back('D');
            --r;
            layer = (f & 4) ? 1 : 0;
        } else if (state == 2) {
            moves.push_back('I');
            --c;
            layer = (f & 8) ? 2 : 0;
        } else {
            moves.push_back(eqx ? (ref_bytes[r - 1] == qry_bytes[c - 1] ? '=' : 'X') : 'M');
            --r;
            --c;
            layer = 0;
        }
    }
    for (; r > 0; --r) moves.push_back('D');
    for (; c > 0; --c) moves.push_back('I');
    cigar.clear();
    for (size_t i = moves.size(); i > 0;) {
        const char op = moves[i - 1];
        size_t j = i;
        while (j > 0 && moves[j - 1] == op) --j;
        cigar += std::to_string(i - j);
        cigar.push_back(op);
        i = j;
    }
    return score;
}

int64_t rfind(const char *hay, int64_t n, const char *needle, int64_t k) {
    if (k == 0) return 0;
    for (int64_t at = n - k; at >= 0; --at)
        if (memcmp(hay + at, needle, (size_t)k) == 0) return at;
    return -1;
}

}  // namespace

/* Local alignments of `pairs` pairs whose letters are codes (`ref_codes`, `qry_codes`) into a 32-column table, the
 * original bytes in `references`/`queries` (for `=`/`X` and the spans), on `threads` threads. Writes each pair's
 * score, spans (4 a pair) and CIGAR (NUL-terminated at cigars + i * cigar_stride when `cigars` is non-null). */
extern "C" int64_t baseline_local_alignments(int64_t pairs, const char *const *references, const int64_t *ref_lens,
                                             const char *const *queries, const int64_t *qry_lens,
                                             const uint8_t *const *ref_codes, const uint8_t *const *qry_codes,
                                             const int8_t *table32, int64_t open, int64_t extend, int64_t eqx,
                                             int64_t threads, int64_t *scores, int64_t *spans, char *cigars,
                                             int64_t cigar_stride) {
    Table s{table32, (int)open, (int)extend};
    int best_entry = -128, least_entry = 127;
    for (int i = 0; i < 32 * 32; ++i) best_entry = std::max<int>(best_entry, table32[i]), least_entry = std::min<int>(least_entry, table32[i]);
    std::vector<int64_t> order(pairs);
    for (int64_t i = 0; i < pairs; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
        return ref_lens[a] != ref_lens[b] ? ref_lens[a] > ref_lens[b] : qry_lens[a] > qry_lens[b];
    });
    const int64_t groups = (pairs + LANES - 1) / LANES;
    std::atomic<int64_t> next{0};
    auto worker = [&]() {
        std::string cigar;
        for (;;) {
            const int64_t g = next.fetch_add(1);
            if (g >= groups) break;
            const int64_t first = g * LANES;
            const int lanes = (int)std::min<int64_t>(LANES, pairs - first);
            const uint8_t *rp[LANES], *qp[LANES];
            int rows[LANES], cols[LANES], best[LANES], er[LANES], ec[LANES];
            int64_t longest = 0;
            for (int l = 0; l < lanes; ++l) {
                const int64_t p = order[first + l];
                rp[l] = ref_codes[p];
                qp[l] = qry_codes[p];
                rows[l] = (int)ref_lens[p];
                cols[l] = (int)qry_lens[p];
                longest = std::max<int64_t>(longest, std::min(ref_lens[p], qry_lens[p]));
            }
            // 16-bit lanes when every score and gap run of the group stays inside them.
            const int64_t reach = (longest + 2) * std::max<int64_t>(std::max(best_entry, -least_entry), -extend) - open;
            if (2 * reach < 15000) local_ends<int16_t>(s, lanes, rp, rows, qp, cols, best, er, ec);
            else local_ends<int32_t>(s, lanes, rp, rows, qp, cols, best, er, ec);
            for (int l = 0; l < lanes; ++l) {
                const int64_t p = order[first + l];
                int64_t *sp = spans + 4 * p;
                sp[0] = sp[1] = sp[2] = sp[3] = 0;
                scores[p] = 0;
                cigar.clear();
                if (best[l] > 0) {
                    int sr, sc;
                    reach_back(s, rp[l], qp[l], er[l], ec[l], best[l], &sr, &sc);
                    const int rn = er[l] - sr, qm = ec[l] - sc;
                    scores[p] = span_alignment(s, rp[l] + sr, rn, qp[l] + sc, qm, references[p] + sr,
                                               queries[p] + sc, eqx != 0, cigar);
                    const int64_t rs = rfind(references[p], ref_lens[p], references[p] + sr, rn);
                    const int64_t qs = rfind(queries[p], qry_lens[p], queries[p] + sc, qm);
                    sp[0] = rs;
                    sp[1] = rs + rn;
                    sp[2] = qs;
                    sp[3] = qs + qm;
                }
                if (cigars) {
                    const size_t k = std::min<size_t>(cigar.size(), (size_t)cigar_stride - 1);
                    memcpy(cigars + p * cigar_stride, cigar.data(), k);
                    cigars[p * cigar_stride + k] = 0;
                }
            }
        }
    };
    const int64_t workers = std::max<int64_t>(1, std::min<int64_t>(threads, groups));
    std::vector<std::thread> pool;
    for (int64_t w = 1; w < workers; ++w) pool.emplace_back(worker);
    worker();
    for (auto &t : pool) t.join();
    return pairs;
}
