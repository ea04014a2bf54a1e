/*
 * A CPU baseline for unit-cost CIGARs: the reference's per-pair method restated, bit-parallel band doubling
 * (reference edit_distance.mojo:274-327 edit_cigar, band.mojo:526-657 band_doubling, bit_parallel.mojo:161-215
 * advance), without its seeds and heuristics (A*PA2's), so it is slower than the reference's own per-pair path.
 * Timed beside the GPU as the reported CPU baseline for the uneven unit-cost batch, and cross-checked against
 * the oracle. Measurement and test infrastructure only; the product never links it.
 *
 * Per pair: a band of W diagonals (a multiple of 64, starting at 64 around the start and end diagonals) swept
 * column by column as a bit-vector window over the rows (vertical differences Pv/Mv, one 64-bit word per 64
 * rows), each column keeping a 2-bit choice per cell for the walk; the band doubles until the strict unit-cost
 * certificate proves it (no path leaving it costs as little). The walk from (n, m) takes a substitution when one
 * is optimal, else a reference letter alone, else a query letter alone, else two equal letters
 * (traceback.mojo:400-575); Ties.LEFT is the walk over both sequences reversed (edit_distance.mojo:295-297).
 * Threads take pairs longest first from a shared counter (api.mojo:1009-1060).
 */
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace {

// The unit-cost price of any path visiting diagonal k (query less reference letters) whose end is on `end`.
inline int64_t off_band(int64_t k, int64_t end) {
    if (k > std::max<int64_t>(0, end)) return 2 * k - end;
    if (k < std::min<int64_t>(0, end)) return end - 2 * k;
    return 0;
}

struct Work {
    std::vector<uint64_t> peq;   // per letter code, the reference's rows as bits, `words` words, offset `shift`
    std::vector<uint64_t> pv, mv, planes;
    std::vector<uint8_t> codes;  // byte -> dense code (255 absent from the reference)
};

// One sweep of band [lo, lo + 64 nw - 1]: the cost at (n, m), and each column's choices (2 x nw words) when
// `choices` is non-null.
int64_t sweep(const uint8_t *x, int64_t n, const uint8_t *y, int64_t m, int64_t lo, int nw, Work &w,
              std::vector<uint64_t> *choices) {
    const int64_t width = 64 * (int64_t)nw, hi = lo + width - 1;
    // Reference rows as bits per letter, row r at bit r + pad (rows above the matrix and below it: no letter).
    const int64_t pad = 64 * ((hi + 64) / 64 + 1);
    const int64_t span = pad + std::max<int64_t>(n, m - lo) + 2 * 64;
    const int64_t words = span / 64 + 2;
    int letters = 0;
    w.codes.assign(256, 255);
    for (int64_t i = 0; i < n; ++i)
        if (w.codes[x[i]] == 255) w.codes[x[i]] = (uint8_t)letters++;
    w.peq.assign((size_t)std::max(letters, 1) * words, 0);
    for (int64_t i = 0; i < n; ++i) {
        const int64_t bit = (i + 1) + pad;
        w.peq[(size_t)w.codes[x[i]] * words + bit / 64] |= uint64_t(1) << (bit % 64);
    }
    w.pv.assign(nw, 0);
    w.mv.assign(nw, 0);
    for (int k = 0; k < nw; ++k)
        for (int b = 0; b < 64; ++b) {
            const int64_t row = -hi + 64 * k + b;  // column 0: cost |row|
            if (row >= 1) w.pv[k] |= uint64_t(1) << b;
            else w.mv[k] |= uint64_t(1) << b;
        }
    if (choices) choices->assign((size_t)2 * nw * m, 0);
    int64_t top = hi;  // the cost at the window's top row
    std::vector<uint64_t> eq(nw), ph(nw), mh(nw);
    for (int64_t j = 1; j <= m; ++j) {
        // Down one row: each word takes the next word's low bit; the bottom takes +1 (Pv) under it.
        for (int k = 0; k < nw; ++k) {
            const uint64_t next_p = k + 1 < nw ? (w.pv[k + 1] & 1) : 1;
            const uint64_t next_m = k + 1 < nw ? (w.mv[k + 1] & 1) : 0;
            w.pv[k] = (w.pv[k] >> 1) | (next_p << 63);
            w.mv[k] = (w.mv[k] >> 1) | (next_m << 63);
        }
        const int64_t r0 = j - hi;  // the window's top row
        const uint8_t code = w.codes[y[j - 1]];
        for (int k = 0; k < nw; ++k) {
            if (code == 255) {
                eq[k] = 0;
                continue;
            }
            const int64_t bit = r0 + 64 * k + pad;
            const uint64_t *p = w.peq.data() + (size_t)code * words;
            const int64_t at = bit / 64, s = bit % 64;
            eq[k] = s ? (p[at] >> s) | (p[at + 1] << (64 - s)) : p[at];
        }
        uint64_t carry = 0, hp_in = 1, hm_in = 0;
        for (int k = 0; k < nw; ++k) {
            const uint64_t pv = w.pv[k], mv = w.mv[k], e = eq[k];
            const uint64_t x1 = (e & pv) + pv;
            const uint64_t c1 = x1 < pv;
            const uint64_t sum = x1 + carry;
            carry = c1 | (carry && x1 == ~uint64_t(0));
            const uint64_t xv = e | mv;
            const uint64_t xh = (sum ^ pv) | e;
            ph[k] = mv | ~(xh | pv);
            mh[k] = pv & xh;
            const uint64_t phs = (ph[k] << 1) | hp_in, mhs = (mh[k] << 1) | hm_in;
            hp_in = ph[k] >> 63;
            hm_in = mh[k] >> 63;
            const uint64_t npv = mhs | ~(xv | phs), nmv = phs & xv;
            if (choices) {
                const uint64_t unequal = ~e & ((npv & ~(phs | mhs)) | (phs & ~(npv | nmv)));
                (*choices)[((size_t)(j - 1) * nw + k) * 2] = ~unequal & (npv | ~ph[k]);
                (*choices)[((size_t)(j - 1) * nw + k) * 2 + 1] = ~unequal & ~npv;
            }
            w.pv[k] = npv;
            w.mv[k] = nmv;
        }
        top += 1 + (int64_t)(w.pv[0] & 1) - (int64_t)(w.mv[0] & 1);
    }
    // Down the last column from the window's top to row n.
    const int64_t last = hi - (m - n);
    int64_t cost = top;
    for (int64_t i = 1; i <= last; ++i) {
        const int k = (int)(i / 64), b = (int)(i % 64);
        cost += (int64_t)((w.pv[k] >> b) & 1) - (int64_t)((w.mv[k] >> b) & 1);
    }
    return cost;
}

std::string walk(const uint8_t *x, int64_t n, const uint8_t *y, int64_t m, int64_t lo, int nw,
                 const std::vector<uint64_t> &choices, bool reversed, bool eqx) {
    const int64_t hi = lo + 64 * (int64_t)nw - 1;
    std::string moves;  // from (n, m) back: 'M' a pair of letters, 'D', 'I'
    int64_t r = n, c = m;
    while (r > 0 && c > 0) {
        const int64_t i = hi - c + r;
        const int k = (int)(i / 64), b = (int)(i % 64);
        const size_t at = ((size_t)(c - 1) * nw + k) * 2;
        const int code = (int)((choices[at] >> b) & 1) | (int)(((choices[at + 1] >> b) & 1) << 1);
        if (code == 1) {
            moves += 'D';
            --r;
        } else if (code == 2) {
            moves += 'I';
            --c;
        } else {
            moves += 'M';
            --r;
            --c;
        }
    }
    moves.append((size_t)r, 'D');
    moves.append((size_t)c, 'I');
    if (!reversed) std::reverse(moves.begin(), moves.end());
    // Runs in the alignment's order; with reversed sequences the walk's order is already the alignment's.
    std::string out;
    int64_t i = 0, j = 0;  // positions in the original (unreversed) sequences
    char run = 0;
    int64_t length = 0;
    auto flush = [&]() {
        if (length) out += std::to_string(length) + run;
    };
    for (char mv : moves) {
        char op = mv;
        if (mv == 'M') {
            const uint8_t a = reversed ? x[n - 1 - i] : x[i], bb = reversed ? y[m - 1 - j] : y[j];
            op = eqx ? (a == bb ? '=' : 'X') : 'M';
            ++i;
            ++j;
        } else if (mv == 'D') {
            ++i;
        } else {
            ++j;
        }
        if (op == run) {
            ++length;
        } else {
            flush();
            run = op;
            length = 1;
        }
    }
    flush();
    return out;
}

}  // namespace

extern "C" {

/* Unit-cost distances (and CIGARs when `cigars` is non-null, Ties.LEFT unless `right_ties`) of `pairs` pairs on
 * `threads` threads: costs[i], and CIGAR i NUL-terminated at cigars + i * cigar_stride (or empty when it would not
 * fit). Returns the pairs done. */
int64_t baseline_edit(int64_t pairs, const char *const *references, const int64_t *reference_lengths,
                      const char *const *queries, const int64_t *query_lengths, int64_t threads, int64_t right_ties,
                      int64_t *costs, char *cigars, int64_t cigar_stride) {
    std::vector<int64_t> order(pairs);
    for (int64_t i = 0; i < pairs; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
        return std::max(reference_lengths[a], query_lengths[a]) > std::max(reference_lengths[b], query_lengths[b]);
    });
    std::atomic<int64_t> next{0}, done{0};
    auto worker = [&]() {
        Work w;
        std::vector<uint64_t> choices;
        std::vector<uint8_t> xr, yr;
        for (int64_t t; (t = next.fetch_add(1)) < pairs;) {
            const int64_t p = order[t];
            const int64_t n = reference_lengths[p], m = query_lengths[p];
            const uint8_t *x = reinterpret_cast<const uint8_t *>(references[p]);
            const uint8_t *y = reinterpret_cast<const uint8_t *>(queries[p]);
            const bool reversed = cigars && !right_ties;
            if (reversed) {
                xr.assign(x, x + n);
                yr.assign(y, y + m);
                std::reverse(xr.begin(), xr.end());
                std::reverse(yr.begin(), yr.end());
                x = xr.data();
                y = yr.data();
            }
            const int64_t end = m - n;
            const int64_t want_lo = std::min<int64_t>(0, end) - 1, want_hi = std::max<int64_t>(0, end) + 1;
            int nw = 1;
            while (64 * (int64_t)nw < want_hi - want_lo + 1) nw *= 2;
            int64_t lo = want_lo - (64 * nw - (want_hi - want_lo + 1)) / 2, cost = 0;
            while (true) {
                cost = sweep(x, n, y, m, lo, nw, w, cigars ? &choices : nullptr);
                const int64_t hi = lo + 64 * (int64_t)nw - 1;
                const int64_t far = INT64_MAX / 4;
                const int64_t above = hi < m ? off_band(hi + 1, end) : far;
                const int64_t below = lo > -n ? off_band(lo - 1, end) : far;
                const int64_t off = std::min(above, below);
                if (cigars ? cost < off : cost <= off) break;
                lo -= 32 * (int64_t)nw;  // twice as wide, centred on the band just swept
                nw *= 2;
            }
            costs[p] = cost;
            if (cigars) {
                const std::string text = walk(x, n, y, m, lo, nw, choices, reversed, true);
                char *out = cigars + p * cigar_stride;
                if ((int64_t)text.size() < cigar_stride) std::memcpy(out, text.c_str(), text.size() + 1);
                else out[0] = 0;
            }
            done.fetch_add(1);
        }
    };
    std::vector<std::thread> pool;
    for (int64_t i = 1; i < std::max<int64_t>(threads, 1); ++i) pool.emplace_back(worker);
    worker();
    for (auto &t : pool) t.join();
    return done.load();
}

}  // extern "C"
