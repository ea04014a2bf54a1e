/*
 * The CPU baseline (see baseline.h): the reference's lanes batch path restated in C++ with GCC vector
 * extensions, 32 pairs side by side in 16-bit lanes. Measurement and test infrastructure only.
 *
 * Follows reference lanes.mojo: dealt (:1007-1106, pairs grouped by end diagonal), the first band over the
 * group's start and end diagonals and one more either side (:1640-1649), band_costs/swept with flags
 * (:543-947), proof (:1337-1392), the second pass over the band each unproven pair's cost needs (:1695-1760),
 * walked (:1410-1460) over both sequences reversed for Ties.LEFT (:1523-1525).
 */
#include "baseline.h"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <thread>
#include <vector>

namespace {

constexpr int LANES = 32;
typedef int16_t V __attribute__((vector_size(LANES * 2)));
typedef int8_t F __attribute__((vector_size(LANES)));
constexpr int16_t FAR = 16384;  // lanes.mojo:107-110
constexpr int64_t HELD = 16000; // lanes.mojo:122-124

inline V splat(int v) { return V{} + (int16_t)v; }
inline V vmin(V a, V b) { return a < b ? a : b; }

struct Costs {
    int64_t mismatch, open, extend;
    int64_t deleted(int64_t k) const { return k <= 0 ? 0 : open + k * extend; }
    int64_t inserted(int64_t k) const { return k <= 0 ? 0 : open + k * extend; }
    int64_t off_band(int64_t k, int64_t end) const {  // lanes.mojo:333-345
        if (k > std::max<int64_t>(0, end)) return inserted(k) + deleted(k - end);
        if (k < std::min<int64_t>(0, end)) return deleted(-k) + inserted(end - k);
        return 0;
    }
    bool fits(int64_t rows, int64_t columns) const {  // lanes.mojo:317-331
        int64_t dearest = open + rows * extend + open + columns * extend;
        return dearest + std::max(mismatch, open + extend) < HELD;
    }
};

struct Batch {
    const char *const *refs;
    const int64_t *rlens;
    const char *const *qrys;
    const int64_t *qlens;
    Costs costs;
    int64_t *costs_out;
    char *cigars;
    int64_t stride;
};

struct Space {
    std::vector<V> scores, dels, rowl, coll;
    std::vector<F> flags;
};

struct Member {
    int64_t index;
    int64_t first_cost;
};

// One group's sweep over lane diagonals low..high (query less reference), both sequences reversed, every flag
// kept: the lanes' band_costs + swept[record] (lanes.mojo:543-947).
void sweep(const Batch &b, Space &sp, const Member *members, int count, int64_t low, int64_t high, int64_t *found) {
    int64_t rows = 0, columns = 0;
    int64_t n[LANES], m[LANES];
    for (int l = 0; l < LANES; ++l) {
        n[l] = l < count ? b.rlens[members[l].index] : -1;
        m[l] = l < count ? b.qlens[members[l].index] : 0;
        rows = std::max(rows, n[l]);
        columns = std::max(columns, m[l]);
    }
    sp.rowl.assign(std::max<int64_t>(rows, 1), V{});
    sp.coll.assign(std::max<int64_t>(columns, 1), V{});
    for (int l = 0; l < count; ++l) {
        const char *r = b.refs[members[l].index];
        const char *q = b.qrys[members[l].index];
        for (int64_t i = 0; i < n[l]; ++i) sp.rowl[i][l] = (unsigned char)r[n[l] - 1 - i];
        for (int64_t i = 0; i < m[l]; ++i) sp.coll[i][l] = (unsigned char)q[m[l] - 1 - i];
    }
    const int64_t span = high - low + 1;
    sp.scores.assign(columns + 2, splat(FAR));
    sp.dels.assign(columns + 2, splat(FAR));
    sp.flags.resize(std::max<int64_t>(rows * span, 1));
    for (int64_t c = std::max<int64_t>(low, 0); c <= std::min(high, columns); ++c)
        sp.scores[c] = splat((int)std::min<int64_t>(b.costs.inserted(c), FAR));
    for (int l = 0; l < LANES; ++l) {
        found[l] = FAR;
        if (l < count && n[l] == 0 && m[l] >= low && m[l] <= high) found[l] = sp.scores[m[l]][l];
    }
    const V mis = splat((int)b.costs.mismatch), first_gap = splat((int)(b.costs.open + b.costs.extend)),
            ext = splat((int)b.costs.extend), zero = V{};
    for (int64_t r = 1; r <= rows; ++r) {
        int64_t first = std::max<int64_t>(r + low, 0), last = std::min(r + high, columns);
        if (first > last) continue;
        const V letter = sp.rowl[r - 1];
        V left = splat(FAR), diag;
        if (first == 0) {
            diag = sp.scores[0];
            left = splat((int)std::min<int64_t>(b.costs.deleted(r), FAR));
            sp.scores[0] = left;
            first = 1;
        } else {
            diag = sp.scores[first - 1];
            sp.scores[first - 1] = splat(FAR);
            sp.dels[first - 1] = splat(FAR);
        }
        V ins = splat(FAR);
        F *row_flags = sp.flags.data() + (r - 1) * span - r - low;
        for (int64_t c = first; c <= last; ++c) {
            const V above = sp.scores[c];
            const V eq = letter == sp.coll[c - 1];
            const V sub = diag + (eq ? zero : mis);
            const V od = above + first_gap, ed = sp.dels[c] + ext;
            const V d = vmin(od, ed);
            const V oi = left + first_gap, ei = ins + ext;
            ins = vmin(oi, ei);
            const V s = vmin(vmin(sub, d), ins);
            // The lowest priority first, the substitution of two unequal letters above all (lanes.mojo:906-913).
            V flag = ins == s ? splat(2) : zero;
            flag = d == s ? splat(1) : flag;
            flag = (~eq & (sub == s)) ? zero : flag;
            flag |= ed <= od ? splat(4) : zero;
            flag |= ei <= oi ? splat(8) : zero;
            row_flags[c] = __builtin_convertvector(flag, F);
            sp.dels[c] = d;
            sp.scores[c] = s;
            diag = above;
            left = s;
        }
        for (int l = 0; l < count; ++l)
            if (n[l] == r && m[l] >= std::max<int64_t>(r + low, 0) && m[l] <= last) found[l] = sp.scores[m[l]][l];
    }
}

// lanes.mojo:1410-1460 over the kept flags; the moves of both reversed sequences, from their corner, run left to
// right over the pair itself: the CIGAR is written as they come.
void walk(const Batch &b, const Space &sp, int lane, int64_t index, int64_t low, int64_t high) {
    const int64_t n = b.rlens[index], m = b.qlens[index], span = high - low + 1;
    const char *ref = b.refs[index], *qry = b.qrys[index];
    char *out = b.cigars + index * b.stride;
    int64_t used = 0, run = 0, ri = 0, qi = 0;
    char op = 0;
    auto flush = [&]() {
        if (!run) return;
        char digits[24];
        int k = 0;
        for (int64_t v = run; v > 0; v /= 10) digits[k++] = (char)('0' + v % 10);
        while (k) out[used++] = digits[--k];
        out[used++] = op;
        run = 0;
    };
    auto add = [&](char o) {
        if (run && o != op) flush();
        op = o;
        ++run;
    };
    auto aligned = [&]() {
        add(ref[ri] == qry[qi] ? '=' : 'X');
        ++ri;
        ++qi;
    };
    int64_t r = n, c = m;
    int layer = 0;
    while (r > 0 && c > 0) {
        const int flag = (uint8_t)sp.flags[(r - 1) * span + (c - r - low)][lane];
        if (layer == 0) {
            layer = flag & 3;
            if (layer == 0) {
                aligned();
                --r;
                --c;
            }
            continue;
        }
        const bool extended = (flag >> (layer == 1 ? 2 : 3)) & 1;
        if (layer == 1) {
            add('D');
            ++ri;
            --r;
        } else {
            add('I');
            ++qi;
            --c;
        }
        if (!extended) layer = 0;
    }
    for (; r > 0; --r, ++ri) add('D');
    for (; c > 0; --c, ++qi) add('I');
    flush();
    out[used] = 0;
}

struct Retry {
    int64_t index, found, low, high;
};

void run_group(const Batch &b, Space &sp, const Member *members, int count, int64_t low, int64_t high,
               bool second, std::vector<Retry> *retries) {
    int64_t found[LANES];
    sweep(b, sp, members, count, low, high, found);
    for (int l = 0; l < count; ++l) {
        const int64_t index = members[l].index;
        const int64_t n = b.rlens[index], m = b.qlens[index];
        int64_t cost = found[l];
        if (second) {
            cost = std::min(cost, members[l].first_cost);
        } else {
            // proof[strict] (lanes.mojo:1337-1392).
            const int64_t end = m - n;
            const int64_t far = INT64_MAX / 4;
            const int64_t bh = high < m ? b.costs.off_band(high + 1, end) : far;
            const int64_t bl = low > -n ? b.costs.off_band(low - 1, end) : far;
            if (cost >= FAR || !(cost < std::min(bh, bl))) {
                const int64_t target = std::min<int64_t>(cost, FAR) + 1;
                int64_t wh = high, wl = low;
                while (wh < m && b.costs.off_band(wh + 1, end) < target) ++wh;
                while (wl > -n && b.costs.off_band(wl - 1, end) < target) --wl;
                retries->push_back({index, cost, wl, wh});
                continue;
            }
        }
        b.costs_out[index] = cost;
        if (b.cigars) walk(b, sp, l, index, low, high);
    }
}

}  // namespace

extern "C" int64_t baseline_alignments(int64_t pairs, const char *const *references,
                                       const int64_t *reference_lengths, const char *const *queries,
                                       const int64_t *query_lengths, int64_t mismatch, int64_t open, int64_t extend,
                                       int64_t threads, int64_t *costs, char *cigars, int64_t cigar_stride) {
    Batch b{references, reference_lengths, queries, query_lengths, {mismatch, open, extend}, costs, cigars,
            cigar_stride};
    // dealt: every pair the lanes hold, by end diagonal (lanes.mojo:1007-1106).
    std::vector<int64_t> order;
    for (int64_t i = 0; i < pairs; ++i) {
        costs[i] = -1;
        if (reference_lengths[i] > 0 && query_lengths[i] > 0 && b.costs.fits(reference_lengths[i], query_lengths[i]))
            order.push_back(i);
    }
    std::stable_sort(order.begin(), order.end(), [&](int64_t x, int64_t y) {
        return query_lengths[x] - reference_lengths[x] < query_lengths[y] - reference_lengths[y];
    });
    const int64_t groups = ((int64_t)order.size() + LANES - 1) / LANES;
    const int workers = (int)std::max<int64_t>(1, std::min<int64_t>(threads, std::max<int64_t>(groups, 1)));
    std::vector<std::vector<Retry>> retries(workers);
    std::atomic<int64_t> next{0};
    auto first_pass = [&](int w) {
        Space sp;
        Member members[LANES];
        for (int64_t g; (g = next.fetch_add(1)) < groups;) {
            int count = 0;
            int64_t low = 0, high = 0;
            for (int64_t s = g * LANES; s < std::min<int64_t>((int64_t)order.size(), (g + 1) * LANES); ++s) {
                const int64_t i = order[s];
                const int64_t end = query_lengths[i] - reference_lengths[i];
                low = std::min(low, std::min<int64_t>(0, end) - 1);
                high = std::max(high, std::max<int64_t>(0, end) + 1);
                members[count++] = {i, 0};
            }
            run_group(b, sp, members, count, low, high, false, &retries[w]);
        }
    };
    std::vector<std::thread> pool;
    for (int w = 0; w < workers; ++w) pool.emplace_back(first_pass, w);
    for (auto &t : pool) t.join();
    // The second pass: the unproven pairs, narrowest band first, each group over its widest (lanes.mojo:1109-1126).
    std::vector<Retry> all;
    for (auto &r : retries) all.insert(all.end(), r.begin(), r.end());
    std::stable_sort(all.begin(), all.end(),
                     [](const Retry &x, const Retry &y) { return x.high - x.low < y.high - y.low; });
    const int64_t second_groups = ((int64_t)all.size() + LANES - 1) / LANES;
    next = 0;
    pool.clear();
    auto second_pass = [&](int) {
        Space sp;
        Member members[LANES];
        std::vector<Retry> unused;
        for (int64_t g; (g = next.fetch_add(1)) < second_groups;) {
            int count = 0;
            int64_t low = 0, high = 0;
            for (int64_t s = g * LANES; s < std::min<int64_t>((int64_t)all.size(), (g + 1) * LANES); ++s) {
                low = std::min(low, all[s].low);
                high = std::max(high, all[s].high);
                members[count++] = {all[s].index, all[s].found};
            }
            run_group(b, sp, members, count, low, high, true, &unused);
        }
    };
    for (int w = 0; w < workers; ++w) pool.emplace_back(second_pass, w);
    for (auto &t : pool) t.join();
    int64_t settled = 0;
    for (int64_t i = 0; i < pairs; ++i) settled += costs[i] >= 0;
    return settled;
}
