// The C ABI's substitution-table calls: scores and alignments under a `dinara_scoring`, global or local, on the
// device. Reference behaviour followed: scoring.mojo:126-238 (the table, gaps and 32-bit reach), common.mojo:267-289
// (letters to codes), api.mojo:843-1001 (the batch calls and their modes), scoring.mojo:544-560 (a local
// alignment: its end, its start, the letters between aligned globally) and :724-745 (spans by rfind).
//
// A batch is validated and packed once, then its passes run on the device: scores take one sweep; a global
// alignment a sweep that keeps every decision and a walk back; a local alignment the best end, the reach back to
// its start, and the span's global alignment with its walk.

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/dinara.h"
#include "device.h"
#include "host.h"
#include "score.h"

using namespace dinara;

namespace {

constexpr uint8_t UNKNOWN_LETTER = 0xFF;

// A device step that failed: DINARA_OUT_OF_MEMORY, and with DINARA_DEBUG set in the environment a line on stderr
// naming the step and the runtime's last error.
int64_t failed(const char *step) {
    if (std::getenv("DINARA_DEBUG")) {
        const hipError_t e = hipGetLastError();
        std::fprintf(stderr, "dinara: %s failed (%s)\n", step, hipGetErrorString(e));
    }
    return DINARA_OUT_OF_MEMORY;
}
constexpr int64_t INT32_TOP = 0x7fffffff;

// A Scoring checked as the reference checks one (scoring.mojo:200-214, gotoh.mojo:47-54): 1 to 32 letters, no
// '-', a table, a gap's parts zero or less and within 32 bits. `code` is nonzero when the batch fails whole.
struct ParsedScoring {
    int64_t code = 0;
    int letters = 0;
    uint8_t codes[256];
    int8_t table[32 * 32];
    int open = 0, extend = 0;  // a gap's first letter (opening + extension), each later one
    int64_t reach_opening = 0, reach_step = 0;
};

ParsedScoring parse_scoring(const dinara_scoring *s) {
    ParsedScoring p;
    std::memset(p.codes, UNKNOWN_LETTER, sizeof p.codes);
    std::memset(p.table, 0, sizeof p.table);
    if (!s || !s->alphabet || !s->table || s->letters < 1 || s->letters > 32) {
        p.code = DINARA_INVALID_COSTS;
        return p;
    }
    p.letters = (int)s->letters;
    for (int i = 0; i < p.letters; ++i) {
        const uint8_t byte = (uint8_t)s->alphabet[i];
        if (byte == '-') {
            p.code = DINARA_INVALID_COSTS;
            return p;
        }
        p.codes[byte] = (uint8_t)i;  // a repeated letter takes its last place
    }
    if (s->opening > 0 || s->extension > 0 || std::min(s->opening, s->extension) < -INT32_TOP ||
        s->opening + s->extension < -INT32_TOP) {
        p.code = DINARA_INVALID_COSTS;
        return p;
    }
    int best = -128, least = 127;
    for (int a = 0; a < p.letters; ++a)
        for (int b = 0; b < p.letters; ++b) {
            const int8_t v = s->table[a * p.letters + b];
            p.table[a * 32 + b] = v;
            best = std::max<int>(best, v);
            least = std::min<int>(least, v);
        }
    p.open = (int)(s->opening + s->extension);
    p.extend = (int)s->extension;
    p.reach_opening = 2 * -(int64_t)p.open;
    p.reach_step = std::max<int64_t>(std::max<int64_t>(best, -least), -(int64_t)p.extend);
    return p;
}

// Whether a pair's scores stay well inside 32 bits (reference scoring.mojo:216-238, ScoreReach.check).
bool within_reach(const ParsedScoring &p, int64_t rows, int64_t cols, bool floored) {
    const int64_t letters = floored ? std::min(rows, cols) : rows + cols;
    return p.reach_opening + 2 * (letters + 2) * p.reach_step < (int64_t(1) << 28);
}

// Global or local (reference api.mojo:853-866): a mode's own match score must be zero; other modes have no device
// path here.
int64_t check_scoring_mode(const dinara_mode *m, bool *local) {
    *local = false;
    if (!m) return 0;
    if (m->match_score != 0) return DINARA_INVALID_MODE;
    if (m->kind == DINARA_LOCAL) {
        *local = true;
        return 0;
    }
    if (m->kind != DINARA_ENDS_FREE) return DINARA_INVALID_MODE;
    if ((m->reference_start | m->reference_end | m->query_start | m->query_end) != 0) return DINARA_INVALID_MODE;
    return 0;
}

int64_t gap_run(const ParsedScoring &p, int64_t letters) {
    return letters == 0 ? 0 : p.open + (letters - 1) * (int64_t)p.extend;
}

// Columns a lane: a strip of at most 512 columns, the strips as even as C allows.
int lane_columns(int cols) {
    const int strips = std::max(1, (cols + SCORE_STRIP_MAX - 1) / SCORE_STRIP_MAX);
    const int per_strip = (cols + strips - 1) / strips;
    return std::min(SCORE_MAX_C, std::max(1, (per_strip + 63) / 64));
}

// The last place `needle` occurs in `hay` (reference scoring.mojo:734-735, String.rfind over bytes).
int64_t rfind(const char *hay, int64_t n, const char *needle, int64_t k) {
    if (k == 0) return 0;
    for (int64_t at = n - k; at >= 0; --at)
        if (std::memcmp(hay + at, needle, (size_t)k) == 0) return at;
    return -1;
}

}  // namespace

// ------------------------------------------------------------------ the batch

struct dinara_scoring_batch {
    int gpu = 0;
    int64_t kind = DINARA_BATCH_DISTANCES;  // DISTANCES: scores; ALIGNMENTS: alignments
    bool local = false;
    int eqx = 1;
    int64_t pairs = 0;
    int64_t threads = 0;
    ParsedScoring scoring;
    DeviceContext *ctx = nullptr;
    const char *const *refs = nullptr;  // the caller's letters, for spans (valid during the call that made it)
    const char *const *qrys = nullptr;
    std::vector<int64_t> code;          // per pair: 0 on the device, else its own result code or a settled host value
    std::vector<int64_t> rows, cols;
    std::vector<int32_t> slot_pair;     // slot -> pair, the pairs the device takes, longest first
    std::vector<int64_t> host_score;    // per pair: a score settled without the device (an empty side)
    std::vector<int32_t> view_h;        // per slot: reference begin, letters, query begin, letters
    std::vector<int32_t> score_h, end_row_h, end_col_h, reach_row_h, reach_col_h, cigar_length_h;
    std::vector<int64_t> cigar_offset_h;
    std::vector<char> cigars_h;
    int64_t max_rows = 0, max_cols = 0;
    // device memory
    DeviceBuffer codes, ref_at, qry_at, view, list, table, score, end_row, end_col, target, walk_base, arena, carry,
        op_scratch, op_base, cigars, cigar_used, cigar_offset, cigar_length;
    // timing of the last run
    double run_ms = 0, pass_ms[4] = {0, 0, 0, 0}, walk_ms = 0;
    int launches = 0;
    bool ran = false;
};

namespace {

using Batch = dinara_scoring_batch;

ScoreArgs score_args(Batch *b, int mode, int reversed) {
    ScoreArgs a{};
    a.codes = static_cast<const uint8_t *>(b->codes.ptr);
    a.ref_at = static_cast<const int64_t *>(b->ref_at.ptr);
    a.qry_at = static_cast<const int64_t *>(b->qry_at.ptr);
    a.view = static_cast<const int32_t *>(b->view.ptr);
    a.reversed = reversed;
    a.table = static_cast<const int8_t *>(b->table.ptr);
    a.letters = b->scoring.letters;
    a.open = b->scoring.open;
    a.extend = b->scoring.extend;
    a.mode = mode;
    a.score = static_cast<int32_t *>(b->score.ptr);
    a.end_row = static_cast<int32_t *>(b->end_row.ptr);
    a.end_col = static_cast<int32_t *>(b->end_col.ptr);
    a.target = static_cast<const int32_t *>(b->target.ptr);
    a.walk_base = static_cast<uint64_t *>(b->walk_base.ptr);
    a.eqx = b->eqx;
    a.op_scratch = static_cast<char *>(b->op_scratch.ptr);
    a.op_base = static_cast<const int64_t *>(b->op_base.ptr);
    a.cigars = static_cast<char *>(b->cigars.ptr);
    a.cigar_used = static_cast<unsigned long long *>(b->cigar_used.ptr);
    a.cigar_offset = static_cast<int64_t *>(b->cigar_offset.ptr);
    a.cigar_length = static_cast<int32_t *>(b->cigar_length.ptr);
    a.slots = (int)b->slot_pair.size();
    return a;
}

// One pass over the slots whose views are not empty: grouped by the columns a lane takes, each group one launch (or
// more, when a trace's decisions pass the budget), each trace launch followed by its walk.
// The reference letters a block can stage whole in LDS beside the profile; a longer reference goes through the
// ring.
int64_t lds_room(const Batch *b) {
    const size_t profile = (size_t)b->scoring.letters * 64 * 8;
    return b->ctx->lds_limit > profile ? (int64_t)(b->ctx->lds_limit - profile) : 0;
}

// The table's best score (at least 0).
int top_score(const Batch *b) {
    int top = 0;
    for (int x = 0; x < b->scoring.letters; ++x)
        for (int y = 0; y < b->scoring.letters; ++y) top = std::max<int>(top, b->scoring.table[x * 32 + y]);
    return top;
}

int64_t run_pass(Batch *b, int mode, int reversed, const std::vector<int32_t> &slots, hipEvent_t start,
                 hipEvent_t stop, double *pass_ms, double *walk_ms) {
    DeviceContext *ctx = b->ctx;
    // Groups by the columns a lane takes and by whether the reference fits LDS whole (group g = c + ring * G).
    constexpr int G = SCORE_MAX_C + 1;
    const int64_t room = lds_room(b);
    std::vector<int32_t> by_c[2 * G];
    int64_t rows_max[2 * G] = {0}, cols_max[2 * G] = {0};
    for (int32_t s : slots) {
        const int32_t rows = b->view_h[4 * s + 1], cols = b->view_h[4 * s + 3];
        if (rows <= 0 || cols <= 0) continue;
        const int g = lane_columns(cols) + (rows > room ? G : 0);
        by_c[g].push_back(s);
        rows_max[g] = std::max<int64_t>(rows_max[g], rows);
        cols_max[g] = std::max<int64_t>(cols_max[g], cols);
    }
    std::vector<int32_t> all;
    std::vector<size_t> at(2 * G, 0);
    for (int g = 0; g < 2 * G; ++g) {
        at[g] = all.size();
        all.insert(all.end(), by_c[g].begin(), by_c[g].end());
    }
    if (all.empty()) return 0;
    if (!b->list.upload(ctx, all.data(), all.size() * 4)) return failed("list upload");
    const int32_t *list = static_cast<const int32_t *>(b->list.ptr);
    ScoreArgs a = score_args(b, mode, reversed);
    hipEventRecord(start, ctx->stream);
    for (int g = 0; g < 2 * G; ++g) {
        const int64_t count = (int64_t)by_c[g].size();
        if (count == 0) continue;
        const int c = g % G;
        a.c = c;
        a.ring = g >= G;
        const int strip_cols = 64 * c;
        const int64_t strips = (cols_max[g] + strip_cols - 1) / strip_cols;
        const size_t lds = score_lds(b->scoring.letters, (int)rows_max[g], a.ring != 0);
        int64_t chunk = count;
        if (mode == SCORE_TRACE) {
            a.arena_stride = strips * (rows_max[g] + 1) * 64;
            chunk = std::max<int64_t>(1, std::min<int64_t>(count, (int64_t)(ARENA_BUDGET / 4 / a.arena_stride)));
        }
        a.carry_stride = strips > 1 ? 2 * (rows_max[g] + 1) : 0;
        // A local end's key: the rows' bits below a score no greater than the table's best times the shorter side.
        a.end_bits = 0;
        if (mode == SCORE_LOCAL) {
            int bits = 1;
            while (bits < 31 && (int64_t(1) << bits) <= rows_max[g]) ++bits;
            const int64_t best = (int64_t)top_score(b) * std::min(rows_max[g], cols_max[g]);
            if (best < (int64_t(1) << (32 - bits))) a.end_bits = bits;
        }
        for (int64_t first = 0; first < count; first += chunk) {
            a.list = list + at[g] + first;
            a.count = (int)std::min(chunk, count - first);
            a.carry = nullptr;
            if (strips > 1) {
                if (!b->carry.reserve((size_t)a.count * a.carry_stride * 4)) return failed("carry");
                a.carry = static_cast<int32_t *>(b->carry.ptr);
            }
            if (mode == SCORE_TRACE) {
                if (!b->arena.reserve((size_t)a.count * a.arena_stride * 4)) return failed("arena");
                a.arena = static_cast<uint32_t *>(b->arena.ptr);
            }
            launch_score(a, lds, ctx->stream);
            if (hipPeekAtLastError() != hipSuccess) return failed("score launch");
            b->launches += 1;
            if (mode == SCORE_TRACE) {
                launch_score_walk(a, ctx->stream);
                if (hipPeekAtLastError() != hipSuccess) return failed("walk launch");
            }
        }
    }
    hipEventRecord(stop, ctx->stream);
    if (hipEventSynchronize(stop) != hipSuccess) return failed("pass");
    float ms = 0;
    hipEventElapsedTime(&ms, start, stop);
    *pass_ms += ms;
    (void)walk_ms;
    return 0;
}

int64_t prepare(Batch *b, const char *const *refs, const int64_t *ref_lens, const char *const *qrys,
                const int64_t *qry_lens) {
    DeviceContext *ctx = b->ctx;
    const int64_t pairs = b->pairs;
    b->code.assign(pairs, 0);
    b->rows.assign(pairs, 0);
    b->cols.assign(pairs, 0);
    b->host_score.assign(pairs, 0);
    parallel_for(pairs, b->threads, [&](int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; ++i) {
            const int64_t n = ref_lens[i], m = qry_lens[i];
            if (n < 0 || m < 0) {
                b->code[i] = DINARA_INVALID_LENGTH;
                continue;
            }
            b->rows[i] = n;
            b->cols[i] = m;
            bool known = true;
            for (int64_t k = 0; k < n && known; ++k) known = b->scoring.codes[(uint8_t)refs[i][k]] != UNKNOWN_LETTER;
            for (int64_t k = 0; k < m && known; ++k) known = b->scoring.codes[(uint8_t)qrys[i][k]] != UNKNOWN_LETTER;
            if (!known) {
                b->code[i] = DINARA_UNSUPPORTED_SYMBOLS;
                continue;
            }
            if (!within_reach(b->scoring, n, m, b->local) || n >= (int64_t(1) << 20) || m >= (int64_t(1) << 20)) {
                b->code[i] = DINARA_INVALID_COSTS;
                continue;
            }
            // Against an empty side: one gap of every letter, or no local alignment (reference
            // scoring.mojo:384-396).
            if (n == 0 || m == 0) b->code[i] = 1, b->host_score[i] = b->local ? 0 : gap_run(b->scoring, n + m);
        }
    });
    b->slot_pair.clear();
    for (int64_t i = 0; i < pairs; ++i)
        if (b->code[i] == 0) b->slot_pair.push_back((int32_t)i);
    std::stable_sort(b->slot_pair.begin(), b->slot_pair.end(), [&](int32_t x, int32_t y) {
        return b->rows[x] * b->cols[x] > b->rows[y] * b->cols[y];
    });
    const int64_t slots = (int64_t)b->slot_pair.size();
    std::vector<int64_t> ref_at(std::max<int64_t>(slots, 1)), qry_at(std::max<int64_t>(slots, 1)), op_base(slots + 1);
    int64_t total = 0, ops = 0, text = 0;
    b->max_rows = b->max_cols = 0;
    for (int64_t s = 0; s < slots; ++s) {
        const int32_t p = b->slot_pair[s];
        ref_at[s] = total;
        total += b->rows[p];
        qry_at[s] = total;
        total += b->cols[p];
        op_base[s] = ops;
        ops += 8 * ((b->rows[p] + b->cols[p]) / 32 + 1);
        text += 2 * (b->rows[p] + b->cols[p]) + 2;
        b->max_rows = std::max(b->max_rows, b->rows[p]);
        b->max_cols = std::max(b->max_cols, b->cols[p]);
    }
    op_base[slots] = ops;
    std::vector<uint8_t> codes((size_t)std::max<int64_t>(total, 1));
    parallel_for(slots, b->threads, [&](int64_t lo, int64_t hi) {
        for (int64_t s = lo; s < hi; ++s) {
            const int32_t p = b->slot_pair[s];
            for (int64_t k = 0; k < b->rows[p]; ++k) codes[ref_at[s] + k] = b->scoring.codes[(uint8_t)refs[p][k]];
            for (int64_t k = 0; k < b->cols[p]; ++k) codes[qry_at[s] + k] = b->scoring.codes[(uint8_t)qrys[p][k]];
        }
    });
    b->view_h.assign(4 * std::max<int64_t>(slots, 1), 0);
    const size_t n4 = (size_t)std::max<int64_t>(slots, 1) * 4;
    if (!b->codes.upload(ctx, codes.data(), codes.size()) ||
        !b->ref_at.upload(ctx, ref_at.data(), ref_at.size() * 8) ||
        !b->qry_at.upload(ctx, qry_at.data(), qry_at.size() * 8) ||
        !b->table.upload(ctx, b->scoring.table, sizeof b->scoring.table) || !b->view.reserve(4 * n4) ||
        !b->score.reserve(n4) || !b->end_row.reserve(n4) || !b->end_col.reserve(n4) || !b->target.reserve(n4))
        return DINARA_OUT_OF_MEMORY;
    if (b->kind == DINARA_BATCH_ALIGNMENTS) {
        if (!b->walk_base.reserve(n4 * 2) || !b->op_base.upload(ctx, op_base.data(), op_base.size() * 8) ||
            !b->op_scratch.reserve(std::max<int64_t>(ops, 8)) || !b->cigars.reserve(std::max<int64_t>(text, 1)) ||
            !b->cigar_used.reserve(8) || !b->cigar_offset.reserve(n4 * 2) || !b->cigar_length.reserve(n4))
            return DINARA_OUT_OF_MEMORY;
    }
    return 0;
}

// Sets every slot's view: both sequences whole, forward.
void whole_views(Batch *b) {
    for (size_t s = 0; s < b->slot_pair.size(); ++s) {
        const int32_t p = b->slot_pair[s];
        int32_t *v = &b->view_h[4 * s];
        v[0] = 0;
        v[1] = (int32_t)b->rows[p];
        v[2] = 0;
        v[3] = (int32_t)b->cols[p];
    }
}

bool read_slots(Batch *b, DeviceBuffer &buffer, std::vector<int32_t> &out) {
    out.resize(std::max<size_t>(b->slot_pair.size(), 1));
    return buffer.read(b->ctx, out.data(), out.size() * 4);
}

int64_t run(Batch *b) {
    DeviceContext *ctx = b->ctx;
    std::lock_guard<std::mutex> serial(ctx->lock);
    DeviceGuard guard(ctx);
    b->launches = 0;
    b->run_ms = b->walk_ms = 0;
    for (double &ms : b->pass_ms) ms = 0;
    const size_t slots = b->slot_pair.size();
    b->ran = true;
    if (slots == 0) return 0;
    (void)hipGetLastError();  // an earlier, unrelated call's error is not this run's
    hipEvent_t start = ctx->event(0), stop = ctx->event(1), begin = ctx->event(2), end = ctx->event(3);
    hipEventRecord(begin, ctx->stream);
    std::vector<int32_t> all(slots);
    for (size_t s = 0; s < slots; ++s) all[s] = (int32_t)s;
    whole_views(b);
    if (!b->view.write(ctx, b->view_h.data(), b->view_h.size() * 4)) return failed("views");
    const bool align = b->kind == DINARA_BATCH_ALIGNMENTS;
    if (align) hipMemsetAsync(b->cigar_used.ptr, 0, 8, ctx->stream);
    int64_t code = 0;
    if (!b->local) {
        code = run_pass(b, align ? SCORE_TRACE : SCORE_GLOBAL, 0, all, start, stop, &b->pass_ms[align ? 3 : 0],
                        &b->walk_ms);
    } else {
        code = run_pass(b, SCORE_LOCAL, 0, all, start, stop, &b->pass_ms[1], &b->walk_ms);
        if (code == 0 && align) {
            // The reach back from every local end above zero, over both prefixes reversed (reference
            // vector_score.mojo:37-162), then the span between start and end aligned globally.
            if (!read_slots(b, b->score, b->score_h) || !read_slots(b, b->end_row, b->end_row_h) ||
                !read_slots(b, b->end_col, b->end_col_h))
                return failed("local ends");
            std::vector<int32_t> reaching;
            for (size_t s = 0; s < slots; ++s) {
                int32_t *v = &b->view_h[4 * s];
                if (b->score_h[s] > 0) {
                    v[0] = 0;
                    v[1] = b->end_row_h[s];
                    v[2] = 0;
                    v[3] = b->end_col_h[s];
                    reaching.push_back((int32_t)s);
                } else {
                    v[1] = v[3] = 0;
                }
            }
            if (!b->view.write(ctx, b->view_h.data(), b->view_h.size() * 4) ||
                !b->target.write(ctx, b->score_h.data(), slots * 4))
                return failed("reach views");
            code = run_pass(b, SCORE_REACH, 1, reaching, start, stop, &b->pass_ms[2], &b->walk_ms);
            if (code == 0) {
                if (!read_slots(b, b->end_row, b->reach_row_h) || !read_slots(b, b->end_col, b->reach_col_h))
                    return failed("reach starts");
                for (int32_t s : reaching) {
                    int32_t *v = &b->view_h[4 * s];
                    const int32_t er = b->end_row_h[s], ec = b->end_col_h[s];
                    const int32_t sr = b->reach_row_h[s] ? er - b->reach_row_h[s] : 0;
                    const int32_t sc = b->reach_row_h[s] ? ec - b->reach_col_h[s] : 0;
                    v[0] = sr;
                    v[1] = er - sr;
                    v[2] = sc;
                    v[3] = ec - sc;
                }
                if (!b->view.write(ctx, b->view_h.data(), b->view_h.size() * 4)) return failed("trace views");
                code = run_pass(b, SCORE_TRACE, 0, reaching, start, stop, &b->pass_ms[3], &b->walk_ms);
            }
        }
    }
    hipEventRecord(end, ctx->stream);
    if (hipEventSynchronize(end) != hipSuccess) return failed("run");
    float ms = 0;
    hipEventElapsedTime(&ms, begin, end);
    b->run_ms = ms;
    return code;
}

int64_t fetch(Batch *b) {
    const size_t slots = b->slot_pair.size();
    if (slots == 0) return 0;
    if (!read_slots(b, b->score, b->score_h)) return failed("scores");
    if (b->kind == DINARA_BATCH_ALIGNMENTS) {
        uint64_t used = 0;
        b->cigar_offset_h.resize(slots);
        b->cigar_length_h.resize(slots);
        if (!b->cigar_used.read(b->ctx, &used, 8) ||
            !b->cigar_offset.read(b->ctx, b->cigar_offset_h.data(), slots * 8) ||
            !b->cigar_length.read(b->ctx, b->cigar_length_h.data(), slots * 4))
            return failed("cigar offsets");
        b->cigars_h.resize(std::max<uint64_t>(used, 1));
        if (used > b->cigars.size) {
            if (std::getenv("DINARA_DEBUG"))
                std::fprintf(stderr, "dinara: CIGARs claim %llu bytes of %zu\n", (unsigned long long)used, b->cigars.size);
            return DINARA_OUT_OF_MEMORY;
        }
        if (used > 0 && !b->cigars.read(b->ctx, b->cigars_h.data(), used)) return failed("cigars");
    }
    return 0;
}

Batch *batch_new(int64_t gpu, int64_t kind, int64_t pairs, const char *const *refs, const int64_t *ref_lens,
                 const char *const *qrys, const int64_t *qry_lens, const dinara_scoring *scoring,
                 const dinara_mode *mode, int64_t eqx, int64_t threads, int64_t *status) {
    auto fail = [&](int64_t code) -> Batch * {
        if (status) *status = code;
        return nullptr;
    };
    ParsedScoring parsed = parse_scoring(scoring);
    if (parsed.code) return fail(parsed.code);
    bool local = false;
    if (const int64_t code = check_scoring_mode(mode, &local)) return fail(code);
    if (kind != DINARA_BATCH_DISTANCES && kind != DINARA_BATCH_ALIGNMENTS) return fail(DINARA_INVALID_MODE);
    DeviceContext *ctx = device_context(gpu);
    if (!ctx) return fail(DINARA_NO_DEVICE);
    auto *b = new Batch();
    b->gpu = (int)gpu;
    b->ctx = ctx;
    b->kind = kind;
    b->local = local;
    b->eqx = eqx != 0;
    b->pairs = std::max<int64_t>(pairs, 0);
    b->threads = threads;
    b->scoring = parsed;
    b->refs = refs;
    b->qrys = qrys;
    int64_t code;
    {
        DeviceGuard guard(ctx);
        code = prepare(b, refs, ref_lens, qrys, qry_lens);
    }
    if (code) {
        {
            DeviceGuard guard(ctx);
            delete b;
        }
        return fail(code);
    }
    if (status) *status = 0;
    return b;
}

}  // namespace

// ------------------------------------------------------------------ the C ABI

extern "C" {

dinara_scoring_batch *dinara_scoring_batch_new(int64_t gpu, int64_t kind, int64_t pairs,
                                               const char *const *references, const int64_t *reference_lengths,
                                               const char *const *queries, const int64_t *query_lengths,
                                               const dinara_scoring *scoring, const dinara_mode *mode, int64_t eqx,
                                               int64_t threads, int64_t *status) {
    return batch_new(gpu, kind, pairs, references, reference_lengths, queries, query_lengths, scoring, mode, eqx,
                     threads, status);
}

int64_t dinara_scoring_batch_run(dinara_scoring_batch *b) {
    if (!b) return DINARA_INVALID_MODE;
    return run(b);
}

int64_t dinara_scoring_batch_scores(dinara_scoring_batch *b, int64_t *scores) {
    if (!b || !b->ran) return DINARA_INVALID_MODE;
    {
        DeviceGuard guard(b->ctx);
        if (const int64_t code = fetch(b)) return code;
    }
    // Scores may be below zero: a pair the batch cannot score fails the call with its code, the first such pair's
    // in the batch's order, as the reference's batch raises the first pair's error (api.mojo:821-841).
    for (int64_t i = 0; i < b->pairs; ++i)
        if (b->code[i] < 0) return b->code[i];
    for (int64_t i = 0; i < b->pairs; ++i) scores[i] = b->host_score[i];
    for (size_t s = 0; s < b->slot_pair.size(); ++s) scores[b->slot_pair[s]] = b->score_h[s];
    return 0;
}

// Alignments in the caller's order; `references`/`queries` must be the batch's own letters (for local spans).
int64_t dinara_scoring_batch_alignments(dinara_scoring_batch *b, const char *const *references,
                                        const char *const *queries, dinara_alignment *alignments,
                                        int64_t *statuses) {
    if (!b || !b->ran || b->kind != DINARA_BATCH_ALIGNMENTS) return DINARA_INVALID_MODE;
    {
        DeviceGuard guard(b->ctx);
        if (const int64_t code = fetch(b)) return code;
    }
    std::vector<int32_t> slot_of(b->pairs, -1);
    for (size_t s = 0; s < b->slot_pair.size(); ++s) slot_of[b->slot_pair[s]] = (int32_t)s;
    std::atomic<bool> oom{false};
    parallel_for(b->pairs, b->threads, [&](int64_t lo, int64_t hi) {
        for (int64_t p = lo; p < hi; ++p) {
            dinara_alignment &out = alignments[p];
            out.cigar = nullptr;
            out.cigar_length = 0;
            const int64_t n = b->rows[p], m = b->cols[p];
            if (b->code[p] < 0) {
                statuses[p] = b->code[p];
                continue;
            }
            int64_t score, rs = 0, re = n, qs = 0, qe = m, length = 0;
            const char *text = nullptr;
            std::string empty_side;
            if (b->code[p] == 1) {
                // An empty side: global, one gap of the other's letters; local, nothing aligned.
                score = b->host_score[p];
                if (b->local) {
                    re = qe = 0;
                } else if (n + m > 0) {
                    empty_side = std::to_string(n + m) + (n > 0 ? "D" : "I");
                    text = empty_side.data();
                    length = (int64_t)empty_side.size();
                }
            } else {
                const int32_t s = slot_of[p];
                const int32_t *v = &b->view_h[4 * s];
                score = b->score_h[s];
                if (b->local) {
                    if (v[1] == 0 && v[3] == 0) {
                        score = 0;
                        re = qe = 0;
                    } else {
                        // A local alignment's spans: the last place its letters occur (reference
                        // scoring.mojo:724-745).
                        rs = rfind(references[p], n, references[p] + v[0], v[1]);
                        re = rs + v[1];
                        qs = rfind(queries[p], m, queries[p] + v[2], v[3]);
                        qe = qs + v[3];
                    }
                }
                if (!(b->local && v[1] == 0 && v[3] == 0)) {
                    text = b->cigars_h.data() + b->cigar_offset_h[s];
                    length = b->cigar_length_h[s];
                }
            }
            char *cigar = static_cast<char *>(std::malloc(length + 1));
            if (!cigar) {
                statuses[p] = DINARA_OUT_OF_MEMORY;
                oom.store(true);
                continue;
            }
            if (length > 0) std::memcpy(cigar, text, length);
            cigar[length] = 0;
            out.cost = -score;
            out.score = score;
            out.reference_start = rs;
            out.reference_end = re;
            out.query_start = qs;
            out.query_end = qe;
            out.cigar = cigar;
            out.cigar_length = length;
            statuses[p] = 0;
        }
    });
    return 0;
}

void dinara_scoring_batch_free(dinara_scoring_batch *b) {
    if (!b) return;
    DeviceGuard guard(b->ctx);
    delete b;
}

double dinara_scoring_batch_stat(dinara_scoring_batch *b, int64_t what) {
    if (!b) return 0;
    switch (what) {
        case 0: return b->run_ms;
        case 2: return b->launches;
        case 10: return b->pass_ms[0];
        case 11: return b->pass_ms[1];
        case 12: return b->pass_ms[2];
        case 13: return b->pass_ms[3];
        default: return 0;
    }
}

int64_t dinara_gpu_scores(int64_t gpu, int64_t pairs, const char *const *references,
                          const int64_t *reference_lengths, const char *const *queries, const int64_t *query_lengths,
                          const dinara_scoring *scoring, const dinara_mode *mode, int64_t threads, int64_t *scores) {
    int64_t status = 0;
    Batch *b = batch_new(gpu, DINARA_BATCH_DISTANCES, pairs, references, reference_lengths, queries, query_lengths,
                         scoring, mode, 1, threads, &status);
    if (!b) return status;
    int64_t code = run(b);
    if (code == 0) code = dinara_scoring_batch_scores(b, scores);
    dinara_scoring_batch_free(b);
    return code;
}

int64_t dinara_gpu_scoring_alignments(int64_t gpu, int64_t pairs, const char *const *references,
                                      const int64_t *reference_lengths, const char *const *queries,
                                      const int64_t *query_lengths, const dinara_scoring *scoring,
                                      const dinara_mode *mode, int64_t eqx, int64_t threads,
                                      dinara_alignment *alignments, int64_t *statuses) {
    int64_t status = 0;
    Batch *b = batch_new(gpu, DINARA_BATCH_ALIGNMENTS, pairs, references, reference_lengths, queries,
                         query_lengths, scoring, mode, eqx, threads, &status);
    if (!b) return status;
    int64_t code = run(b);
    if (code == 0) code = dinara_scoring_batch_alignments(b, references, queries, alignments, statuses);
    dinara_scoring_batch_free(b);
    return code;
}

}  // extern "C"
