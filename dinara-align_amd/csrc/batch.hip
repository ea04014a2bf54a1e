// The C ABI's batch calls under `dinara_costs`: validation with the reference's result codes, packing into the
// device tape, the sweep launches class by class until every pair is proven, and results in the caller's
// order. Reference behaviour followed: c/dinara.mojo:111-190 (argument parsing and codes), :96-108 (per-pair
// refusals), :563-819 (batch results), lanes.mojo:1163-1252 (first band), api.mojo:1188-1213 (alignments).

#include <algorithm>
#include <chrono>
#include <functional>
#include <atomic>
#include <memory>
#include <new>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <thread>
#include <vector>

#include "../../include/dinara.h"
#include "carved.h"
#include "common.h"
#include "device.h"
#include "edit.h"
#include "full.h"
#include "host.h"
#include "learned.h"

using namespace dinara;

namespace {

constexpr int64_t UNBOUNDED = int64_t(1) << 60;
constexpr int64_t MAX_COST = int64_t(1) << 32;

// The costs a batch asked for, checked as the reference checks them (modes.mojo:84-143); `code` is nonzero
// when the batch fails whole.
struct ParsedCosts {
    int64_t code = 0;
    int64_t mismatch = 1, opening = 0, extension = 1, del_opening = 0, del_extension = 1;
    bool two_piece = false;
    // The second pieces when two_piece (reference modes.mojo:99-143): a side given none counts its one piece twice.
    int64_t opening2 = 0, extension2 = 1, del_opening2 = 0, del_extension2 = 1;
};

ParsedCosts parse_costs(const dinara_costs *c) {
    ParsedCosts p;
    if (!c) return p;  // unit costs
    auto invalid = [&]() {
        p.code = DINARA_INVALID_COSTS;
        return p;
    };
    if (c->mismatch <= 0 || c->extension <= 0 || c->opening < 0) return invalid();
    if (c->mismatch > MAX_COST || c->opening > MAX_COST || c->extension > MAX_COST) return invalid();
    p.mismatch = c->mismatch;
    p.opening = c->opening;
    p.extension = c->extension;
    p.del_opening = c->opening;
    p.del_extension = c->extension;
    if (c->opening2 >= 0) {
        if (c->extension2 <= 0) return invalid();
        if (c->opening2 > MAX_COST || c->extension2 > MAX_COST) return invalid();
        p.two_piece = true;
        p.opening2 = c->opening2;
        p.extension2 = c->extension2;
    }
    p.del_opening2 = p.opening2;
    p.del_extension2 = p.extension2;
    if (c->deletion_extension != 0) {
        if (c->deletion_extension <= 0 || c->deletion_opening < 0) return invalid();
        if (c->deletion_opening2 >= 0 && c->deletion_extension2 <= 0) return invalid();
        if (c->deletion_opening > MAX_COST || c->deletion_extension > MAX_COST) return invalid();
        if (c->deletion_opening2 > MAX_COST || c->deletion_extension2 > MAX_COST) return invalid();
        p.del_opening = c->deletion_opening;
        p.del_extension = c->deletion_extension;
        if (c->deletion_opening2 >= 0) {
            p.del_opening2 = c->deletion_opening2;
            p.del_extension2 = c->deletion_extension2;
            if (!p.two_piece) {  // insertions gain a second piece: their own once more
                p.opening2 = p.opening;
                p.extension2 = p.extension;
            }
            p.two_piece = true;
        } else {
            p.del_opening2 = p.del_opening;  // deletions lack one: their own once more
            p.del_extension2 = p.del_extension;
        }
    }
    return p;
}

// The mode: global is all this library serves under Costs. Invalid modes are refused as the reference refuses
// them (modes.mojo:261-386, c/dinara.mojo:130-151); valid ones it does not serve as DINARA_INVALID_MODE.
int64_t check_mode(const dinara_mode *m) {
    if (!m) return 0;
    if (m->kind == DINARA_EXTENSION) {
        if (m->match_score < 0 || m->match_score > MAX_COST) return DINARA_INVALID_COSTS;
        return DINARA_INVALID_MODE;
    }
    if (m->kind == DINARA_LOCAL) {
        if (m->match_score < 0 || m->match_score > MAX_COST) return DINARA_INVALID_COSTS;
        return DINARA_INVALID_MODE;
    }
    if (m->kind != DINARA_ENDS_FREE) return DINARA_INVALID_MODE;
    if (std::min(std::min(m->reference_start, m->reference_end), std::min(m->query_start, m->query_end)) < 0)
        return DINARA_INVALID_MODE;
    if (m->match_score < 0 || m->match_score > MAX_COST) return DINARA_INVALID_COSTS;
    // Free ends that minimize the costs are served; a match score (maximized) is not.
    if (m->match_score > 0) return DINARA_INVALID_MODE;
    return 0;
}

// Whether a valid mode frees any letters at an end (reference modes.mojo:261-386: ends_free and its presets).
bool free_mode(const dinara_mode *m) {
    return m && (m->reference_start | m->reference_end | m->query_start | m->query_end) != 0;
}

struct Options {
    int64_t band_low = -UNBOUNDED, band_high = UNBOUNDED;  // library terms: reference less query letters
    int64_t max_cost = -1;
    bool eqx = true;
    bool right_ties = false;
};

Options parse_options(const dinara_options *o) {
    Options p;
    if (!o) return p;
    p.band_low = std::min(std::max(o->band_low, -UNBOUNDED), UNBOUNDED);
    p.band_high = std::min(std::max(o->band_high, -UNBOUNDED), UNBOUNDED);
    p.max_cost = o->max_cost;
    p.eqx = o->eqx != 0;
    p.right_ties = o->right_ties != 0;
    return p;
}

// Zero for a pair the library takes, else why not (reference c/dinara.mojo:96-108).
int64_t pair_code(const char *ref, int64_t n, const char *qry, int64_t m) {
    if (n < 0 || m < 0) return DINARA_INVALID_LENGTH;
    auto plain = [](const char *s, int64_t len) {
        const unsigned char *u = reinterpret_cast<const unsigned char *>(s);
        unsigned refused = 0;  // no early exit: the loop vectorizes
        for (int64_t i = 0; i < len; ++i) refused |= u[i] >= 0xFE;
        return refused == 0;
    };
    if ((n > 0 && !plain(ref, n)) || (m > 0 && !plain(qry, m))) return DINARA_UNSUPPORTED_SYMBOLS;
    return 0;
}

// The class a pair's first band takes: the narrowest group band holding `need` diagonals (one fewer than its width:
// a group band starts on an even diagonal), else the narrowest block band, else the narrowest wave band (as the
// device's class_for).
int first_class_for(int need, bool groups, int block_mask, bool user_band) {
    if (groups && !user_band)
        for (int cls = CLASS_GROUP32; cls <= CLASS_GROUP256; ++cls)
            if (need < class_width(cls)) return cls;
    if (!user_band)
        for (int cls = CLASS_BLOCK_FIRST; cls < CLASSES; ++cls)
            if (((block_mask >> (cls - CLASS_BLOCK_FIRST)) & 1) && need < class_width(cls)) return cls;
    for (int cls = CLASS_WAVE_FIRST; cls <= CLASS_WAVE_LAST; ++cls)
        if (need <= class_width(cls)) return cls;
    return CLASS_TOO_WIDE;
}

// LDS bytes one block of class `cls` stages: four pairs' letters for a group, one pair's and the reads past it
// for a block.
size_t block_letters(int max_letters, int cls) { return ((size_t)max_letters + class_width(cls) / 2 + 16 + 3) & ~size_t(3); }

}  // namespace

// ------------------------------------------------------------------ the batch

struct dinara_batch {
    int gpu = 0;
    int64_t kind = DINARA_BATCH_DISTANCES;
    int64_t pairs = 0;
    int64_t threads = 0;
    Options options;
    ParsedCosts parsed;
    CostParams costs{};
    bool user_band = false;
    int user_low = 0, user_high = 0;
    int initial_width = 32;
    std::vector<int64_t> code;            // per pair: 0 on the device, else its own result code
    std::vector<int64_t> reference_length, query_length;
    std::vector<int32_t> slot_pair;       // slot -> pair
    std::vector<int32_t> rows, cols;      // per slot
    std::vector<int32_t> first_low, first_high;
    std::vector<int32_t> first_lists[CLASSES];
    std::vector<int64_t> scratch_base;
    int bits = 2;
    bool reversed = false;                // the tape holds both sequences back to front (Ties.LEFT)
    bool doubling = false;                // retries double the band rather than take the band the cost proves
    size_t lds_limit = 0;
    DeviceContext *ctx = nullptr;
    int max_letters = 0;                  // the most letters of a pair, both sequences
    int32_t max_rows = 0;
    int group_letters = 4;                // LDS bytes a pair takes in a group launch
    int group_max_letters = 0;            // the longest pair a group launch stages
    bool groups_fit = false;              // four pairs' letters fit one block's LDS, their costs the tags
    int group16_top = 0;                  // > 0: packed 16-bit group cells, this the unreachable value
    int block_mask = 0;                   // the block classes whose letters and tags fit
    int64_t dearest = 1;                  // the dearest single step: a substitution or a gap's first letter
    // Free ends (wave sweeps): the letters each end of each sequence may leave out (reference start, end, query
    // start, end), the phase a run is in (1 the forward phase finding cost and end, 2 the anchored phase keeping
    // decisions), and per slot its views, free letters and first bands of each phase.
    bool free_ends = false;
    int64_t ends[4] = {0, 0, 0, 0};
    int phase = 0;
    std::vector<int32_t> a_rows, a_cols, a_at, a_free, a_first_low, a_first_high, a_first_all;
    std::vector<int64_t> a_first_offset;
    std::vector<int32_t> a_status_h, a_cost_h, a_end_row_h, a_end_col_h, b_end_row_h, b_end_col_h;
    std::vector<int32_t> b_rows_h, b_cols_h;
    std::vector<DeviceBuffer *> arenas;   // one per launch of a run, kept from run to run
    std::vector<int64_t> first_offset;    // where each class's first list starts in first_lists_d
    // device memory
    DeviceBuffer tape, group_base, group_ref_words, rows_d, cols_d, want_low, want_high, first_low_d,
        first_high_d, cost, status, walk_low, walk_class, cigar_scratch, cigar_scratch_base, cigars, cigar_used,
        cigar_offset, cigar_length, retry_lists, retry_counts, first_lists_d, walk_flags, view_rows_d, view_cols_d,
        view_ref_at_d, view_qry_at_d, free_d, end_row_d, end_col_d;
    ~dinara_batch() {
        for (DeviceBuffer *buffer : arenas) delete buffer;
        if (edit_graph) (void)hipGraphExecDestroy(edit_graph);
        if (pinned_counts) (void)hipHostFree(pinned_counts);
    }
    // Alignments launched one by one: the queue counts a run reads back while its last walk runs (pinned, made at
    // the first such run).
    int32_t *pinned_counts = nullptr;
    // the last run's results, by slot
    std::vector<int32_t> cost_h, status_h, cigar_length_h;
    std::vector<int64_t> cigar_offset_h;
    std::vector<char> cigars_h;
    bool ran = false;
    // statistics of the last run
    double run_ms = 0, sweep_ms = 0, walk_ms = 0, launches = 0, first_proven = 0;
    double class_ms[CLASSES] = {}, class_launches[CLASSES] = {}, class_pairs[CLASSES] = {};

    // The unit-cost path (edit.hip): unit costs times `unit_scale` and no user band.
    bool unit = false;
    int unit_scale = 1;
    bool use_graph = true;                // replay the run's launches as one captured hipGraph
    bool persistent = true;               // pair groups refill from their queue; else one pair a group
    bool keep_first = false;              // alignments: every sweep keeps its decisions (one pass when proven)
    bool side_by_side_bins = true;        // the bins' first queues (and tracing sweeps) on streams of their own
    bool timed = true;                    // launches made one by one record an event pair each
    int grid_per_cu = 8;                  // persistent workgroups a compute unit
    int start_bin[EDIT_LENGTH_CLASSES] = {};  // the least first bin, by the pair's longer side
    uint32_t first_bins = EDIT_ALL_BINS;  // the bins a first band may take (bit b for bin b): the first pass's edges
    std::vector<int32_t> edit_first_low;  // per slot: its first band's lowest diagonal
    std::vector<int32_t> edit_first_bin;
    int64_t first_count[EDIT_BINS] = {};  // pairs each bin's first queue holds
    int64_t retry_reach[EDIT_BINS] = {};  // pairs a bin's retry queue may get: first bin narrower, widest no narrower
    int64_t trace_reach[EDIT_BINS] = {};  // pairs a bin's tracing sweep may meet
    DeviceBuffer features_d, predictions_d;  // learned seeding on the device
    int seeded_start[EDIT_LENGTH_CLASSES] = {-1};  // the start bins the device seeding's launch shape was made for
    DeviceBuffer edit_band_low, edit_first_low_d, edit_lists, edit_retry_lists, edit_retry_counts, edit_later_lists,
        edit_later_counts, edit_counts, edit_first_lists, edit_first_counts, trace_lists,
        trace_counts, trace_offset, walk_ready, walk_moves, walk_early, deferred, deferred_count, deferred_held,
        arena_used, edit_stats, edit_work;
    size_t arena_capacity = 0;            // the bytes of the context's arena this batch's runs use
    const void *arena_ptr = nullptr;      // where that arena was when the captured graph was made
    size_t arena_want = 0;                // the bytes it asks for (0: not worked out yet)
    hipGraphExec_t edit_graph = nullptr;
    double edit_ms[EDIT_BINS] = {}, trace_ms[EDIT_BINS] = {}, edit_sweeps[EDIT_BINS] = {},
           trace_sweeps[EDIT_BINS] = {}, graph_used = 0, trace_passes = 0, walked_first = 0;
    double sweep_work[EDIT_BINS] = {}, trace_work[EDIT_BINS] = {};  // word-columns swept by each bin
    bool stats_pending = false;  // the unit path's device counters of the last run, read when first asked for
    int64_t stats_slots = 0;
    // The whole-matrix path (full.hip) for pairs the band sweeps leave unsettled: whether any pair is long enough
    // to need it (-1 not worked out yet), and the last run's count and time.
    int may_need_full = -1;
    bool whole = false;  // every pair takes the whole-matrix path, no band sweep (for checking and comparison)
    double full_pairs = 0, full_ms = 0;
    DeviceBuffer full_list, full_count, full_decisions, full_carry, full_claim, full_slots, full_strips,
        full_decision_at, full_carry_at;
};

namespace {

int64_t unit_lists(dinara_batch *b);
int64_t unit_buffers(dinara_batch *b, const std::vector<int32_t> &lists, const std::vector<int32_t> &counts,
                     size_t cap);

// Builds the tape: pairs by length, longest first, 64 to a group, each group's words transposed.
int64_t prepare(dinara_batch *b, const char *const *refs, const char *const *qrys) {
    const int64_t pairs = b->pairs;
    std::vector<int32_t> order;
    order.reserve(pairs);
    for (int64_t i = 0; i < pairs; ++i)
        if (b->code[i] == 0) order.push_back((int32_t)i);
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
        return std::max(b->reference_length[x], b->query_length[x]) >
               std::max(b->reference_length[y], b->query_length[y]);
    });
    const int64_t slots = (int64_t)order.size();
    b->slot_pair = order;
    b->first_offset.assign(CLASSES + 1, 0);
    if (slots == 0) return 0;  // every pair answered by its own code
    b->rows.resize(slots);
    b->cols.resize(slots);
    for (int64_t s = 0; s < slots; ++s) {
        b->rows[s] = (int32_t)b->reference_length[order[s]];
        b->cols[s] = (int32_t)b->query_length[order[s]];
    }
    // One pass over every pair's letters: which of A, C, G, T it holds and whether it holds any other byte. Two
    // bits a letter when every pair holds at most four distinct bytes (each pair its own codes: only equality
    // counts), else a byte a letter.
    static const ByteClasses classes;
    std::vector<uint8_t> plain(slots, 0);  // 1: the pair's letters are A, C, G and T alone
    std::atomic<bool> wide{false};
    parallel_for(slots, b->threads, [&](int64_t lo, int64_t hi) {
        for (int64_t s = lo; s < hi; ++s) {
            const int32_t p = order[s];
            const uint32_t seen = classes.seen(refs[p], b->reference_length[p]) |
                                  classes.seen(qrys[p], b->query_length[p]);
            if (!(seen & ByteClasses::OTHER)) {
                plain[s] = 1;
                continue;
            }
            if (wide.load(std::memory_order_relaxed)) continue;
            bool present[256] = {false};
            int distinct = 0;
            auto scan = [&](const char *text, int64_t len) {
                const unsigned char *u = reinterpret_cast<const unsigned char *>(text);
                for (int64_t i = 0; i < len && distinct <= 4; ++i)
                    if (!present[u[i]]) {
                        present[u[i]] = true;
                        ++distinct;
                    }
            };
            scan(refs[p], b->reference_length[p]);
            scan(qrys[p], b->query_length[p]);
            if (distinct > 4) wide.store(true);
        }
    });
    b->bits = wide.load() ? 8 : 2;
    const int per = 32 / b->bits;
    const int64_t groups = (slots + 63) / 64;
    std::vector<int64_t> gbase(std::max<int64_t>(groups, 1));
    std::vector<int32_t> gref(std::max<int64_t>(groups, 1)), gqry(std::max<int64_t>(groups, 1));
    int64_t words = 0;
    for (int64_t g = 0; g < groups; ++g) {
        int32_t rw = 0, qw = 0;
        for (int64_t s = g * 64; s < std::min(slots, (g + 1) * 64); ++s) {
            rw = std::max<int32_t>(rw, (b->rows[s] + per - 1) / per);
            qw = std::max<int32_t>(qw, (b->cols[s] + per - 1) / per);
        }
        gbase[g] = words;
        gref[g] = rw;
        gqry[g] = qw;
        words += (int64_t)(rw + qw) * 64;
    }
    // Every word is written below (a pair's words past its letters are zero), so the tape needs no clearing but
    // its padding: the unit-cost sweeps' letter cursors read up to three words past a sequence's last without a bounds
    // check (edit.hip), which past the last group's words lands here.
    const int64_t padded = words + TAPE_PADDING_WORDS;
    std::unique_ptr<uint32_t[]> tape(new (std::nothrow) uint32_t[padded]);
    if (!tape) return DINARA_OUT_OF_MEMORY;
    std::fill(tape.get() + words, tape.get() + padded, 0u);
    // Ties.LEFT is the rule over both sequences reversed, so their tape is reversed (lanes.mojo:1523-1525).
    const bool reversed = b->kind == DINARA_BATCH_ALIGNMENTS && !b->options.right_ties && !b->free_ends;
    b->reversed = reversed;
    parallel_for(groups, b->threads, [&](int64_t lo, int64_t hi) {
        for (int64_t g = lo; g < hi; ++g) {
            const int64_t first = g * 64, last = std::min(slots, (g + 1) * 64);
            for (int64_t s = first; s < (g + 1) * 64; ++s) {
                uint32_t *ref_out = tape.get() + gbase[g] + (s & 63);
                uint32_t *qry_out = ref_out + (int64_t)gref[g] * 64;
                if (s >= last) {  // the group's empty slots
                    for (int32_t w = 0; w < gref[g]; ++w) ref_out[(int64_t)w * 64] = 0;
                    for (int32_t w = 0; w < gqry[g]; ++w) qry_out[(int64_t)w * 64] = 0;
                    continue;
                }
                const int32_t p = order[s];
                const int64_t n = b->reference_length[p], m = b->query_length[p];
                if (b->bits == 2 && plain[s]) {
                    // A, C, G and T are codes 0, 1, 2, 3 in every pair (the k-mer features read them so).
                    classes.pack2(refs[p], n, reversed, ref_out, gref[g]);
                    classes.pack2(qrys[p], m, reversed, qry_out, gqry[g]);
                    continue;
                }
                // Two-bit codes for a pair with other bytes: A, C, G and T keep 0, 1, 2, 3, any other byte the
                // least code its pair leaves free; or a byte a letter. Only equality counts to the sweeps.
                int16_t codes[256];
                for (int i = 0; i < 256; ++i) codes[i] = b->bits == 8 ? (int16_t)i : -1;
                if (b->bits == 2) {
                    bool used[4] = {false, false, false, false};
                    auto fixed = [](unsigned char byte) { return byte == 'A' ? 0 : byte == 'C' ? 1 : byte == 'G' ? 2 : byte == 'T' ? 3 : -1; };
                    auto note = [&](const char *text, int64_t len, bool fixed_pass) {
                        const unsigned char *u = reinterpret_cast<const unsigned char *>(text);
                        for (int64_t i = 0; i < len; ++i) {
                            if (codes[u[i]] >= 0) continue;
                            const int f = fixed(u[i]);
                            if (fixed_pass && f >= 0) {
                                codes[u[i]] = (int16_t)f;
                                used[f] = true;
                            } else if (!fixed_pass) {
                                int free_code = 0;
                                while (used[free_code]) ++free_code;
                                codes[u[i]] = (int16_t)free_code;
                                used[free_code] = true;
                            }
                        }
                    };
                    note(refs[p], n, true);
                    note(qrys[p], m, true);
                    note(refs[p], n, false);
                    note(qrys[p], m, false);
                }
                auto pack = [&](const char *text, int64_t len, uint32_t *out, int32_t words_here) {
                    const unsigned char *u = reinterpret_cast<const unsigned char *>(text);
                    for (int32_t w = 0; w < words_here; ++w) {
                        uint32_t word = 0;
                        for (int i = 0; i < per; ++i) {
                            const int64_t at = (int64_t)w * per + i;
                            if (at >= len) break;
                            const unsigned char byte = reversed ? u[len - 1 - at] : u[at];
                            word |= (uint32_t)codes[byte] << (b->bits * i);
                        }
                        out[(int64_t)w * 64] = word;
                    }
                };
                pack(refs[p], n, ref_out, gref[g]);
                pack(qrys[p], m, qry_out, gqry[g]);
            }
        }
    });
    b->max_rows = 0;
    b->max_letters = 0;
    for (int64_t s = 0; s < slots; ++s) {
        b->max_rows = std::max(b->max_rows, b->rows[s]);
        b->max_letters = std::max<int>(b->max_letters, b->rows[s] + b->cols[s]);
    }
    // A group pair's LDS region: both sequences and room for reads up to half the widest group band (256) past
    // the query's end (group.hip). Pairs up to a quarter of the LDS take group bands; longer ones take blocks.
    const int group_cap = (int)(b->lds_limit / 4) - 128 - 16 - 4;
    b->group_max_letters = 0;
    for (int64_t s = 0; s < slots; ++s)
        if (b->rows[s] + b->cols[s] <= group_cap)
            b->group_max_letters = std::max(b->group_max_letters, b->rows[s] + b->cols[s]);
    b->group_letters = (b->group_max_letters + 128 + 16 + 3) & ~3;
    // The group sweep holds costs times 16 with a tag (group.hip) and runs up to half its widest band past the
    // corner: every value it makes stays under 2^30.
    // The group and block sweeps hold one gap piece: two-piece batches take the wave sweeps alone.
    b->groups_fit = !b->parsed.two_piece && !b->free_ends && group_cap > 0 && (size_t)4 * b->group_letters <= b->lds_limit &&
                    ((int64_t)b->group_max_letters + 2 * 256 + 8) * b->dearest < REAL_LIMIT;
    // Packed 16-bit group cells (group.hip) for shifted costs (equal extensions e) whose mismatch is no dearer than
    // two extension letters: a tagged value then lies in [-16 e (n + m), 32 opening + 15] and a cell no path
    // reaches stays above `top` - 16 e (n + m), so both fit 16 bits, apart, while 16 e (n + m) + 32 opening + 15
    // stays under `top` (less a margin); `top` leaves room for an opening's add. DINARA_GROUP_CELLS32 set keeps
    // 32-bit cells, for comparison.
    b->group16_top = 0;
    {
        const CostParams &k = b->costs;
        const int64_t e = k.del_ext, opening = std::max(k.del_open, k.ins_open);
        const int64_t top = (32767 - 16 * opening - 32) & ~int64_t(15);
        if (b->groups_fit && !std::getenv("DINARA_GROUP_CELLS32") && k.pieces != 2 && k.del_ext == k.ins_ext &&
            e >= 0 && k.mismatch >= 0 && k.mismatch <= 2 * e && std::min(k.del_open, k.ins_open) >= 0 &&
            16 * e * b->group_max_letters + 32 * opening + 15 < top - 64 && (size_t)8 * b->group_letters <= b->lds_limit)
            b->group16_top = (int)top;
    }
    b->block_mask = 0;
    for (int cls = CLASS_BLOCK_FIRST; cls < CLASSES && !b->parsed.two_piece && !b->free_ends; ++cls)
        if (block_letters(b->max_letters, cls) <= b->lds_limit &&
            ((int64_t)b->max_letters + class_width(cls) + 8) * b->dearest < REAL_LIMIT &&
            !(block_s(cls) > 8 && b->kind == DINARA_BATCH_ALIGNMENTS))  // the widest keep no decisions
            b->block_mask |= 1 << (cls - CLASS_BLOCK_FIRST);
    // The first band: the lanes' start and end diagonals and one either side (lanes.mojo:1207-1215), within the
    // user band; swept at no less than the initial width, centred.
    b->first_low.resize(slots);
    b->first_high.resize(slots);
    for (auto &l : b->first_lists) l.clear();
    if (b->free_ends) b->a_free.assign((size_t)slots * 4, 0);
    b->scratch_base.resize(slots + 1);
    int64_t scratch = 0, text = 0;
    for (int64_t s = 0; s < slots; ++s) {
        const int32_t n = b->rows[s], m = b->cols[s];
        const int32_t end = m - n;
        int32_t lo = std::min(0, end) - 1, hi = std::max(0, end) + 1;
        if (b->user_band) {
            lo = std::max(lo, b->user_low);
            hi = std::min(hi, b->user_high);
        }
        if (b->free_ends) {
            // The forward phase: Ties.LEFT over the pair as it is, Ties.RIGHT over both sequences reversed with each
            // sequence's free letters swapped end for end (reference tests/test_alignment.mojo:2054-2093). Its
            // band holds every diagonal a path may start or end on, and one either side.
            const bool right = b->options.right_ties;
            const int64_t e[4] = {right ? b->ends[1] : b->ends[0], right ? b->ends[0] : b->ends[1],
                                  right ? b->ends[3] : b->ends[2], right ? b->ends[2] : b->ends[3]};
            const int32_t f[4] = {(int32_t)std::min<int64_t>(e[0], n), (int32_t)std::min<int64_t>(e[1], n),
                                  (int32_t)std::min<int64_t>(e[2], m), (int32_t)std::min<int64_t>(e[3], m)};
            for (int k = 0; k < 4; ++k) b->a_free[4 * s + k] = f[k];
            lo = std::max(std::min(-f[0], end - f[3]) - 1, -n - 1);
            hi = std::min(std::max(f[2], end + f[1]) + 1, m + 1);
        }
        b->first_low[s] = lo;
        b->first_high[s] = hi;
        const int need = std::max(hi - lo + 1, b->initial_width);
        const int cls = first_class_for(need, b->groups_fit && n + m <= b->group_max_letters, b->block_mask,
                                        b->user_band);
        // A pair whose start and end diagonals lie further apart than the widest band is left unswept: its
        // status stays unsettled and it reports DINARA_OUT_OF_MEMORY.
        if (cls != CLASS_TOO_WIDE) b->first_lists[cls].push_back((int32_t)s);
        // A walk's moves, 2 bits each in 64-bit words; the CIGAR text, at most two bytes a move.
        b->scratch_base[s] = scratch;
        if (b->kind == DINARA_BATCH_ALIGNMENTS) {
            scratch += 8 * (((int64_t)n + m) / 32 + 1);
            text += 2 * ((int64_t)n + m) + 2;
        }
    }
    b->scratch_base[slots] = scratch;
    std::vector<int32_t> first_all;
    b->first_offset.assign(CLASSES + 1, 0);
    for (int c = 0; c < CLASSES; ++c) {
        b->first_offset[c] = (int64_t)first_all.size();
        first_all.insert(first_all.end(), b->first_lists[c].begin(), b->first_lists[c].end());
    }
    b->first_offset[CLASSES] = (int64_t)first_all.size();
    if (first_all.empty()) first_all.push_back(0);  // every pair too wide: the device list still needs room
    if (b->free_ends) {
        // The forward phase's lists and bands, uploaded again at each run (the anchored phase writes its own).
        b->a_first_low = b->first_low;
        b->a_first_high = b->first_high;
        b->a_first_all = first_all;
        b->a_first_offset = b->first_offset;
        b->a_rows = b->rows;
        b->a_cols = b->cols;
        b->a_at.assign(std::max<int64_t>(slots, 1), 0);
        const size_t cap = (size_t)std::max<int64_t>(slots, 1);
        if (!b->view_rows_d.reserve(cap * 4) || !b->view_cols_d.reserve(cap * 4) || !b->view_ref_at_d.reserve(cap * 4) ||
            !b->view_qry_at_d.reserve(cap * 4) || !b->free_d.reserve(cap * 16) || !b->end_row_d.reserve(cap * 4) ||
            !b->end_col_d.reserve(cap * 4))
            return DINARA_OUT_OF_MEMORY;
    }

    // Upload.
    DeviceContext *ctx = b->ctx;
    if (!b->tape.upload(ctx, tape.get(), (size_t)padded * 4) ||
        !b->group_base.upload(ctx, gbase.data(), gbase.size() * 8) ||
        !b->group_ref_words.upload(ctx, gref.data(), gref.size() * 4) ||
        !b->rows_d.upload(ctx, b->rows.data(), std::max<int64_t>(slots, 1) * 4) ||
        !b->cols_d.upload(ctx, b->cols.data(), std::max<int64_t>(slots, 1) * 4) ||
        !b->first_low_d.upload(ctx, b->first_low.data(), std::max<int64_t>(slots, 1) * 4) ||
        !b->first_high_d.upload(ctx, b->first_high.data(), std::max<int64_t>(slots, 1) * 4) ||
        !b->want_low.reserve(std::max<int64_t>(slots, 1) * 4) || !b->want_high.reserve(std::max<int64_t>(slots, 1) * 4) ||
        !b->cost.reserve(std::max<int64_t>(slots, 1) * 4) || !b->status.reserve(std::max<int64_t>(slots, 1) * 4) ||
        !b->walk_low.reserve(std::max<int64_t>(slots, 1) * 4) ||
        !b->walk_class.reserve(std::max<int64_t>(slots, 1) * 4) ||
        !b->retry_lists.reserve((size_t)2 * CLASSES * std::max<int64_t>(slots, 1) * 4) ||
        !b->retry_counts.reserve(2 * CLASSES * 4) ||
        !b->first_lists_d.upload(ctx, first_all.data(), std::max<size_t>(first_all.size(), 1) * 4) ||
        !b->walk_flags.reserve(std::max<int64_t>(slots, 1) * 8))
        return DINARA_OUT_OF_MEMORY;
    if (b->unit)
        if (const int64_t code = unit_lists(b)) return code;
    if (b->kind == DINARA_BATCH_ALIGNMENTS) {
        if (!b->cigar_scratch.reserve(std::max<int64_t>(scratch, 1)) ||
            !b->cigar_scratch_base.upload(ctx, b->scratch_base.data(), b->scratch_base.size() * 8) ||
            !b->cigars.reserve(std::max<int64_t>(text, 1)) || !b->cigar_used.reserve(8) ||
            !b->cigar_offset.reserve(std::max<int64_t>(slots, 1) * 8) ||
            !b->cigar_length.reserve(std::max<int64_t>(slots, 1) * 4))
            return DINARA_OUT_OF_MEMORY;
    }
    return 0;
}

SweepArgs sweep_args(dinara_batch *b) {
    SweepArgs a{};
    a.tape.words = static_cast<const uint32_t *>(b->tape.ptr);
    a.tape.group_base = static_cast<const int64_t *>(b->group_base.ptr);
    a.tape.group_ref_words = static_cast<const int32_t *>(b->group_ref_words.ptr);
    a.tape.rows = static_cast<const int32_t *>(b->rows_d.ptr);
    a.tape.cols = static_cast<const int32_t *>(b->cols_d.ptr);
    a.tape.bits = b->bits;
    a.slots = (int)b->slot_pair.size();
    a.want_low = static_cast<int32_t *>(b->want_low.ptr);
    a.want_high = static_cast<int32_t *>(b->want_high.ptr);
    a.costs = b->costs;
    a.strict = b->kind == DINARA_BATCH_ALIGNMENTS;
    a.cap = b->options.max_cost >= 0 ? b->options.max_cost : INT64_MAX / 4;
    a.user_band = b->user_band;
    a.user_low = b->user_low;
    a.user_high = b->user_high;
    a.eqx = b->options.eqx;
    a.forward_cigar = b->kind == DINARA_BATCH_ALIGNMENTS && !b->options.right_ties;
    a.record = b->kind == DINARA_BATCH_ALIGNMENTS;
    a.lds_limit = (int)b->lds_limit;
    a.cost = static_cast<int32_t *>(b->cost.ptr);
    a.status = static_cast<int32_t *>(b->status.ptr);
    a.walk_low = static_cast<int32_t *>(b->walk_low.ptr);
    a.walk_class = static_cast<int32_t *>(b->walk_class.ptr);
    a.walk_flags = static_cast<uint64_t *>(b->walk_flags.ptr);
    a.group_letters = b->group_letters;
    a.doubling = b->doubling ? 1 : 0;
    a.groups_fit = b->groups_fit;
    a.group_max_letters = b->group_max_letters;
    a.group16_top = b->group16_top;
    a.block_mask = b->block_mask;
    a.cigar_scratch = static_cast<char *>(b->cigar_scratch.ptr);
    a.cigar_scratch_base = static_cast<const int64_t *>(b->cigar_scratch_base.ptr);
    a.cigars = static_cast<char *>(b->cigars.ptr);
    a.cigar_used = static_cast<unsigned long long *>(b->cigar_used.ptr);
    a.cigar_offset = static_cast<int64_t *>(b->cigar_offset.ptr);
    a.cigar_length = static_cast<int32_t *>(b->cigar_length.ptr);
    a.retry_capacity = (int)std::max<size_t>(b->slot_pair.size(), 1);
    a.free_ends = b->free_ends ? 1 : 0;
    if (b->free_ends) {
        // The forward phase reads the pair as Ties.RIGHT's mirror or as it is; the anchored phase the other way,
        // as a global alignment's tape is (reference lanes.mojo:1523-1525).
        const bool right = b->options.right_ties;
        a.view_reversed = b->phase == 2 ? !right : right;
        a.lowest_end = b->phase == 2;
        a.record = b->kind == DINARA_BATCH_ALIGNMENTS && b->phase == 2;
        a.view_rows = static_cast<const int32_t *>(b->view_rows_d.ptr);
        a.view_cols = static_cast<const int32_t *>(b->view_cols_d.ptr);
        a.view_ref_at = static_cast<const int32_t *>(b->view_ref_at_d.ptr);
        a.view_qry_at = static_cast<const int32_t *>(b->view_qry_at_d.ptr);
        a.free = static_cast<const int32_t *>(b->free_d.ptr);
        a.end_row = static_cast<int32_t *>(b->end_row_d.ptr);
        a.end_col = static_cast<int32_t *>(b->end_col_d.ptr);
    }
    return a;
}


// Launches class `cls` over `count` slots listed at `list` (device memory), filing retries into `retry_lists` /
// `retry_counts`. Each launch keeps its decisions in an arena of its own, a fixed stride a listed pair, until the
// walk; a launch whose arena would pass the budget is cut in pieces. `arena_index` counts the run's arenas.
int64_t launch_class(dinara_batch *b, int cls, const int32_t *list, int64_t count, int32_t *retry_lists,
                     int32_t *retry_counts, size_t &arena_index) {
    DeviceContext *ctx = b->ctx;
    SweepArgs a = sweep_args(b);
    a.cls = cls;
    a.retry_lists = retry_lists;
    a.retry_counts = retry_counts;
    const bool wave = class_is_wave(cls);
    const size_t lds = wave                     ? (size_t)std::min<int>(std::max(b->max_letters, 4), (int)b->lds_limit)
                       : band_is_block(cls) ? block_letters(b->max_letters, cls)
                                            : (size_t)(b->group16_top > 0 && cls >= 1 ? 8 : 4) * b->group_letters;
    int64_t chunk = count;
    size_t stride = 0;
    if (a.record) {
        // The batch's largest pair bounds every pair's decisions: rows for a wave band, tiles of 32
        // anti-diagonals (16 bytes a diagonal pair) for a group's.
        const int64_t bytes = wave ? (int64_t)b->max_rows * class_width(cls)
                                   : ((int64_t)b->max_letters / 32 + 1) * (class_width(cls) / 2) * 16;
        stride = ((size_t)bytes + 255) & ~size_t(255);
        chunk = std::max<int64_t>(1, (int64_t)std::min<size_t>(count, ARENA_BUDGET / std::max<size_t>(stride, 1)));
    }
    for (int64_t start = 0; start < count; start += chunk) {
        a.list = list + start;
        a.count = (int)std::min(chunk, count - start);
        if (a.record) {
            if (arena_index == b->arenas.size()) b->arenas.push_back(new DeviceBuffer());
            DeviceBuffer *arena = b->arenas[arena_index++];
            if (!arena->reserve(std::max<size_t>(stride * a.count, 256))) return DINARA_OUT_OF_MEMORY;
            a.arena = static_cast<uint8_t *>(arena->ptr);
            a.arena_stride = (int64_t)stride;
        }
        launch_sweep(a, lds, ctx->stream);
        if (hipGetLastError() != hipSuccess) return DINARA_OUT_OF_MEMORY;
        b->launches += 1;
        b->class_launches[cls] += 1;
        b->class_pairs[cls] += a.count;
    }
    return 0;
}

// ------------------------------------------------------------------ the unit-cost path

// The unit path's first queues: each pair in the narrowest bin of the first pass's set that holds its first band,
// no narrower than its length class's start bin (under band doubling, the first such bin of any set whose width is a
// power of two), placed around the start and end diagonals. Caller holds the device lock.
int64_t unit_lists(dinara_batch *b) {
    const int64_t slots = (int64_t)b->slot_pair.size();
    const size_t cap = (size_t)std::max<int64_t>(slots, 1);
    std::vector<int32_t> lists(EDIT_BINS * cap, 0), counts(EDIT_BINS, 0);
    b->edit_first_low.assign(cap, 0);
    b->edit_first_bin.assign(cap, -1);
    int64_t first_count[EDIT_BINS] = {}, retry_reach[EDIT_BINS] = {}, trace_reach[EDIT_BINS] = {};
    for (int64_t s = 0; s < slots; ++s) {
        const int n = b->rows[s], m = b->cols[s];
        const int lo = b->first_low[s], hi = b->first_high[s];
        const int from = b->start_bin[edit_length_class(std::max(n, m))];
        // A band placed inside the matrix's diagonals and n + m + 3 wide proves any pair, and no traced band is
        // wider than n + m + 1: the bins a pair can reach.
        const int widest = edit_widest(n + m);
        const int bin = b->doubling ? edit_power_bin_for(hi - lo + 1, from)
                                    : edit_first_bin(hi - lo + 1, from, b->first_bins, std::max(from, widest));
        if (bin < 0) continue;  // wider than the widest band: left unsettled
        b->edit_first_low[s] = edit_place(lo, hi, edit_width(bin), n, m);
        b->edit_first_bin[s] = bin;
        lists[bin * cap + counts[bin]++] = (int32_t)s;
        ++first_count[bin];
        for (int k = bin + 1; k <= widest; ++k) ++retry_reach[k];
        for (int k = 0; k <= widest; ++k) ++trace_reach[k];
    }
    bool same_reach = true;
    for (int k = 0; k < EDIT_BINS; ++k) {
        same_reach = same_reach && first_count[k] == b->first_count[k] && retry_reach[k] == b->retry_reach[k] &&
                     trace_reach[k] == b->trace_reach[k];
        b->first_count[k] = first_count[k];
        b->retry_reach[k] = retry_reach[k];
        b->trace_reach[k] = trace_reach[k];
    }
    b->seeded_start[0] = -1;  // host-filed queues: the device seeding's shape no longer holds
    if (!same_reach && b->edit_graph) {  // the launches' shape changes
        (void)hipGraphExecDestroy(b->edit_graph);
        b->edit_graph = nullptr;
    }
    // The captured graph reads these buffers where they were: it is kept unless one moves.
    DeviceBuffer *held[] = {&b->edit_first_lists, &b->edit_first_counts, &b->edit_first_low_d, &b->edit_band_low,
                            &b->edit_lists, &b->edit_counts, &b->edit_retry_lists, &b->edit_retry_counts,
                            &b->edit_later_lists, &b->edit_later_counts, &b->trace_lists, &b->trace_counts, &b->trace_offset,
                            &b->walk_ready, &b->walk_moves, &b->walk_early, &b->deferred, &b->deferred_count, &b->deferred_held,
                            &b->arena_used, &b->edit_stats, &b->edit_work};
    void *before[sizeof(held) / sizeof(held[0])];
    for (size_t i = 0; i < sizeof(held) / sizeof(held[0]); ++i) before[i] = held[i]->ptr;
    const int64_t code = unit_buffers(b, lists, counts, cap);
    for (size_t i = 0; i < sizeof(held) / sizeof(held[0]); ++i)
        if (held[i]->ptr != before[i] && b->edit_graph) {
            (void)hipGraphExecDestroy(b->edit_graph);
            b->edit_graph = nullptr;
        }
    return code;
}

int64_t unit_buffers(dinara_batch *b, const std::vector<int32_t> &lists, const std::vector<int32_t> &counts,
                     size_t cap) {
    DeviceContext *ctx = b->ctx;
    if (!b->edit_first_lists.upload(ctx, lists.data(), lists.size() * 4) ||
        !b->edit_first_counts.upload(ctx, counts.data(), counts.size() * 4) ||
        !b->edit_first_low_d.upload(ctx, b->edit_first_low.data(), cap * 4) || !b->edit_band_low.reserve(cap * 4) ||
        !b->edit_lists.reserve(EDIT_BINS * cap * 4) || !b->edit_counts.reserve(2 * EDIT_BINS * 4) ||
        !b->edit_retry_lists.reserve(EDIT_BINS * cap * 4) || !b->edit_retry_counts.reserve(2 * EDIT_BINS * 4) ||
        !b->edit_later_lists.reserve(EDIT_BINS * cap * 4) || !b->edit_later_counts.reserve(3 * EDIT_BINS * 4) ||
        !b->trace_lists.reserve(EDIT_BINS * cap * 4) || !b->trace_counts.reserve(2 * EDIT_BINS * 4) ||
        !b->trace_offset.reserve(cap * 8) || !b->walk_ready.reserve(cap * 4) ||
        !b->walk_moves.reserve(cap * 4) || !b->walk_early.reserve(cap * 4) ||
        !b->deferred.reserve(cap * 4) || !b->deferred_count.reserve(4) || !b->deferred_held.reserve(cap * 4) ||
        !b->arena_used.reserve(8) || !b->edit_stats.reserve((2 * EDIT_BINS + 2) * 4) ||
        !b->edit_work.reserve(2 * EDIT_BINS * 8))
        return DINARA_OUT_OF_MEMORY;
    return 0;
}

EditArgs edit_args(dinara_batch *b) {
    EditArgs e{};
    e.s = sweep_args(b);
    // The certificate and the cap in unit terms: every cost is `unit_scale` times the unit cost.
    e.s.costs = CostParams{1, 0, 1, 0, 1, 1, 1, 1, 0, 1, 0, 1, 1, 1};
    e.s.cap = b->options.max_cost >= 0 ? b->options.max_cost / b->unit_scale : INT64_MAX / 4;
    e.scale = b->unit_scale;
    e.capacity = (int)std::max<size_t>(b->slot_pair.size(), 1);
    e.band_low = static_cast<int32_t *>(b->edit_band_low.ptr);
    e.first_band_low = static_cast<const int32_t *>(b->edit_first_low_d.ptr);
    e.lists = static_cast<int32_t *>(b->edit_lists.ptr);
    e.counts = static_cast<int32_t *>(b->edit_counts.ptr);
    e.file_lists = static_cast<int32_t *>(b->edit_retry_lists.ptr);
    e.file_counts = static_cast<int32_t *>(b->edit_retry_counts.ptr);
    e.spare_counts = static_cast<int32_t *>(b->edit_later_counts.ptr);
    e.first_lists = static_cast<int32_t *>(b->edit_first_lists.ptr);
    e.first_counts = static_cast<const int32_t *>(b->edit_first_counts.ptr);
    e.first_counts_out = static_cast<int32_t *>(b->edit_first_counts.ptr);
    e.first_low_out = static_cast<int32_t *>(b->edit_first_low_d.ptr);
    e.trace_lists = static_cast<int32_t *>(b->trace_lists.ptr);
    e.trace_counts = static_cast<int32_t *>(b->trace_counts.ptr);
    e.trace_offset = static_cast<int64_t *>(b->trace_offset.ptr);
    e.walk_ready = static_cast<int32_t *>(b->walk_ready.ptr);
    e.walk_moves = static_cast<int32_t *>(b->walk_moves.ptr);
    e.walk_early = static_cast<int32_t *>(b->walk_early.ptr);
    e.deferred = static_cast<int32_t *>(b->deferred.ptr);
    e.deferred_count = static_cast<int32_t *>(b->deferred_count.ptr);
    e.arena = static_cast<uint8_t *>(const_cast<void *>(b->arena_ptr));
    e.arena_used = static_cast<unsigned long long *>(b->arena_used.ptr);
    e.arena_capacity = b->arena_capacity;
    e.stats = static_cast<uint32_t *>(b->edit_stats.ptr);
    e.work = static_cast<unsigned long long *>(b->edit_work.ptr);
    e.one_pair = b->persistent ? 0 : 1;
    e.keep_first = b->keep_first && b->kind == DINARA_BATCH_ALIGNMENTS ? 1 : 0;
    return e;
}

// The slots walked a wave a pair: those whose longer side is at least EDIT_WALK_LONG letters, the first ones (slots
// are longest first).
int long_walks(const dinara_batch *b) {
    int count = 0;
    while (count < (int)b->rows.size() && std::max(b->rows[count], b->cols[count]) >= EDIT_WALK_LONG) ++count;
    return count;
}

// The walk of every slot ready for it: the long slots on `aside` and the others on `stream` at once (forked and joined
// by marks `mark` and `mark` + 1; `mark` + 2 marks both ended; one after the other when `aside` is `stream`), then
// the CIGAR text on `stream`. `aside` is the
// narrowest bin's stream, which no retry uses (retries go to wider bins): a stream of its own would be one more than
// the hardware queues bench.py asks for, and a stream sharing a queue runs after the other's work.
void walk_slots(dinara_batch *b, const EditArgs &a, hipStream_t stream, hipStream_t aside, int mark, bool early) {
    DeviceContext *ctx = b->ctx;
    const int slots = (int)b->slot_pair.size(), longs = long_walks(b);
    if (longs > 0 && aside == stream) {
        launch_edit_walk_long(a, longs, stream, early);
    } else if (longs > 0) {
        (void)hipEventRecord(ctx->mark(mark), stream);
        (void)hipStreamWaitEvent(aside, ctx->mark(mark), 0);
        launch_edit_walk_long(a, longs, aside, early);
        (void)hipEventRecord(ctx->mark(mark + 1), aside);
    }
    launch_edit_walk_short(a, longs, slots, stream, early);
    if (longs > 0 && aside != stream) (void)hipStreamWaitEvent(stream, ctx->mark(mark + 1), 0);
    (void)hipEventRecord(ctx->mark(mark + 2), stream);
    launch_edit_text(a, slots, stream, early);
}

// Workgroups of a bin's launch that may meet `reach` pairs: enough for every one at once, at most `grid_per_cu` a
// compute unit when the groups refill from the queue; 0 when it meets none.
int edit_grid(const dinara_batch *b, int bin, int64_t reach) {
    if (reach == 0) return 0;
    const int64_t every = (reach * edit_lanes(bin) + 255) / 256;
    if (!b->persistent) return (int)std::max<int64_t>(1, every);
    return (int)std::max<int64_t>(1, std::min<int64_t>(every, (int64_t)b->ctx->compute_units * b->grid_per_cu));
}

// The arena the traced bands' decisions take: the batch's bound (no traced band is wider than its whole matrix),
// at most 0.6 of the free memory, at least 1 MiB. DINARA_TRACE_ARENA sets it in bytes.
size_t arena_bytes(const dinara_batch *b) {
    if (const char *env = std::getenv("DINARA_TRACE_ARENA")) return std::max<size_t>(std::strtoull(env, nullptr, 10), 4096);
    size_t bound = 0;
    for (size_t s = 0; s < b->slot_pair.size(); ++s) {
        const int n = b->rows[s], m = b->cols[s];
        const int bin = edit_widest(n + m);
        bound += (size_t)edit_trace_bytes(m, bin);
    }
    size_t free_bytes = 0, total = 0;
    (void)hipMemGetInfo(&free_bytes, &total);
    if (b->ctx->edit_arena) free_bytes += b->ctx->edit_arena->size;  // the arena itself is there to be used
    const size_t most = free_bytes / 10 * 6;
    if (b->keep_first) bound *= 3;  // every attempt keeps its decisions
    return std::max<size_t>(std::min(bound, most), size_t(1) << 20);
}

// Timing events of a run launched kernel by kernel: a pair for each bin's first sweep, retry sweep and tracing
// sweep, and for the walk.
constexpr int EV_FIRST = 0, EV_RETRY = 2 * EDIT_BINS, EV_TRACE = 4 * EDIT_BINS, EV_WALK = 6 * EDIT_BINS,
              EV_LATER = 6 * EDIT_BINS + 2;
// Marks (untimed events) of the early walk's fork and join; the bins side by side use marks 0 .. 4 * EDIT_BINS + 2.
constexpr int EARLY_WALK = 5 * EDIT_BINS;
// Marks of the walks' forks and joins (the long slots' launch beside the others'): the early walk's three, the last
// walk's three; then the queue counts' read-back.
constexpr int WALK_MARKS = 5 * EDIT_BINS + 2;
constexpr int COUNTS_MARK = WALK_MARKS + 6;

// Runs `launch(bin, stream)` for each bin `use(bin)` picks, every one on a stream of its own forked from the main
// stream and joined back to it (`mark_base` picks the ordering events), each between timing events EV + 2 bin and
// EV + 2 bin + 1 when `timed`. Side by side, the widest bin is launched first: its long pairs are the run's
// longest chains, and the hardware queues take the streams in order.
template <typename Use, typename Launch>
void side_by_side(DeviceContext *ctx, int mark_base, bool timed, int ev, bool apart, Use use, Launch launch) {
    if (apart) (void)hipEventRecord(ctx->mark(mark_base), ctx->stream);
    for (int k = 0; k < EDIT_BINS; ++k) {
        const int bin = apart ? EDIT_BINS - 1 - k : k;
        if (!use(bin)) continue;
        if (!apart) {  // one after another on the main stream
            if (timed) (void)hipEventRecord(ctx->event(ev + 2 * bin), ctx->stream);
            launch(bin, ctx->stream);
            if (timed) (void)hipEventRecord(ctx->event(ev + 2 * bin + 1), ctx->stream);
            continue;
        }
        hipStream_t side = ctx->side(bin);
        (void)hipStreamWaitEvent(side, ctx->mark(mark_base), 0);
        if (timed) (void)hipEventRecord(ctx->event(ev + 2 * bin), side);
        launch(bin, side);
        if (timed) (void)hipEventRecord(ctx->event(ev + 2 * bin + 1), side);
        (void)hipEventRecord(ctx->mark(mark_base + 1 + bin), side);
        (void)hipStreamWaitEvent(ctx->stream, ctx->mark(mark_base + 1 + bin), 0);
    }
}

// The run's launches: reset; every bin's first queue, the bins side by side; the retry queues, narrowest bin
// first; then (alignments) every bin's tracing sweep side by side, and the walk. `timed` records an event pair
// around each launch (not under capture); `may_wait` lets the host read queue counts back between launches.
int64_t enqueue_unit(dinara_batch *b, const EditArgs &a, bool timed, bool trace, bool may_wait) {
    DeviceContext *ctx = b->ctx;
    const int slots = (int)b->slot_pair.size();
    launch_edit_reset(a, slots, ctx->stream);
    side_by_side(ctx, 0, timed, EV_FIRST, b->side_by_side_bins, [&](int bin) { return b->first_count[bin] > 0; },
                 [&](int bin, hipStream_t st) {
                     launch_edit_sweep(a, bin, false, edit_grid(b, bin, b->first_count[bin]), st);
                 });
    // Kept decisions: the pairs the first pass proved are walked on a stream of their own while the retries sweep.
    const bool early = trace && b->keep_first;
    if (early) {
        launch_edit_snapshot(a, slots, ctx->stream);
        (void)hipEventRecord(ctx->mark(EARLY_WALK), ctx->stream);
        hipStream_t walker = ctx->side(EDIT_BINS);
        (void)hipStreamWaitEvent(walker, ctx->mark(EARLY_WALK), 0);
        // The early walk's two launches one after the other: side by side they slow the retries beside them,
        // whose long pairs are the run's tail.
        walk_slots(b, a, walker, walker, WALK_MARKS, true);
        (void)hipEventRecord(ctx->mark(EARLY_WALK + 1), walker);
    }
    // The first retries side by side, filing into the later set; the later retries one bin after another.
    EditArgs retry = a;
    retry.lists = a.file_lists;
    retry.counts = a.file_counts;
    retry.file_lists = static_cast<int32_t *>(b->edit_later_lists.ptr);
    retry.file_counts = static_cast<int32_t *>(b->edit_later_counts.ptr);
    side_by_side(ctx, 3 * EDIT_BINS + 2, timed, EV_RETRY, b->side_by_side_bins,
                 [&](int bin) { return b->retry_reach[bin] > 0; },
                 [&](int bin, hipStream_t st) {
                     launch_edit_sweep(retry, bin, false, edit_grid(b, bin, b->retry_reach[bin]), st);
                 });
    EditArgs later = retry;
    later.lists = retry.file_lists;
    later.counts = retry.file_counts;
    // Launched one by one (`may_wait`, not captured): the later retries begin at the narrowest bin whose queue holds
    // a pair once the retries have ended (a later retry files only into wider bins), and a tracing sweep is made only
    // for a queue that holds a pair; the counts are read back, the main stream waiting (the early walk goes on). An
    // empty launch costs about 0.08 ms of the main stream. A launch not made has its timing events back to back.
    // Launched one by one (`may_wait`, not captured), the later retries begin at the narrowest bin whose queue holds
    // a pair once the retries have ended (a later retry files only into wider bins), and a tracing sweep is made only
    // for a queue that holds a pair: one read brings both sets of counts, the tracing queues' copied beside the
    // later retries' first (the later-retry counts' buffer has room past its lengths and heads). An empty launch
    // costs about 0.08 ms of the main stream. A launch not made has its timing events back to back.
    //
    // Alignments launched one by one walk the pairs ready once the retries end at once, while the counts come back
    // (pinned, waited for on an event): the host's round trip, about 0.6 ms, is then hidden under that walk. Only
    // when a later retry or a tracing sweep runs does a second walk follow them.
    int32_t pending[3 * EDIT_BINS] = {};
    int32_t *const trace_pending = pending + 2 * EDIT_BINS;
    int from = 0;
    bool walked = false;
    if (may_wait) {
        int32_t *later_counts = static_cast<int32_t *>(b->edit_later_counts.ptr);
        if (trace && hipMemcpyAsync(later_counts + 2 * EDIT_BINS, b->trace_counts.ptr, EDIT_BINS * 4,
                                    hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess)
            return DINARA_OUT_OF_MEMORY;
        if (trace && !b->pinned_counts &&
            hipHostMalloc(reinterpret_cast<void **>(&b->pinned_counts), sizeof(pending)) != hipSuccess)
            b->pinned_counts = nullptr;
        if (trace && b->pinned_counts) {
            if (hipMemcpyAsync(b->pinned_counts, later_counts, sizeof(pending), hipMemcpyDeviceToHost, ctx->stream) !=
                    hipSuccess ||
                hipEventRecord(ctx->mark(COUNTS_MARK), ctx->stream) != hipSuccess)
                return DINARA_OUT_OF_MEMORY;
            if (early) (void)hipStreamWaitEvent(ctx->stream, ctx->mark(WALK_MARKS + 2), 0);
            if (timed) (void)hipEventRecord(ctx->event(EV_WALK), ctx->stream);
            walk_slots(b, a, ctx->stream, ctx->side(0), WALK_MARKS + 3, false);
            walked = true;
            if (hipEventSynchronize(ctx->mark(COUNTS_MARK)) != hipSuccess) return DINARA_OUT_OF_MEMORY;
            std::memcpy(pending, b->pinned_counts, sizeof(pending));
        } else if (!b->edit_later_counts.read(ctx, pending, sizeof(pending))) {
            return DINARA_OUT_OF_MEMORY;
        }
        while (from < EDIT_BINS && pending[from] == 0) ++from;
    }
    for (int bin = 0; bin < EDIT_BINS; ++bin) {
        const int grid = edit_grid(b, bin, b->retry_reach[bin]);
        if (!grid) continue;
        if (timed) (void)hipEventRecord(ctx->event(EV_LATER + 2 * bin), ctx->stream);
        if (bin >= from) launch_edit_sweep(later, bin, false, grid, ctx->stream);
        if (timed) (void)hipEventRecord(ctx->event(EV_LATER + 2 * bin + 1), ctx->stream);
    }
    if (trace) {
        if (may_wait && from < EDIT_BINS && !b->trace_counts.read(ctx, trace_pending, EDIT_BINS * 4))
            return DINARA_OUT_OF_MEMORY;
        const auto traced = [&](int bin) { return b->trace_reach[bin] > 0 && (!may_wait || trace_pending[bin] > 0); };
        bool any_traced = false;
        for (int bin = 0; bin < EDIT_BINS; ++bin) any_traced = any_traced || traced(bin);
        if (timed)
            for (int bin = 0; bin < EDIT_BINS; ++bin)
                if (b->trace_reach[bin] > 0 && !traced(bin)) {
                    (void)hipEventRecord(ctx->event(EV_TRACE + 2 * bin), ctx->stream);
                    (void)hipEventRecord(ctx->event(EV_TRACE + 2 * bin + 1), ctx->stream);
                }
        side_by_side(ctx, EDIT_BINS + 1, timed, EV_TRACE, b->side_by_side_bins, traced,
                     [&](int bin, hipStream_t st) {
                         launch_edit_sweep(a, bin, true, edit_grid(b, bin, b->trace_reach[bin]), st);
                     });
        // The last walk follows the early walk's kernels (it writes moves of slots the early walk marked unwalked)
        // and runs while the early walk's text is made; the run ends when both texts are made. After a walk already
        // made, a second one only for what later retries and tracing sweeps proved (it marks the first's slots
        // unwalked, whose text the first made before it on this stream).
        if (!walked || from < EDIT_BINS || any_traced) {
            if (early && !walked) (void)hipStreamWaitEvent(ctx->stream, ctx->mark(WALK_MARKS + 2), 0);
            if (timed && !walked) (void)hipEventRecord(ctx->event(EV_WALK), ctx->stream);
            walk_slots(b, a, ctx->stream, ctx->side(0), WALK_MARKS + 3, false);
        }
        if (timed) (void)hipEventRecord(ctx->event(EV_WALK + 1), ctx->stream);
        if (early) (void)hipStreamWaitEvent(ctx->stream, ctx->mark(EARLY_WALK + 1), 0);
    }
    return hipPeekAtLastError() == hipSuccess ? 0 : DINARA_OUT_OF_MEMORY;
}

// The unit-cost run: one captured graph (or the same launches one by one, timed), then further tracing passes for
// pairs the arena had no room for. `trace` false skips the tracing (the autotuner's runs).
int64_t run_unit(dinara_batch *b, bool trace) {
    DeviceContext *ctx = b->ctx;
    const int64_t slots = (int64_t)b->slot_pair.size();
    trace = trace && b->kind == DINARA_BATCH_ALIGNMENTS;
    b->launches = b->first_proven = b->graph_used = b->trace_passes = 0;
    b->run_ms = b->sweep_ms = b->walk_ms = 0;
    for (int bin = 0; bin < EDIT_BINS; ++bin) b->edit_ms[bin] = b->trace_ms[bin] = b->edit_sweeps[bin] = b->trace_sweeps[bin] = 0;
    if (slots == 0) {
        b->ran = true;
        return 0;
    }
    if (b->kind == DINARA_BATCH_ALIGNMENTS) {
        // The context's arena, grown when this batch wants more than it holds; the graph is made again when the
        // arena it was captured with moved or this batch's share of it changed.
        if (!ctx->edit_arena) ctx->edit_arena = new DeviceBuffer();
        if (!b->arena_want) b->arena_want = arena_bytes(b);
        const size_t bytes = b->arena_want;
        if (ctx->edit_arena->size < bytes) {
            ctx->edit_arena->release();
            if (!ctx->edit_arena->reserve(bytes)) return DINARA_OUT_OF_MEMORY;
        }
        if (b->arena_ptr != ctx->edit_arena->ptr || b->arena_capacity != bytes) {
            b->arena_ptr = ctx->edit_arena->ptr;
            b->arena_capacity = bytes;
            if (b->edit_graph) {
                (void)hipGraphExecDestroy(b->edit_graph);
                b->edit_graph = nullptr;
            }
        }
    }
    const EditArgs a = edit_args(b);
    const bool graph = b->use_graph && trace == (b->kind == DINARA_BATCH_ALIGNMENTS);
    if (graph) {
        if (!b->edit_graph) {
            hipGraph_t captured = nullptr;
            if (hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
                return DINARA_OUT_OF_MEMORY;
            const int64_t code = enqueue_unit(b, a, false, trace, false);
            const hipError_t ended = hipStreamEndCapture(ctx->stream, &captured);
            if (code || ended != hipSuccess) return DINARA_OUT_OF_MEMORY;
            const hipError_t made = hipGraphInstantiate(&b->edit_graph, captured, nullptr, nullptr, 0);
            (void)hipGraphDestroy(captured);
            if (made != hipSuccess) {
                b->edit_graph = nullptr;
                return DINARA_OUT_OF_MEMORY;
            }
        }
        if (hipEventRecord(ctx->start, ctx->stream) != hipSuccess ||
            hipGraphLaunch(b->edit_graph, ctx->stream) != hipSuccess)
            return DINARA_OUT_OF_MEMORY;
        b->graph_used = 1;
    } else {
        if (hipEventRecord(ctx->start, ctx->stream) != hipSuccess) return DINARA_OUT_OF_MEMORY;
        if (const int64_t code = enqueue_unit(b, a, b->timed, trace, true)) return code;
    }
    if (hipEventRecord(ctx->sweep_stop, ctx->stream) != hipSuccess) return DINARA_OUT_OF_MEMORY;
    // Pairs whose decisions found no room: trace them in further passes, the arena emptied each time.
    int32_t deferred = 0;
    if (trace) {
        if (!b->deferred_count.read(ctx, &deferred, 4)) return DINARA_OUT_OF_MEMORY;
        while (deferred > 0) {
            b->trace_passes += 1;
            if (hipMemcpyAsync(b->deferred_held.ptr, b->deferred.ptr, (size_t)deferred * 4, hipMemcpyDeviceToDevice,
                               ctx->stream) != hipSuccess)
                return DINARA_OUT_OF_MEMORY;
            launch_edit_requeue(a, static_cast<const int32_t *>(b->deferred_held.ptr), deferred, ctx->stream);
            for (int bin = 0; bin < EDIT_BINS; ++bin)
                if (const int grid = edit_grid(b, bin, b->trace_reach[bin]))
                    launch_edit_sweep(a, bin, true, grid, ctx->stream);
            walk_slots(b, a, ctx->stream, ctx->side(0), WALK_MARKS + 3, false);
            int32_t left = 0;
            if (hipPeekAtLastError() != hipSuccess || !b->deferred_count.read(ctx, &left, 4)) return DINARA_OUT_OF_MEMORY;
            if (left >= deferred) return DINARA_OUT_OF_MEMORY;  // a pair larger than the whole arena
            deferred = left;
        }
    }
    if (hipEventRecord(ctx->stop, ctx->stream) != hipSuccess || hipEventSynchronize(ctx->stop) != hipSuccess)
        return DINARA_OUT_OF_MEMORY;
    float total = 0, sweeps = 0;
    (void)hipEventElapsedTime(&total, ctx->start, ctx->stop);
    (void)hipEventElapsedTime(&sweeps, ctx->start, ctx->sweep_stop);
    b->run_ms = total;
    b->sweep_ms = sweeps;
    // The device counters (sweeps, word-columns, first bands proven) are read back only when asked for: a
    // short batch's run is then one graph launch and one wait.
    b->stats_pending = true;
    b->stats_slots = slots;
    for (int bin = 0; bin < EDIT_BINS; ++bin) {
        if (!graph && b->timed) {
            float ms = 0;
            if (b->first_count[bin] > 0 &&
                hipEventElapsedTime(&ms, ctx->event(EV_FIRST + 2 * bin), ctx->event(EV_FIRST + 2 * bin + 1)) == hipSuccess)
                b->edit_ms[bin] += ms;
            if (b->retry_reach[bin] > 0 &&
                hipEventElapsedTime(&ms, ctx->event(EV_RETRY + 2 * bin), ctx->event(EV_RETRY + 2 * bin + 1)) == hipSuccess)
                b->edit_ms[bin] += ms;
            if (b->retry_reach[bin] > 0 &&
                hipEventElapsedTime(&ms, ctx->event(EV_LATER + 2 * bin), ctx->event(EV_LATER + 2 * bin + 1)) == hipSuccess)
                b->edit_ms[bin] += ms;
            if (trace && b->trace_reach[bin] > 0 &&
                hipEventElapsedTime(&ms, ctx->event(EV_TRACE + 2 * bin), ctx->event(EV_TRACE + 2 * bin + 1)) == hipSuccess)
                b->trace_ms[bin] = ms;
        }
    }
    if (!graph && trace && b->timed) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, ctx->event(EV_WALK), ctx->event(EV_WALK + 1));
        b->walk_ms = ms;
    }
    b->launches = graph ? 1 : 1 + EDIT_BINS * (trace ? 2 : 1) + (trace ? 1 : 0);
    b->ran = true;
    return 0;
}

// Sweeps every slot of the first lists until each is settled or left unsettled: want bands from first_low_d /
// first_high_d, round 0 the first lists, each later round the lists the round before filed, kept on the device,
// two sets taking turns. Only the counts come back to the host. Caller holds the device lock.
int64_t sweep_rounds(dinara_batch *b, size_t &arena_index) {
    DeviceContext *ctx = b->ctx;
    const int64_t slots = (int64_t)b->slot_pair.size();
    if (slots == 0) return 0;
    if (hipMemcpyAsync(b->want_low.ptr, b->first_low_d.ptr, slots * 4, hipMemcpyDeviceToDevice, ctx->stream) !=
            hipSuccess ||
        hipMemcpyAsync(b->want_high.ptr, b->first_high_d.ptr, slots * 4, hipMemcpyDeviceToDevice, ctx->stream) !=
            hipSuccess ||
        hipMemsetAsync(b->status.ptr, 0xFF, slots * 4, ctx->stream) != hipSuccess)
        return DINARA_OUT_OF_MEMORY;
    int64_t counts[CLASSES] = {};
    for (int c = 0; c < CLASSES; ++c) counts[c] = b->first_offset[c + 1] - b->first_offset[c];
    int32_t *lists = static_cast<int32_t *>(b->retry_lists.ptr);
    int32_t *count_d = static_cast<int32_t *>(b->retry_counts.ptr);
    const size_t capacity = (size_t)std::max<int64_t>(slots, 1);
    for (int round = 0; round < 64; ++round) {
        int64_t pending = 0;
        for (int c = 0; c < CLASSES; ++c) pending += counts[c];
        if (pending == 0) break;
        int32_t *out_lists = lists + (size_t)(round % 2) * CLASSES * capacity;
        int32_t *out_counts = count_d + (round % 2) * CLASSES;
        const int32_t *in_lists = lists + (size_t)((round + 1) % 2) * CLASSES * capacity;
        if (hipMemsetAsync(out_counts, 0, CLASSES * 4, ctx->stream) != hipSuccess) return DINARA_OUT_OF_MEMORY;
        (void)hipEventRecord(ctx->sweep_start, ctx->stream);
        for (int c = 0; c < CLASSES; ++c) {
            if (counts[c] == 0) continue;
            const int32_t *list = round == 0 ? static_cast<const int32_t *>(b->first_lists_d.ptr) + b->first_offset[c]
                                             : in_lists + (size_t)c * capacity;
            // Each class's launches between two events of their own, for the per-kernel time.
            (void)hipEventRecord(ctx->event(2 * c), ctx->stream);
            if (const int64_t code = launch_class(b, c, list, counts[c], out_lists, out_counts, arena_index))
                return code;
            (void)hipEventRecord(ctx->event(2 * c + 1), ctx->stream);
        }
        (void)hipEventRecord(ctx->sweep_stop, ctx->stream);
        int32_t filed[CLASSES];
        if (hipMemcpyAsync(filed, out_counts, sizeof(filed), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess)
            return DINARA_OUT_OF_MEMORY;
        float ms = 0;
        (void)hipEventElapsedTime(&ms, ctx->sweep_start, ctx->sweep_stop);
        b->sweep_ms += ms;
        int64_t retried = 0;
        for (int c = 0; c < CLASSES; ++c) {
            if (counts[c] > 0) {
                float class_ms = 0;
                (void)hipEventElapsedTime(&class_ms, ctx->event(2 * c), ctx->event(2 * c + 1));
                b->class_ms[c] += class_ms;
            }
            counts[c] = filed[c];
            retried += filed[c];
        }
        if (round == 0 && b->phase <= 1) b->first_proven = (double)(slots - retried);
    }
    return 0;
}

// The forward phase's views, free letters, bands and lists onto the device. Caller holds the device lock.
int64_t free_phase_a(dinara_batch *b) {
    DeviceContext *ctx = b->ctx;
    const size_t slots = b->slot_pair.size();
    if (slots == 0) return 0;
    b->phase = 1;
    b->first_offset = b->a_first_offset;
    if (!b->view_rows_d.write(ctx, b->a_rows.data(), slots * 4) || !b->view_cols_d.write(ctx, b->a_cols.data(), slots * 4) ||
        !b->view_ref_at_d.write(ctx, b->a_at.data(), slots * 4) || !b->view_qry_at_d.write(ctx, b->a_at.data(), slots * 4) ||
        !b->free_d.write(ctx, b->a_free.data(), slots * 16) ||
        !b->first_low_d.write(ctx, b->a_first_low.data(), slots * 4) ||
        !b->first_high_d.write(ctx, b->a_first_high.data(), slots * 4) ||
        !b->first_lists_d.upload(ctx, b->a_first_all.data(), b->a_first_all.size() * 4))
        return DINARA_OUT_OF_MEMORY;
    return 0;
}

// The anchored phase of free-end alignments, from the forward phase's results: each settled pair (within the cap)
// sweeps the letters before its end, back from the end (no free start), free to end where the forward phase was
// free to start, its band the one its known cost proves; a pair it settles is walked from the end it finds.
// Caller holds the device lock.
int64_t free_phase_b(dinara_batch *b) {
    DeviceContext *ctx = b->ctx;
    const int64_t slots = (int64_t)b->slot_pair.size();
    b->a_status_h.resize(slots);
    b->a_cost_h.resize(slots);
    b->a_end_row_h.resize(slots);
    b->a_end_col_h.resize(slots);
    if (!b->status.read(ctx, b->a_status_h.data(), slots * 4) || !b->cost.read(ctx, b->a_cost_h.data(), slots * 4) ||
        !b->end_row_d.read(ctx, b->a_end_row_h.data(), slots * 4) ||
        !b->end_col_d.read(ctx, b->a_end_col_h.data(), slots * 4))
        return DINARA_OUT_OF_MEMORY;
    const CostParams &c = b->costs;
    auto gap = [&](int64_t k, int open, int ext, int open2, int ext2) -> int64_t {
        if (k <= 0) return 0;
        const int64_t one = (int64_t)open + k * ext;
        return c.pieces == 2 ? std::min(one, (int64_t)open2 + k * ext2) : one;
    };
    const bool right = b->options.right_ties;
    const int64_t cap = b->options.max_cost >= 0 ? b->options.max_cost : INT64_MAX / 4;
    std::vector<int32_t> rows(slots), cols(slots), ref_at(slots), qry_at(slots), fr((size_t)slots * 4, 0);
    b->b_rows_h.assign(slots, 0);
    b->b_cols_h.assign(slots, 0);
    for (auto &l : b->first_lists) l.clear();
    for (int64_t s = 0; s < slots; ++s) {
        const int32_t n = b->rows[s], m = b->cols[s];
        const int32_t re = b->a_end_row_h[s], ce = b->a_end_col_h[s];
        const bool go = b->a_status_h[s] == 0 && b->a_cost_h[s] <= cap;
        rows[s] = go ? re : 0;
        cols[s] = go ? ce : 0;
        ref_at[s] = go && right ? n - re : 0;
        qry_at[s] = go && right ? m - ce : 0;
        b->b_rows_h[s] = rows[s];
        b->b_cols_h[s] = cols[s];
        if (!go) continue;
        // Free to end where the forward phase was free to start.
        const int32_t er = std::min<int32_t>(b->a_free[4 * s], re), ec = std::min<int32_t>(b->a_free[4 * s + 2], ce);
        fr[4 * s + 1] = er;
        fr[4 * s + 3] = ec;
        const int64_t least = b->a_cost_h[s], target = least + 1, end = (int64_t)ce - re;
        auto off = [&](int64_t k) -> int64_t {  // off_band_ends with no free start
            const int64_t high = std::max<int64_t>(0, end + er), low = std::min<int64_t>(0, end - ec);
            if (k > high) return gap(k, c.ins_open, c.ins_ext, c.ins_open2, c.ins_ext2) +
                                 gap(k - (end + er), c.del_open, c.del_ext, c.del_open2, c.del_ext2);
            if (k < low) return gap(-k, c.del_open, c.del_ext, c.del_open2, c.del_ext2) +
                                gap(end - ec - k, c.ins_open, c.ins_ext, c.ins_open2, c.ins_ext2);
            return 0;
        };
        int64_t h = std::max<int64_t>(0, end + er), l = std::min<int64_t>(0, end - ec);
        while (h < ce && off(h + 1) < target) ++h;
        while (l > -re && off(l - 1) < target) --l;
        b->first_low[s] = (int32_t)l;
        b->first_high[s] = (int32_t)h;
        const int cls = first_class_for((int)(h - l + 1), false, 0, false);
        if (cls != CLASS_TOO_WIDE) b->first_lists[cls].push_back((int32_t)s);
    }
    std::vector<int32_t> first_all;
    b->first_offset.assign(CLASSES + 1, 0);
    for (int k = 0; k < CLASSES; ++k) {
        b->first_offset[k] = (int64_t)first_all.size();
        first_all.insert(first_all.end(), b->first_lists[k].begin(), b->first_lists[k].end());
    }
    b->first_offset[CLASSES] = (int64_t)first_all.size();
    if (first_all.empty()) first_all.push_back(0);
    if (!b->view_rows_d.write(ctx, rows.data(), slots * 4) || !b->view_cols_d.write(ctx, cols.data(), slots * 4) ||
        !b->view_ref_at_d.write(ctx, ref_at.data(), slots * 4) || !b->view_qry_at_d.write(ctx, qry_at.data(), slots * 4) ||
        !b->free_d.write(ctx, fr.data(), slots * 16) || !b->first_low_d.write(ctx, b->first_low.data(), slots * 4) ||
        !b->first_high_d.write(ctx, b->first_high.data(), slots * 4) ||
        !b->first_lists_d.upload(ctx, first_all.data(), first_all.size() * 4))
        return DINARA_OUT_OF_MEMORY;
    b->phase = 2;
    return 0;
}

// The decision bytes one pair may take on the whole-matrix path: past it the pair stays unsettled.
constexpr size_t FULL_BUDGET = size_t(16) << 30;

// After a run, every pair the band sweeps left unsettled (wider than the widest band, or too many letters for the
// kernels that stage them in LDS) is swept whole, one pair after another, and walked (full.hip); a pair under free
// ends or a user band, or whose decisions pass FULL_BUDGET, stays unsettled. A batch none of
// whose pairs could need it skips the pass; otherwise the unsettled slots are listed on the device and only their
// count (and list) come back. Caller holds the device lock.
int64_t full_pass(dinara_batch *b) {
    DeviceContext *ctx = b->ctx;
    const int64_t slots = (int64_t)b->slot_pair.size();
    b->full_pairs = b->full_ms = 0;
    if (b->free_ends || b->user_band || slots == 0) return 0;
    if (b->may_need_full < 0) {
        // Under both the unit bins and the band classes a band of 8192 diagonals proves any pair of fewer letters,
        // and such a pair fits every kernel's LDS.
        b->may_need_full = 0;
        for (int64_t s = 0; s < slots; ++s)
            if ((int64_t)b->rows[s] + b->cols[s] + 3 > 8192) b->may_need_full = 1;
    }
    if (!b->may_need_full && !b->whole) return 0;
    const auto began = std::chrono::steady_clock::now();
    if (!b->full_list.reserve((size_t)slots * 4) || !b->full_count.reserve(4) || !b->full_claim.reserve(8) ||
        hipMemsetAsync(b->full_count.ptr, 0, 4, ctx->stream) != hipSuccess)
        return DINARA_OUT_OF_MEMORY;
    launch_full_collect(static_cast<const int32_t *>(b->status.ptr), (int)slots,
                        static_cast<int32_t *>(b->full_list.ptr), static_cast<int32_t *>(b->full_count.ptr),
                        ctx->stream);
    int32_t count = 0;
    if (hipPeekAtLastError() != hipSuccess || !b->full_count.read(ctx, &count, 4)) return DINARA_OUT_OF_MEMORY;
    if (count == 0) return 0;
    std::vector<int32_t> list(count);
    if (!b->full_list.read(ctx, list.data(), (size_t)count * 4)) return DINARA_OUT_OF_MEMORY;
    std::sort(list.begin(), list.end());
    const bool record = b->kind == DINARA_BATCH_ALIGNMENTS;
    const int pieces = b->costs.pieces == 2 ? 2 : 1;
    const size_t record_bytes = pieces == 2 ? 8 : 4;
    // The listed pairs in launches of as many as fit FULL_BUDGET together (a pair past it alone stays unsettled).
    size_t next = 0;
    bool launched = false;
    while (next < list.size()) {
        // The buffers below are refilled (and may move) for each launch: the one before must be done with them.
        if (launched && hipStreamSynchronize(ctx->stream) != hipSuccess) return DINARA_OUT_OF_MEMORY;
        std::vector<int32_t> pair_slot, strip_first;
        std::vector<int64_t> decision_at, carry_at;
        int64_t strips = 0, records = 0, carry_words = 0;
        for (; next < list.size(); ++next) {
            const int32_t slot = list[next];
            if (slot < 0 || slot >= slots) continue;
            const int64_t n = b->rows[slot], m = b->cols[slot];
            const int64_t own = std::max<int64_t>(1, (m + FULL_STRIP - 1) / FULL_STRIP);
            const int64_t own_records = record ? own * n * 64 : 0;
            if ((size_t)own_records * record_bytes > FULL_BUDGET) continue;
            if (!pair_slot.empty() && (size_t)(records + own_records) * record_bytes > FULL_BUDGET) break;
            pair_slot.push_back(slot);
            strip_first.push_back((int32_t)strips);
            decision_at.push_back(records);
            carry_at.push_back(carry_words);
            strips += own;
            records += own_records;
            carry_words += own * (n + 1) * pieces;
        }
        if (pair_slot.empty()) continue;
        strip_first.push_back((int32_t)strips);
        const int64_t count = (int64_t)pair_slot.size();
        if ((record && !b->full_decisions.reserve((size_t)records * record_bytes)) ||
            !b->full_carry.reserve((size_t)carry_words * 8) || !b->full_claim.reserve((size_t)(count + 1) * 4) ||
            !b->full_slots.upload(ctx, pair_slot.data(), (size_t)count * 4) ||
            !b->full_strips.upload(ctx, strip_first.data(), (size_t)(count + 1) * 4) ||
            !b->full_decision_at.upload(ctx, decision_at.data(), (size_t)count * 8) ||
            !b->full_carry_at.upload(ctx, carry_at.data(), (size_t)count * 8) ||
            hipMemsetAsync(b->full_carry.ptr, 0, (size_t)carry_words * 8, ctx->stream) != hipSuccess ||
            hipMemsetAsync(b->full_claim.ptr, 0, (size_t)(count + 1) * 4, ctx->stream) != hipSuccess)
            return DINARA_OUT_OF_MEMORY;
        FullArgs fa{};
        fa.s = sweep_args(b);
        fa.pairs = (int)count;
        fa.pair_slot = static_cast<const int32_t *>(b->full_slots.ptr);
        fa.strip_first = static_cast<const int32_t *>(b->full_strips.ptr);
        fa.decision_at = static_cast<const int64_t *>(b->full_decision_at.ptr);
        fa.carry_at = static_cast<const int64_t *>(b->full_carry_at.ptr);
        fa.record = record ? 1 : 0;
        fa.decisions = record ? b->full_decisions.ptr : nullptr;
        fa.carry = static_cast<unsigned long long *>(b->full_carry.ptr);
        fa.claim = static_cast<int *>(b->full_claim.ptr);
        fa.failed = static_cast<int *>(b->full_claim.ptr) + 1;
        launch_full_sweep(fa, (int)strips, ctx->stream);
        if (record) launch_full_walk(fa, ctx->stream);
        if (hipPeekAtLastError() != hipSuccess) return DINARA_OUT_OF_MEMORY;
        b->full_pairs += (double)count;
        launched = true;
    }
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return DINARA_OUT_OF_MEMORY;
    b->full_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - began).count();
    return 0;
}

int64_t run(dinara_batch *b) {
    DeviceContext *ctx = b->ctx;
    std::lock_guard<std::mutex> serial(ctx->lock);
    DeviceGuard guard(ctx);
    (void)hipGetLastError();  // an earlier, unrelated call's error is not this run's
    if (b->whole) {
        // No band sweep: every slot unsettled, then the whole-matrix pass takes them all.
        const int64_t slots = (int64_t)b->slot_pair.size();
        b->launches = b->first_proven = 0;
        b->run_ms = b->sweep_ms = b->walk_ms = 0;
        if (slots > 0 && (hipMemsetAsync(b->status.ptr, 0xFF, slots * 4, ctx->stream) != hipSuccess ||
                          (b->kind == DINARA_BATCH_ALIGNMENTS &&
                           hipMemsetAsync(b->cigar_used.ptr, 0, 8, ctx->stream) != hipSuccess)))
            return DINARA_OUT_OF_MEMORY;
        const int64_t code = full_pass(b);
        b->run_ms = b->full_ms;
        b->ran = code == 0;
        return code;
    }
    if (b->unit) {
        if (const int64_t code = run_unit(b, true)) return code;
        return full_pass(b);
    }
    const int64_t slots = (int64_t)b->slot_pair.size();
    b->launches = b->first_proven = 0;
    b->run_ms = b->sweep_ms = b->walk_ms = 0;
    for (int c = 0; c < CLASSES; ++c) b->class_ms[c] = b->class_launches[c] = b->class_pairs[c] = 0;
    if (hipEventRecord(ctx->start, ctx->stream) != hipSuccess) return DINARA_OUT_OF_MEMORY;
    if (slots > 0 && b->kind == DINARA_BATCH_ALIGNMENTS &&
        hipMemsetAsync(b->cigar_used.ptr, 0, 8, ctx->stream) != hipSuccess)
        return DINARA_OUT_OF_MEMORY;
    size_t arena_index = 0;
    if (b->free_ends) {
        // Free ends: the forward phase finds each pair's cost and end (reference tests/test_alignment.mojo:1983-2000);
        // alignments sweep again from that end, anchored, keeping decisions, for the start and the walk.
        if (const int64_t code = free_phase_a(b)) return code;
        b->phase = 1;
        if (const int64_t code = sweep_rounds(b, arena_index)) return code;
        if (b->kind == DINARA_BATCH_ALIGNMENTS && slots > 0) {
            if (const int64_t code = free_phase_b(b)) return code;
            b->phase = 2;
            if (const int64_t code = sweep_rounds(b, arena_index)) return code;
        }
    } else if (const int64_t code = sweep_rounds(b, arena_index)) {
        return code;
    }
    if (b->kind == DINARA_BATCH_ALIGNMENTS && slots > 0) {
        SweepArgs a = sweep_args(b);
        (void)hipEventRecord(ctx->event(2 * CLASSES), ctx->stream);
        launch_walk(a, ctx->stream);
        (void)hipEventRecord(ctx->event(2 * CLASSES + 1), ctx->stream);
    }
    if (hipEventRecord(ctx->stop, ctx->stream) != hipSuccess || hipEventSynchronize(ctx->stop) != hipSuccess)
        return DINARA_OUT_OF_MEMORY;
    if (b->kind == DINARA_BATCH_ALIGNMENTS && slots > 0) {
        float walk_ms = 0;
        (void)hipEventElapsedTime(&walk_ms, ctx->event(2 * CLASSES), ctx->event(2 * CLASSES + 1));
        b->walk_ms = walk_ms;
    }
    float total = 0;
    (void)hipEventElapsedTime(&total, ctx->start, ctx->stop);
    b->run_ms = total;
    b->ran = true;
    return full_pass(b);
}

int64_t fetch(dinara_batch *b) {
    DeviceContext *ctx = b->ctx;
    std::lock_guard<std::mutex> serial(ctx->lock);
    DeviceGuard guard(ctx);
    const int64_t slots = (int64_t)b->slot_pair.size();
    b->cost_h.resize(slots);
    b->status_h.resize(slots);
    if (slots == 0) return 0;
    if (!b->cost.read(ctx, b->cost_h.data(), slots * 4) || !b->status.read(ctx, b->status_h.data(), slots * 4))
        return DINARA_OUT_OF_MEMORY;
    if (b->free_ends && b->kind == DINARA_BATCH_ALIGNMENTS) {
        b->b_end_row_h.resize(slots);
        b->b_end_col_h.resize(slots);
        if (!b->end_row_d.read(ctx, b->b_end_row_h.data(), slots * 4) ||
            !b->end_col_d.read(ctx, b->b_end_col_h.data(), slots * 4))
            return DINARA_OUT_OF_MEMORY;
    }
    if (b->kind == DINARA_BATCH_ALIGNMENTS) {
        unsigned long long used = 0;
        b->cigar_offset_h.resize(slots);
        b->cigar_length_h.resize(slots);
        if (!b->cigar_used.read(ctx, &used, 8) || !b->cigar_offset.read(ctx, b->cigar_offset_h.data(), slots * 8) ||
            !b->cigar_length.read(ctx, b->cigar_length_h.data(), slots * 4))
            return DINARA_OUT_OF_MEMORY;
        b->cigars_h.resize(std::max<unsigned long long>(used, 1));
        if (used > 0 && !b->cigars.read(ctx, b->cigars_h.data(), used)) return DINARA_OUT_OF_MEMORY;
    }
    return 0;
}

// A pair's code when the device left it unsettled (a nonzero status): no result.
int64_t unsettled_code(int32_t) { return DINARA_OUT_OF_MEMORY; }

dinara_batch *batch_new(int64_t gpu, int64_t kind, int64_t pairs, const char *const *refs, const int64_t *ref_lens,
                        const char *const *qrys, const int64_t *qry_lens, const dinara_costs *costs,
                        const dinara_mode *mode, const dinara_options *options, int64_t threads, int64_t *status) {
    auto fail = [&](int64_t code) -> dinara_batch * {
        if (status) *status = code;
        return nullptr;
    };
    ParsedCosts parsed = parse_costs(costs);
    if (parsed.code) return fail(parsed.code);
    if (const int64_t code = check_mode(mode)) return fail(code);
    if (kind != DINARA_BATCH_DISTANCES && kind != DINARA_BATCH_ALIGNMENTS) return fail(DINARA_INVALID_MODE);
    DeviceContext *ctx = device_context(gpu);
    if (!ctx) return fail(DINARA_NO_DEVICE);
    auto *b = new dinara_batch();
    b->gpu = (int)gpu;
    b->ctx = ctx;
    b->kind = kind;
    b->pairs = std::max<int64_t>(pairs, 0);
    // Packing threads: as many as asked, but no more than one a half-million letters (a thread's start costs more
    // than packing fewer letters).
    {
        int64_t letters = 0;
        for (int64_t i = 0; i < b->pairs; ++i)
            letters += std::max<int64_t>(ref_lens[i], 0) + std::max<int64_t>(qry_lens[i], 0);
        threads = std::max<int64_t>(1, std::min<int64_t>(threads, letters >> 19));
    }
    b->threads = threads;
    b->parsed = parsed;
    b->options = parse_options(options);
    b->lds_limit = ctx->lds_limit;
    if (const char *w = std::getenv("DINARA_INITIAL_WIDTH")) b->initial_width = std::max(1, atoi(w));
    CostParams &c = b->costs;
    c.mismatch = (int)parsed.mismatch;
    c.del_open = (int)parsed.del_opening;
    c.del_ext = (int)parsed.del_extension;
    c.ins_open = (int)parsed.opening;
    c.ins_ext = (int)parsed.extension;
    c.del_first = c.del_open + c.del_ext;
    c.ins_first = c.ins_open + c.ins_ext;
    c.pieces = parsed.two_piece ? 2 : 1;
    c.del_open2 = (int)parsed.del_opening2;
    c.del_ext2 = (int)parsed.del_extension2;
    c.ins_open2 = (int)parsed.opening2;
    c.ins_ext2 = (int)parsed.extension2;
    c.del_first2 = c.del_open2 + c.del_ext2;
    c.ins_first2 = c.ins_open2 + c.ins_ext2;
    // The lanes' diagonals are the query's letters less the reference's: the library band mirrored
    // (lanes.mojo:1151-1155).
    const Options &o = b->options;
    b->user_band = o.band_low > -UNBOUNDED || o.band_high < UNBOUNDED;
    const int64_t clamp = int64_t(1) << 30;
    b->user_low = (int)std::max<int64_t>(-o.band_high, -clamp);
    b->user_high = (int)std::min<int64_t>(-o.band_low, clamp);
    // Unit costs times a scale, with no user band, take the bit-parallel path (DINARA_GOTOH_UNIT=1 keeps them on the
    // gap-cost sweeps, for comparison).
    b->free_ends = free_mode(mode);
    if (b->free_ends) {
        // Free ends under a user band are not served (the band would move with every view).
        if (b->user_band) {
            delete b;
            return fail(DINARA_INVALID_MODE);
        }
        b->ends[0] = mode->reference_start;
        b->ends[1] = mode->reference_end;
        b->ends[2] = mode->query_start;
        b->ends[3] = mode->query_end;
    }
    b->unit = !b->user_band && !b->free_ends && !parsed.two_piece && parsed.opening == 0 && parsed.del_opening == 0 &&
              parsed.extension == parsed.mismatch && parsed.del_extension == parsed.mismatch &&
              !(std::getenv("DINARA_GOTOH_UNIT") && std::atoi(std::getenv("DINARA_GOTOH_UNIT")) != 0);
    b->unit_scale = (int)parsed.mismatch;
    int64_t dearest = std::max<int64_t>(
        parsed.mismatch, std::max(parsed.opening + parsed.extension, parsed.del_opening + parsed.del_extension));
    if (parsed.two_piece)
        dearest = std::max(dearest, std::max(parsed.opening2 + parsed.extension2, parsed.del_opening2 + parsed.del_extension2));
    b->dearest = std::max<int64_t>(dearest, 1);
    b->code.assign(b->pairs, 0);
    b->reference_length.assign(b->pairs, 0);
    b->query_length.assign(b->pairs, 0);
    const bool capped = o.max_cost >= 0;
    parallel_for(b->pairs, threads, [&](int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; ++i) {
            const int64_t n = ref_lens[i], m = qry_lens[i];
            int64_t code = pair_code(refs[i], n, qrys[i], m);
            if (code == 0) {
                // Every cost the sweep holds stays well inside 32 bits.
                if ((n + m + 4) * dearest >= REAL_LIMIT || n >= (int64_t(1) << 28) || m >= (int64_t(1) << 28))
                    code = DINARA_OUT_OF_MEMORY;
            }
            if (code == 0 && b->user_band) {
                // No alignment inside a band missing the start's or the end's diagonal (lanes.mojo:1363-1367).
                const int64_t end = m - n;
                if (b->user_low > 0 || b->user_high < 0 || b->user_low > end || b->user_high < end)
                    code = capped ? DINARA_ABOVE_MAX : DINARA_OUTSIDE_BAND;
            }
            b->code[i] = code;
            b->reference_length[i] = std::max<int64_t>(n, 0);
            b->query_length[i] = std::max<int64_t>(m, 0);
        }
    });
    int64_t code;
    {
        std::lock_guard<std::mutex> serial(ctx->lock);
        DeviceGuard guard(ctx);
        code = prepare(b, refs, qrys);
    }
    if (code) {
        delete b;
        return fail(code);
    }
    if (status) *status = 0;
    return b;
}

}  // namespace

// ------------------------------------------------------------------ the C ABI

extern "C" {

void dinara_free(char *cigar) { std::free(cigar); }

int64_t dinara_device_count(void) { return device_count(); }

dinara_batch *dinara_batch_new(int64_t gpu, int64_t kind, int64_t pairs, const char *const *references,
                               const int64_t *reference_lengths, const char *const *queries,
                               const int64_t *query_lengths, const dinara_costs *costs, const dinara_mode *mode,
                               const dinara_options *options, int64_t threads, int64_t *status) {
    return batch_new(gpu, kind, pairs, references, reference_lengths, queries, query_lengths, costs, mode, options,
                     threads, status);
}

int64_t dinara_batch_run(dinara_batch *batch) {
    if (!batch) return DINARA_INVALID_MODE;
    return run(batch);
}

int64_t dinara_batch_distances(dinara_batch *b, int64_t *results) {
    if (!b || b->kind != DINARA_BATCH_DISTANCES || !b->ran) return DINARA_INVALID_MODE;
    if (const int64_t code = fetch(b)) return code;
    const bool capped = b->options.max_cost >= 0;
    for (int64_t i = 0; i < b->pairs; ++i) results[i] = b->code[i];
    for (size_t s = 0; s < b->slot_pair.size(); ++s) {
        const int32_t p = b->slot_pair[s];
        if (b->status_h[s] != 0) {
            results[p] = unsettled_code(b->status_h[s]);
            continue;
        }
        const int64_t cost = b->cost_h[s];
        results[p] = capped && cost > b->options.max_cost ? DINARA_ABOVE_MAX : cost;
    }
    return 0;
}

int64_t dinara_batch_alignments(dinara_batch *b, dinara_alignment *alignments, int64_t *statuses) {
    if (!b || b->kind != DINARA_BATCH_ALIGNMENTS || !b->ran) return DINARA_INVALID_MODE;
    if (const int64_t code = fetch(b)) return code;
    const bool capped = b->options.max_cost >= 0;
    for (int64_t i = 0; i < b->pairs; ++i) {
        statuses[i] = b->code[i];
        if (b->code[i] != 0) {
            alignments[i].cigar = nullptr;
            alignments[i].cigar_length = 0;
        }
    }
    std::atomic<bool> oom{false};
    parallel_for((int64_t)b->slot_pair.size(), b->threads, [&](int64_t lo, int64_t hi) {
        for (int64_t s = lo; s < hi; ++s) {
            const int32_t p = b->slot_pair[s];
            dinara_alignment &out = alignments[p];
            out.cigar = nullptr;
            out.cigar_length = 0;
            // Free ends: the cost is the forward phase's, the CIGAR the anchored phase's.
            const int32_t status = b->free_ends ? b->a_status_h[s] : b->status_h[s];
            if (status != 0) {
                statuses[p] = unsettled_code(status);
                continue;
            }
            const int64_t cost = b->free_ends ? b->a_cost_h[s] : b->cost_h[s];
            if (capped && cost > b->options.max_cost) {
                statuses[p] = DINARA_ABOVE_MAX;
                continue;
            }
            if (b->free_ends && b->status_h[s] != 0) {
                statuses[p] = unsettled_code(b->status_h[s]);
                continue;
            }
            const int64_t length = b->cigar_length_h[s];
            char *text = static_cast<char *>(std::malloc(length + 1));
            if (!text) {
                statuses[p] = DINARA_OUT_OF_MEMORY;
                oom.store(true);
                continue;
            }
            std::memcpy(text, b->cigars_h.data() + b->cigar_offset_h[s], length);
            text[length] = 0;
            out.cost = cost;
            out.score = -cost;
            out.reference_start = 0;
            out.reference_end = b->reference_length[p];
            out.query_start = 0;
            out.query_end = b->query_length[p];
            if (b->free_ends) {
                // The forward phase's end and the anchored phase's, as spans of the pair (Ties.RIGHT's mirrored).
                const int64_t n = b->rows[s], m = b->cols[s];
                const int64_t re = b->a_end_row_h[s], ce = b->a_end_col_h[s];
                const int64_t rb = b->b_end_row_h[s], cb = b->b_end_col_h[s];
                if (b->options.right_ties) {
                    out.reference_start = n - re;
                    out.reference_end = n - re + rb;
                    out.query_start = m - ce;
                    out.query_end = m - ce + cb;
                } else {
                    out.reference_start = re - rb;
                    out.reference_end = re;
                    out.query_start = ce - cb;
                    out.query_end = ce;
                }
            }
            out.cigar = text;
            out.cigar_length = length;
            statuses[p] = 0;
        }
    });
    return 0;
}

int64_t dinara_batch_features(dinara_batch *b, float *features) {
    if (!b) return DINARA_INVALID_MODE;
    const int64_t slots = (int64_t)b->slot_pair.size();
    std::fill(features, features + b->pairs * FEATURE_STRIDE, 0.0f);
    if (slots == 0) return 0;
    DeviceContext *ctx = b->ctx;
    std::lock_guard<std::mutex> serial(ctx->lock);
    DeviceGuard guard(ctx);
    DeviceBuffer out;
    if (!out.reserve((size_t)slots * FEATURE_STRIDE * 4)) return DINARA_OUT_OF_MEMORY;
    SweepArgs a = sweep_args(b);
    launch_features(a.tape, (int)slots, b->reversed ? 1 : 0, static_cast<float *>(out.ptr), ctx->stream);
    std::vector<float> host((size_t)slots * FEATURE_STRIDE);
    if (hipGetLastError() != hipSuccess || !out.read(ctx, host.data(), host.size() * 4)) return DINARA_OUT_OF_MEMORY;
    for (int64_t s = 0; s < slots; ++s)
        std::memcpy(features + (int64_t)b->slot_pair[s] * FEATURE_STRIDE, host.data() + s * FEATURE_STRIDE,
                    FEATURE_STRIDE * 4);
    return 0;
}

}  // extern "C"

namespace {

// Every slot's first band from its guessed cost (`guess_of(slot)`, below zero: no guess), classes and device lists.
template <typename GuessOf>
int64_t seed_slots(dinara_batch *b, GuessOf guess_of) {
    const int64_t slots = (int64_t)b->slot_pair.size();
    if (slots == 0) return 0;
    // The lanes' off-band price (lanes.mojo:333-345) in the batch's costs, as the device's off_band.
    const CostParams &c = b->costs;
    auto ins = [&](int64_t k) {
        if (k <= 0) return int64_t(0);
        const int64_t one = (int64_t)c.ins_open + k * (int64_t)c.ins_ext;
        return c.pieces == 2 ? std::min(one, (int64_t)c.ins_open2 + k * (int64_t)c.ins_ext2) : one;
    };
    auto del = [&](int64_t k) {
        if (k <= 0) return int64_t(0);
        const int64_t one = (int64_t)c.del_open + k * (int64_t)c.del_ext;
        return c.pieces == 2 ? std::min(one, (int64_t)c.del_open2 + k * (int64_t)c.del_ext2) : one;
    };
    auto off = [&](int64_t k, int64_t end) {
        if (k > std::max<int64_t>(end, 0)) return ins(k) + del(k - end);
        if (k < std::min<int64_t>(end, 0)) return del(-k) + ins(end - k);
        return int64_t(0);
    };
    for (auto &l : b->first_lists) l.clear();
    for (int64_t s = 0; s < slots; ++s) {
        const int32_t n = b->rows[s], m = b->cols[s], end = m - n;
        const int64_t guess = guess_of(s);
        int32_t lo = std::min(0, end) - 1, hi = std::max(0, end) + 1;
        if (guess >= 0 && !b->user_band) {
            // The narrowest band whose every outside path costs more than the guess: a pair costing no more than
            // its guess is proven by it at once.
            auto widen = [&](int32_t from, int32_t step, int32_t limit) {
                int32_t k = from, span = 1;
                while (span < (1 << 22) && k + step * span >= -limit - 1 && k + step * span <= limit + 1 &&
                       off(k + step * span, end) <= guess)
                    span *= 2;
                for (; span > 0; span /= 2)
                    if (off(k + step * span, end) <= guess) k += step * span;
                return k;
            };
            hi = std::min(std::max(hi, widen(std::max(0, end), 1, m) + 1), m + 1);
            lo = std::max(std::min(lo, widen(std::min(0, end), -1, n) - 1), -n - 1);
        }
        b->first_low[s] = lo;
        b->first_high[s] = hi;
        const int need = std::max(hi - lo + 1, 32);
        const int cls = first_class_for(need, b->groups_fit && n + m <= b->group_max_letters, b->block_mask,
                                        b->user_band);
        if (cls != CLASS_TOO_WIDE) b->first_lists[cls].push_back((int32_t)s);
    }
    std::vector<int32_t> first_all;
    b->first_offset.assign(CLASSES + 1, 0);
    for (int c2 = 0; c2 < CLASSES; ++c2) {
        b->first_offset[c2] = (int64_t)first_all.size();
        first_all.insert(first_all.end(), b->first_lists[c2].begin(), b->first_lists[c2].end());
    }
    b->first_offset[CLASSES] = (int64_t)first_all.size();
    if (first_all.empty()) first_all.push_back(0);  // every pair too wide: the device list still needs room
    DeviceContext *ctx = b->ctx;
    std::lock_guard<std::mutex> serial(ctx->lock);
    DeviceGuard guard(ctx);
    if (!b->first_low_d.write(ctx, b->first_low.data(), slots * 4) ||
        !b->first_high_d.write(ctx, b->first_high.data(), slots * 4) ||
        !b->first_lists_d.upload(ctx, first_all.data(), std::max<size_t>(first_all.size(), 1) * 4))
        return DINARA_OUT_OF_MEMORY;
    if (b->unit) return unit_lists(b);
    return 0;
}

// The launch shape when the device files each pair's first band: a first queue may hold any pair whose class starts
// at or below the bin and whose widest band reaches it. Caller holds the device lock.
void seeded_shape(dinara_batch *b) {
    bool same = true;
    for (int c = 0; c < EDIT_LENGTH_CLASSES; ++c) same = same && b->seeded_start[c] == b->start_bin[c];
    if (same) return;
    int64_t first[EDIT_BINS] = {}, retry[EDIT_BINS] = {}, trace[EDIT_BINS] = {};
    for (size_t s = 0; s < b->slot_pair.size(); ++s) {
        const int n = b->rows[s], m = b->cols[s];
        const int start = b->start_bin[edit_length_class(std::max(n, m))];
        const int widest = std::max(start, edit_widest(n + m));
        for (int k = start; k <= widest; ++k) ++first[k];
        for (int k = start + 1; k <= widest; ++k) ++retry[k];
        for (int k = 0; k <= widest; ++k) ++trace[k];
    }
    for (int k = 0; k < EDIT_BINS; ++k) {
        b->first_count[k] = first[k];
        b->retry_reach[k] = retry[k];
        b->trace_reach[k] = trace[k];
    }
    for (int c = 0; c < EDIT_LENGTH_CLASSES; ++c) b->seeded_start[c] = b->start_bin[c];
    if (b->edit_graph) {
        (void)hipGraphExecDestroy(b->edit_graph);
        b->edit_graph = nullptr;
    }
}

// Learned seeds without leaving the device (unit-cost batches): features, the network's predictions and each
// pair's first band and queue, all on the batch's stream.
int64_t seed_unit_learned(dinara_batch *b, dinara_mlp *net, double margin, double slope) {
    const int64_t slots = (int64_t)b->slot_pair.size();
    if (slots == 0) return 0;
    DeviceContext *ctx = b->ctx;
    std::lock_guard<std::mutex> serial(ctx->lock);
    DeviceGuard guard(ctx);
    (void)hipGetLastError();
    if (!b->features_d.reserve((size_t)slots * FEATURE_STRIDE * 4) || !b->predictions_d.reserve((size_t)slots * 4))
        return DINARA_OUT_OF_MEMORY;
    SweepArgs sa = sweep_args(b);
    launch_features(sa.tape, (int)slots, b->reversed ? 1 : 0, static_cast<float *>(b->features_d.ptr), ctx->stream);
    if (const int64_t code = mlp_predict_into(net, static_cast<const float *>(b->features_d.ptr), slots,
                                              static_cast<float *>(b->predictions_d.ptr)))
        return code;
    EditStart start{};
    for (int c = 0; c < EDIT_LENGTH_CLASSES; ++c) start.bin[c] = b->start_bin[c];
    start.bins = b->first_bins;
    launch_edit_seed(edit_args(b), static_cast<const float *>(b->predictions_d.ptr), margin, slope, start,
                     (int)slots, ctx->stream);
    seeded_shape(b);
    return hipPeekAtLastError() == hipSuccess ? 0 : DINARA_OUT_OF_MEMORY;
}

void drop_graph(dinara_batch *b) {
    if (b->edit_graph) {
        DeviceGuard guard(b->ctx);
        (void)hipGraphExecDestroy(b->edit_graph);
        b->edit_graph = nullptr;
    }
}

// Tunes the unit-cost sweep's launch shape on the batch itself: first the workgroups a compute unit, then each
// length class's least first bin (heaviest class first, widening while the first pass gets faster). Each
// candidate is timed as the faster of two first passes (no tracing).
int64_t autotune(dinara_batch *b, int64_t *chosen) {
    const int64_t slots = (int64_t)b->slot_pair.size();
    auto report = [&]() {
        if (!chosen) return;
        chosen[0] = b->grid_per_cu;
        for (int c = 0; c < EDIT_LENGTH_CLASSES; ++c) chosen[1 + c] = b->start_bin[c];
        chosen[1 + EDIT_LENGTH_CLASSES] = b->side_by_side_bins ? 1 : 0;
    };
    if (!b->unit || slots == 0) {
        report();
        return 0;
    }
    DeviceContext *ctx = b->ctx;
    std::lock_guard<std::mutex> serial(ctx->lock);
    DeviceGuard guard(ctx);
    (void)hipGetLastError();
    int64_t failure = 0;
    auto measure = [&]() {
        double best = 1e300;
        for (int i = 0; i < 2; ++i) {
            if (const int64_t code = run_unit(b, false)) {
                failure = code;
                return 1e300;
            }
            best = std::min(best, b->run_ms);
        }
        return best;
    };
    double best = measure();
    int best_grid = b->grid_per_cu;
    for (int grid : {2, 4, 8, 16}) {
        if (grid == best_grid) continue;
        b->grid_per_cu = grid;
        drop_graph(b);
        const double t = measure();
        if (t < best) {
            best = t;
            best_grid = grid;
        }
    }
    b->grid_per_cu = best_grid;
    drop_graph(b);
    // The bins side by side or one after another.
    b->side_by_side_bins = !b->side_by_side_bins;
    drop_graph(b);
    if (const double t = measure(); t < best)
        best = t;
    else
        b->side_by_side_bins = !b->side_by_side_bins;
    drop_graph(b);
    // Queues the device filed from learned seeds keep their bands: the start bins stay.
    if (b->seeded_start[0] >= 0) {
        report();
        return failure;
    }
    // Length classes by weight (the sum of their pairs' squared longer sides), heaviest first.
    double weight[EDIT_LENGTH_CLASSES] = {};
    for (int64_t s = 0; s < slots; ++s) {
        const double longer = std::max(b->rows[s], b->cols[s]);
        weight[edit_length_class((int)longer)] += longer * longer;
    }
    int order[EDIT_LENGTH_CLASSES];
    std::iota(order, order + EDIT_LENGTH_CLASSES, 0);
    std::sort(order, order + EDIT_LENGTH_CLASSES, [&](int x, int y) { return weight[x] > weight[y]; });
    for (int c : order) {
        if (weight[c] == 0) continue;
        int best_start = b->start_bin[c];
        for (int start = best_start + 1; start < EDIT_BINS; ++start) {
            b->start_bin[c] = start;
            if (const int64_t code = unit_lists(b)) return code;
            const double t = measure();
            if (t >= best) break;
            best = t;
            best_start = start;
        }
        b->start_bin[c] = best_start;
        if (const int64_t code = unit_lists(b)) return code;
    }
    report();
    return failure;
}

}  // namespace

extern "C" {

int64_t dinara_batch_seed(dinara_batch *b, const int64_t *predicted) {
    if (!b) return DINARA_INVALID_MODE;
    return seed_slots(b, [&](int64_t s) { return predicted[b->slot_pair[s]]; });
}

int64_t dinara_batch_seed_learned(dinara_batch *b, dinara_mlp *net, double margin) {
    return dinara_batch_seed_learned_by_length(b, net, margin, 0.0);
}

int64_t dinara_batch_seed_learned_by_length(dinara_batch *b, dinara_mlp *net, double margin, double slope) {
    if (!b || !net) return DINARA_INVALID_MODE;
    if (b->unit) return seed_unit_learned(b, net, margin, slope);
    const int64_t slots = (int64_t)b->slot_pair.size();
    if (slots == 0) return 0;
    std::vector<float> predicted((size_t)slots);
    {
        DeviceContext *ctx = b->ctx;
        std::lock_guard<std::mutex> serial(ctx->lock);
        DeviceGuard guard(ctx);
        DeviceBuffer features;
        if (!features.reserve((size_t)slots * FEATURE_STRIDE * 4)) return DINARA_OUT_OF_MEMORY;
        SweepArgs a = sweep_args(b);
        launch_features(a.tape, (int)slots, b->reversed ? 1 : 0, static_cast<float *>(features.ptr), ctx->stream);
        if (hipPeekAtLastError() != hipSuccess) return DINARA_OUT_OF_MEMORY;
        if (const int64_t code = mlp_predict_device(net, static_cast<const float *>(features.ptr), slots,
                                                    predicted.data()))
            return code;
    }
    return seed_slots(b, [&](int64_t s) {
        const double longer = std::max(1000.0, (double)std::max(b->rows[s], b->cols[s]));
        const double guess = std::expm1((double)predicted[s] + margin + slope * std::log(longer / 1000.0));
        // Clamped before the cast: past 2^40 every band is the whole matrix anyway, and an infinite or huge
        // prediction must not reach a conversion it does not fit (NaN fails the test and gives 0).
        constexpr double most = 1099511627776.0;
        return guess > 0 ? (int64_t)(guess < most ? guess : most) : int64_t(0);
    });
}

int64_t dinara_batch_set_retry(dinara_batch *b, int64_t doubling) {
    if (!b) return DINARA_INVALID_MODE;
    const bool changed = b->doubling != (doubling != 0);
    b->doubling = doubling != 0;
    drop_graph(b);
    // The host-filed first queues follow the policy (doubling keeps to the powers of two); seeded ones are filed
    // again by the next seeding.
    if (!changed || !b->unit || b->slot_pair.empty()) return 0;
    std::lock_guard<std::mutex> serial(b->ctx->lock);
    DeviceGuard guard(b->ctx);
    return unit_lists(b);
}

int64_t dinara_batch_set_graph(dinara_batch *b, int64_t on) {
    if (!b) return DINARA_INVALID_MODE;
    b->use_graph = on != 0;
    return 0;
}

int64_t dinara_batch_set_keep_first(dinara_batch *b, int64_t on) {
    if (!b) return DINARA_INVALID_MODE;
    b->keep_first = on != 0;
    drop_graph(b);
    b->arena_want = 0;  // the arena it asks for follows the setting: worked out again at the next run
    return 0;
}

int64_t dinara_batch_set_side_by_side(dinara_batch *b, int64_t on) {
    if (!b) return DINARA_INVALID_MODE;
    b->side_by_side_bins = on != 0;
    drop_graph(b);
    return 0;
}

int64_t dinara_batch_set_whole(dinara_batch *b, int64_t on) {
    if (!b || b->free_ends || b->user_band) return DINARA_INVALID_MODE;
    b->whole = on != 0;
    return 0;
}

int64_t dinara_batch_set_persistent(dinara_batch *b, int64_t on) {
    if (!b) return DINARA_INVALID_MODE;
    b->persistent = on != 0;
    drop_graph(b);
    return 0;
}

int64_t dinara_batch_set_tuning(dinara_batch *b, int64_t grid_per_cu, const int64_t *start_bins) {
    if (!b || grid_per_cu < 1 || grid_per_cu > 64) return DINARA_INVALID_MODE;
    if (start_bins)
        for (int c = 0; c < EDIT_LENGTH_CLASSES; ++c)
            if (start_bins[c] < 0 || start_bins[c] >= EDIT_BINS) return DINARA_INVALID_MODE;
    b->grid_per_cu = (int)grid_per_cu;
    if (start_bins)
        for (int c = 0; c < EDIT_LENGTH_CLASSES; ++c) b->start_bin[c] = (int)start_bins[c];
    drop_graph(b);
    if (!b->unit || b->slot_pair.empty()) return 0;
    std::lock_guard<std::mutex> serial(b->ctx->lock);
    DeviceGuard guard(b->ctx);
    return unit_lists(b);
}

int64_t dinara_batch_set_bins(dinara_batch *b, int64_t bins) {
    if (!b || bins <= 0 || bins > (int64_t)EDIT_ALL_BINS) return DINARA_INVALID_MODE;
    b->first_bins = (uint32_t)bins;
    drop_graph(b);
    // The host-filed first queues follow the set; seeded ones are filed again by the next seeding.
    if (!b->unit || b->slot_pair.empty()) return 0;
    std::lock_guard<std::mutex> serial(b->ctx->lock);
    DeviceGuard guard(b->ctx);
    return unit_lists(b);
}

int64_t dinara_batch_autotune(dinara_batch *b, int64_t *chosen) {
    if (!b) return DINARA_INVALID_MODE;
    return autotune(b, chosen);
}

void dinara_batch_free(dinara_batch *batch) {
    if (!batch) return;
    {
        DeviceGuard guard(batch->ctx);
        delete batch;
    }
}

double dinara_batch_stat(dinara_batch *b, int64_t what) {
    if (!b) return 0;
    const bool counted = what == 4 || what == 9 ||
                         (what >= 700 && what < 1200 && what / 100 != 8 && what % 100 < EDIT_BINS);
    if (counted && b->stats_pending) {  // the unit path's counters of the last run
        DeviceContext *ctx = b->ctx;
        std::lock_guard<std::mutex> serial(ctx->lock);
        DeviceGuard guard(ctx);
        uint32_t stats[2 * EDIT_BINS + 2] = {};
        unsigned long long work[2 * EDIT_BINS] = {};
        if (b->edit_stats.read(ctx, stats, sizeof(stats)) && b->edit_work.read(ctx, work, sizeof(work))) {
            for (int bin = 0; bin < EDIT_BINS; ++bin) {
                b->edit_sweeps[bin] = stats[bin];
                b->trace_sweeps[bin] = stats[EDIT_BINS + bin];
                b->sweep_work[bin] = (double)work[bin];
                b->trace_work[bin] = (double)work[EDIT_BINS + bin];
            }
            b->first_proven = (double)(b->stats_slots - stats[2 * EDIT_BINS]);
            b->walked_first = stats[2 * EDIT_BINS + 1];
        }
        b->stats_pending = false;
    }
    switch (what) {
        case 0: return b->run_ms;
        case 1: return b->sweep_ms;
        case 2: return b->launches;
        case 4: return b->first_proven;
        case 5: return b->walk_ms;
        case 6: return b->unit ? 1 : 0;
        case 7: return b->graph_used;
        case 8: return b->trace_passes;
        case 9: return b->walked_first;
        case 10: return b->full_pairs;
        case 11: return b->full_ms;
        case 12: return EDIT_BINS;
        case 13: return b->first_bins;
        case 14: return b->group16_top;
        default: break;
    }
    if (what >= 600 && what < 1300 && what % 100 < EDIT_BINS) {
        const int bin = (int)(what % 100);
        switch (what / 100) {
            case 12: return edit_width(bin);
            case 10: return b->sweep_work[bin];
            case 11: return b->trace_work[bin];
            case 6: return b->edit_ms[bin];
            case 7: return b->edit_sweeps[bin];
            case 8: return b->trace_ms[bin];
            case 9: return b->trace_sweeps[bin];
        }
    }
    if (what >= 100 && what < 500 && what % 100 < CLASSES) {
        const int c = (int)(what % 100);
        switch (what / 100) {
            case 1: return b->class_ms[c];
            case 2: return b->class_launches[c];
            case 4: return b->class_pairs[c];
        }
    }
    switch (what) {
        default: return 0;
    }
}

// A batch run once (the drop-in calls): the unit path's launches go one by one without timing events, since a
// captured graph would be instantiated for a single replay.
void one_shot(dinara_batch *b) {
    if (!b->unit) return;
    b->use_graph = false;
    b->timed = false;
}

int64_t dinara_gpu_distances(int64_t gpu, int64_t pairs, const char *const *references,
                             const int64_t *reference_lengths, const char *const *queries,
                             const int64_t *query_lengths, const dinara_costs *costs, const dinara_mode *mode,
                             const dinara_options *options, int64_t threads, int64_t *results) {
    ParsedCosts parsed = parse_costs(costs);
    if (parsed.code) return parsed.code;
    if (const int64_t code = check_mode(mode)) return code;
    if (pairs <= 0) return 0;
    int64_t status = 0;
    dinara_batch *b = batch_new(gpu, DINARA_BATCH_DISTANCES, pairs, references, reference_lengths, queries,
                                query_lengths, costs, mode, options, threads, &status);
    if (!b) return status;
    one_shot(b);
    int64_t code = run(b);
    if (code == 0) code = dinara_batch_distances(b, results);
    dinara_batch_free(b);
    return code;
}

int64_t dinara_gpu_alignments(int64_t gpu, int64_t pairs, const char *const *references,
                              const int64_t *reference_lengths, const char *const *queries,
                              const int64_t *query_lengths, const dinara_costs *costs, const dinara_mode *mode,
                              const dinara_options *options, int64_t threads, dinara_alignment *alignments,
                              int64_t *statuses) {
    ParsedCosts parsed = parse_costs(costs);
    if (parsed.code) return parsed.code;
    if (const int64_t code = check_mode(mode)) return code;
    if (pairs <= 0) return 0;
    int64_t status = 0;
    dinara_batch *b = batch_new(gpu, DINARA_BATCH_ALIGNMENTS, pairs, references, reference_lengths, queries,
                                query_lengths, costs, mode, options, threads, &status);
    if (!b) return status;
    one_shot(b);
    int64_t code = run(b);
    if (code == 0) code = dinara_batch_alignments(b, alignments, statuses);
    dinara_batch_free(b);
    return code;
}

int64_t dinara_distances(int64_t pairs, const char *const *references, const int64_t *reference_lengths,
                         const char *const *queries, const int64_t *query_lengths, const dinara_costs *costs,
                         const dinara_mode *mode, const dinara_options *options, int64_t threads, int64_t *results) {
    return dinara_gpu_distances(0, pairs, references, reference_lengths, queries, query_lengths, costs, mode, options,
                                threads, results);
}

int64_t dinara_alignments(int64_t pairs, const char *const *references, const int64_t *reference_lengths,
                          const char *const *queries, const int64_t *query_lengths, const dinara_costs *costs,
                          const dinara_mode *mode, const dinara_options *options, int64_t threads,
                          dinara_alignment *alignments, int64_t *statuses) {
    return dinara_gpu_alignments(0, pairs, references, reference_lengths, queries, query_lengths, costs, mode,
                                 options, threads, alignments, statuses);
}

}  // extern "C"

namespace dinara {

// dinara_gpu_distances under unit costs with every device buffer taken from `carve` (see carved.h).
int64_t carved_unit_distances(Carve *carve, int64_t gpu, int64_t pairs, const char *const *references,
                              const int64_t *reference_lengths, const char *const *queries,
                              const int64_t *query_lengths, const dinara_costs *costs, const dinara_mode *mode,
                              const dinara_options *options, int64_t threads, int64_t *results) {
    ParsedCosts parsed = parse_costs(costs);
    if (parsed.code) return parsed.code;
    if (const int64_t code = check_mode(mode)) return code;
    if (parsed.two_piece || free_mode(mode)) return DINARA_INVALID_MODE;
    const Options o = parse_options(options);
    const bool unit = parsed.opening == 0 && parsed.del_opening == 0 && parsed.extension == parsed.mismatch &&
                      parsed.del_extension == parsed.mismatch && o.band_low <= -UNBOUNDED && o.band_high >= UNBOUNDED;
    if (!unit) return DINARA_INVALID_MODE;
    if (pairs <= 0) return 0;
    CarveScope scope(carve);
    int64_t status = 0;
    dinara_batch *b = batch_new(gpu, DINARA_BATCH_DISTANCES, pairs, references, reference_lengths, queries,
                                query_lengths, costs, mode, options, threads, &status);
    if (!b) return carve->exhausted ? DINARA_SCRATCH_FULL : status;
    int64_t code = 0;
    if (!b->unit) {
        code = DINARA_INVALID_MODE;  // the bit-parallel path switched off
    } else {
        // One stream, launches one by one: no graph to instantiate, no side streams or timing events to make.
        b->use_graph = false;
        b->side_by_side_bins = false;
        b->timed = false;
        code = run(b);
        if (code == 0) code = dinara_batch_distances(b, results);
        if (code != 0 && carve->exhausted) code = DINARA_SCRATCH_FULL;
    }
    dinara_batch_free(b);
    return code;
}

}  // namespace dinara
