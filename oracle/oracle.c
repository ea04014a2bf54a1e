/*
 * The oracle: a plain-C restatement of dinara-align's CPU alignment path. TEST INFRASTRUCTURE ONLY: only
 * tests/, the smoke check and bench.py's cpu_baseline leg load it, as the checker; the product never does.
 * Every function keeps the whole matrix and int64 arithmetic, so it is slow and simple on purpose.
 *
 * Each function cites the reference file:line it restates. The reference is Mojo and cannot be built or
 * imported here; the restatement is pinned by the reference's own literals and by brute force in tests/.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

#define INF (INT64_C(1) << 50)

static int64_t min64(int64_t a, int64_t b) { return a < b ? a : b; }
static int64_t max64(int64_t a, int64_t b) { return a > b ? a : b; }

/* ---------------------------------------------------------------- unit costs */

int64_t orc_edit_distance(const uint8_t *ref, int64_t n, const uint8_t *qry, int64_t m) {
    int64_t *row = malloc(sizeof(int64_t) * (size_t)(m + 1));
    for (int64_t c = 0; c <= m; c++) row[c] = c;
    for (int64_t r = 1; r <= n; r++) {
        int64_t diag = row[0];
        row[0] = r;
        for (int64_t c = 1; c <= m; c++) {
            int64_t up = row[c];
            int64_t best = diag + (ref[r - 1] != qry[c - 1]);
            best = min64(best, up + 1);
            best = min64(best, row[c - 1] + 1);
            diag = up;
            row[c] = best;
        }
    }
    int64_t out = row[m];
    free(row);
    return out;
}

/* Reference device_edit.mojo:29-77 (myers_kernel) and :176-207 (match rows): the shorter side is the pattern,
 * `words` 64-bit words of it; each text letter steps every word with the carry of the word below, and the
 * score follows the pattern's last letter, starting at its length. */
int64_t orc_myers_distance(const uint8_t *ref, int64_t n, const uint8_t *qry, int64_t m) {
    const uint8_t *pattern = ref, *text = qry;
    int64_t plen = n, tlen = m;
    if (m < n) {
        pattern = qry;
        plen = m;
        text = ref;
        tlen = n;
    }
    if (plen == 0) return tlen;
    int64_t words = (plen + 63) / 64;
    uint64_t *peq = calloc((size_t)(256 * words), sizeof(uint64_t));
    for (int64_t i = 0; i < plen; i++) peq[pattern[i] * words + i / 64] |= UINT64_C(1) << (i % 64);
    uint64_t *pv = malloc(sizeof(uint64_t) * (size_t)words), *mv = calloc((size_t)words, sizeof(uint64_t));
    for (int64_t w = 0; w < words; w++) pv[w] = ~UINT64_C(0);
    int64_t score = plen;
    int last_bit = (int)((plen - 1) % 64);
    for (int64_t j = 0; j < tlen; j++) {
        const uint64_t *eqrow = peq + text[j] * words;
        int hin = 1; /* the top row grows by one a column: a step of +1 into the first word */
        for (int64_t w = 0; w < words; w++) {
            uint64_t eq = eqrow[w], p = pv[w], mm = mv[w];
            uint64_t xv = eq | mm;
            if (hin < 0) eq |= 1;
            uint64_t xh = (((eq & p) + p) ^ p) | eq;
            uint64_t ph = mm | ~(xh | p);
            uint64_t mh = p & xh;
            int hout = 0;
            uint64_t top = UINT64_C(1) << (w == words - 1 ? last_bit : 63);
            if (ph & top) hout = 1;
            else if (mh & top) hout = -1;
            ph <<= 1;
            mh <<= 1;
            if (hin < 0) mh |= 1;
            else if (hin > 0) ph |= 1;
            pv[w] = mh | ~(xv | ph);
            mv[w] = ph & xv;
            hin = hout;
        }
        score += hin;
    }
    free(peq);
    free(pv);
    free(mv);
    return score;
}

/* ---------------------------------------------------------------- gap costs (Costs) */

/* lanes.mojo:268-286: a run of `letters` at its cheaper piece, nothing for none. */
static int64_t deleted(const orc_costs *c, int64_t letters) {
    if (letters <= 0) return 0;
    int64_t cost = c->del_open + letters * c->del_ext;
    if (c->pieces == 2) cost = min64(cost, c->del_open2 + letters * c->del_ext2);
    return cost;
}

static int64_t inserted(const orc_costs *c, int64_t letters) {
    if (letters <= 0) return 0;
    int64_t cost = c->ins_open + letters * c->ins_ext;
    if (c->pieces == 2) cost = min64(cost, c->ins_open2 + letters * c->ins_ext2);
    return cost;
}

int64_t orc_off_band(const orc_costs *costs, int64_t diagonal, int64_t end) {
    /* lanes.mojo:333-345 with no free letters: every start on lane diagonal 0, every end on `end`. */
    if (diagonal > max64(0, end)) return inserted(costs, diagonal) + deleted(costs, diagonal - end);
    if (diagonal < min64(0, end)) return deleted(costs, -diagonal) + inserted(costs, end - diagonal);
    return 0;
}

int orc_proof(const orc_costs *costs, int64_t rows, int64_t columns, int64_t found, int64_t low, int64_t high,
              int strict, int64_t *wide_low, int64_t *wide_high) {
    /* lanes.mojo:1337-1392, global, uncapped, the whole matrix the user band. */
    int64_t end = columns - rows;
    int64_t top = columns, bottom = -rows;
    int64_t beyond_high = high < top ? orc_off_band(costs, high + 1, end) : INF;
    int64_t beyond_low = low > bottom ? orc_off_band(costs, low - 1, end) : INF;
    int64_t off = min64(beyond_high, beyond_low);
    int beaten = strict ? found < off : found <= off;
    *wide_low = low;
    *wide_high = high;
    if (beaten) return 1;
    int64_t target = strict ? found + 1 : found;
    while (*wide_high < top && orc_off_band(costs, *wide_high + 1, end) < target) ++*wide_high;
    while (*wide_low > bottom && orc_off_band(costs, *wide_low - 1, end) < target) --*wide_low;
    return 0;
}

/* Flag bits, as lanes.mojo:528-540 and gap_affine.mojo:128-136/511-518 lay them out in meaning: the low three
 * bits the source the aligned layer takes, a bit per gap layer that extends into the cell. */
enum { SRC_ALIGNED = 0, SRC_DEL = 1, SRC_INS = 2, SRC_DEL2 = 3, SRC_INS2 = 4 };
#define EXT_BIT(layer) (1u << (2 + (layer)))

typedef struct {
    int64_t n, m, stride;
    int64_t *h;
    uint8_t *flags;
} Matrix;

/* The band-restricted Gotoh recurrence of lanes.mojo:795-947 over the whole matrix, rows the reference: cells
 * off diagonals band_low..band_high (library terms, reference letters less query letters) reached by nothing.
 * With `flags`, each cell's flag by the priority of lanes.mojo:905-923. Returns the corner's cost, or INF. */
static int64_t costs_sweep(const uint8_t *ref, int64_t n, const uint8_t *qry, int64_t m, const orc_costs *costs,
                           int64_t band_low, int64_t band_high, Matrix *keep) {
    int64_t stride = m + 1;
    size_t cells = (size_t)((n + 1) * stride);
    int64_t *h = malloc(sizeof(int64_t) * cells);
    int64_t *d = malloc(sizeof(int64_t) * cells), *ins = malloc(sizeof(int64_t) * cells);
    int64_t *d2 = malloc(sizeof(int64_t) * cells), *ins2 = malloc(sizeof(int64_t) * cells);
    uint8_t *flags = keep ? calloc(cells, 1) : NULL;
#define AT(r, c) ((size_t)((r) * stride + (c)))
#define IN_BAND(r, c) ((r) - (c) >= band_low && (r) - (c) <= band_high)
    /* A band missing the origin or the corner holds no alignment (lanes.mojo:1363-1367); the borders below
     * would otherwise read as reached from an origin outside it. */
    int whole = IN_BAND(0, 0) && IN_BAND(n, m);
    for (int64_t r = 0; r <= n; r++) {
        for (int64_t c = 0; c <= m; c++) {
            size_t at = AT(r, c);
            d[at] = ins[at] = d2[at] = ins2[at] = INF;
            if (!whole || !IN_BAND(r, c)) {
                h[at] = INF;
                continue;
            }
            if (r == 0) {
                h[at] = inserted(costs, c);
                continue;
            }
            if (c == 0) {
                h[at] = deleted(costs, r);
                continue;
            }
            int equal = ref[r - 1] == qry[c - 1];
            int64_t substituted = h[AT(r - 1, c - 1)] + (equal ? 0 : costs->mismatch);
            int64_t opened_del = h[AT(r - 1, c)] + costs->del_open + costs->del_ext;
            int64_t extended_del = d[AT(r - 1, c)] + costs->del_ext;
            int64_t del = min64(opened_del, extended_del);
            int64_t opened_ins = h[AT(r, c - 1)] + costs->ins_open + costs->ins_ext;
            int64_t extended_ins = ins[AT(r, c - 1)] + costs->ins_ext;
            int64_t in = min64(opened_ins, extended_ins);
            int64_t score = min64(min64(substituted, del), in);
            int64_t opened_del2 = INF, extended_del2 = INF, del2 = INF;
            int64_t opened_ins2 = INF, extended_ins2 = INF, in2 = INF;
            if (costs->pieces == 2) {
                opened_del2 = h[AT(r - 1, c)] + costs->del_open2 + costs->del_ext2;
                extended_del2 = d2[AT(r - 1, c)] + costs->del_ext2;
                del2 = min64(opened_del2, extended_del2);
                opened_ins2 = h[AT(r, c - 1)] + costs->ins_open2 + costs->ins_ext2;
                extended_ins2 = ins2[AT(r, c - 1)] + costs->ins_ext2;
                in2 = min64(opened_ins2, extended_ins2);
                score = min64(score, min64(del2, in2));
            }
            /* Sums past INF stay past it: no source that far is ever the cell's. */
            if (score >= INF) score = INF;
            h[at] = score;
            d[at] = min64(del, INF);
            ins[at] = min64(in, INF);
            d2[at] = min64(del2, INF);
            ins2[at] = min64(in2, INF);
            if (flags) {
                /* The lowest priority first, each source as cheap as the cell taking over (lanes.mojo:906-913). */
                uint8_t flag = in == score ? SRC_INS : SRC_ALIGNED;
                if (costs->pieces == 2 && in2 == score) flag = SRC_INS2;
                if (del == score) flag = SRC_DEL;
                if (costs->pieces == 2 && del2 == score) flag = SRC_DEL2;
                if (!equal && substituted == score) flag = SRC_ALIGNED;
                if (extended_del <= opened_del) flag |= EXT_BIT(SRC_DEL);
                if (extended_ins <= opened_ins) flag |= EXT_BIT(SRC_INS);
                if (costs->pieces == 2) {
                    if (extended_del2 <= opened_del2) flag |= EXT_BIT(SRC_DEL2);
                    if (extended_ins2 <= opened_ins2) flag |= EXT_BIT(SRC_INS2);
                }
                flags[at] = flag;
            }
        }
    }
    int64_t corner = h[AT(n, m)];
#undef AT
#undef IN_BAND
    free(d);
    free(ins);
    free(d2);
    free(ins2);
    if (keep) {
        keep->n = n;
        keep->m = m;
        keep->stride = stride;
        keep->h = h;
        keep->flags = flags;
    } else {
        free(h);
    }
    return corner;
}

int64_t orc_costs_cost(const uint8_t *ref, int64_t n, const uint8_t *qry, int64_t m, const orc_costs *costs,
                       int64_t band_low, int64_t band_high) {
    int64_t cost = costs_sweep(ref, n, qry, m, costs, band_low, band_high, NULL);
    return cost >= INF ? ORC_OUTSIDE : cost;
}

/* A CIGAR written a move at a time in order: runs of one letter, each its length then its letter. */
typedef struct {
    char *out;
    int64_t cap, used;
    char run_op;
    int64_t run_len;
    int overflow;
} Cigar;

static void cigar_flush(Cigar *w) {
    if (w->run_len == 0) return;
    char digits[24];
    int count = 0;
    int64_t v = w->run_len;
    while (v > 0) {
        digits[count++] = (char)('0' + v % 10);
        v /= 10;
    }
    if (w->used + count + 1 >= w->cap) {
        w->overflow = 1;
    } else {
        while (count > 0) w->out[w->used++] = digits[--count];
        w->out[w->used++] = w->run_op;
    }
    w->run_len = 0;
}

static void cigar_add(Cigar *w, char op) {
    if (w->run_len > 0 && w->run_op != op) cigar_flush(w);
    w->run_op = op;
    w->run_len++;
}

static int64_t cigar_finish(Cigar *w) {
    cigar_flush(w);
    if (w->overflow || w->used >= w->cap) return -2;
    w->out[w->used] = 0;
    return w->used;
}

enum { MOVE_ALIGNED = 0, MOVE_DEL = 1, MOVE_INS = 2 };

int64_t orc_costs_align(const uint8_t *ref, int64_t n, const uint8_t *qry, int64_t m, const orc_costs *costs,
                        int64_t band_low, int64_t band_high, int right_ties, int eqx, char *cigar, int64_t cap) {
    /* Ties.LEFT is the rule over both sequences reversed (lanes.mojo:1523-1525), the band mirrored about the
     * end diagonal (modes.mojo:477-480). */
    int left = !right_ties;
    uint8_t *a = malloc((size_t)n + 1), *b = malloc((size_t)m + 1);
    for (int64_t i = 0; i < n; i++) a[i] = left ? ref[n - 1 - i] : ref[i];
    for (int64_t i = 0; i < m; i++) b[i] = left ? qry[m - 1 - i] : qry[i];
    int64_t low = band_low, high = band_high;
    if (left) {
        int64_t target = n - m;
        low = band_low <= ORC_NO_BAND_LOW ? ORC_NO_BAND_LOW : target - band_high;
        high = band_high >= ORC_NO_BAND_HIGH ? ORC_NO_BAND_HIGH : target - band_low;
        if (band_high >= ORC_NO_BAND_HIGH) low = ORC_NO_BAND_LOW;
        if (band_low <= ORC_NO_BAND_LOW) high = ORC_NO_BAND_HIGH;
    }
    Matrix mat;
    int64_t cost = costs_sweep(a, n, b, m, costs, low, high, &mat);
    if (cost >= INF) {
        free(mat.h);
        free(mat.flags);
        free(a);
        free(b);
        return ORC_OUTSIDE;
    }
    /* The walk of lanes.mojo:1410-1460 from the corner, moves right to left over the swept sequences. */
    uint8_t *moves = malloc((size_t)(n + m + 1));
    int64_t count = 0;
    int64_t r = n, c = m;
    int layer = SRC_ALIGNED;
    while (r > 0 && c > 0) {
        uint8_t flag = mat.flags[r * mat.stride + c];
        if (layer == SRC_ALIGNED) {
            layer = flag & 7;
            if (layer == SRC_ALIGNED) {
                moves[count++] = MOVE_ALIGNED;
                r--;
                c--;
            }
            continue;
        }
        int extended = (flag & EXT_BIT(layer)) != 0;
        if (layer % 2 == 1) {
            moves[count++] = MOVE_DEL;
            r--;
        } else {
            moves[count++] = MOVE_INS;
            c--;
        }
        if (!extended) layer = SRC_ALIGNED;
    }
    while (r > 0) {
        moves[count++] = MOVE_DEL;
        r--;
    }
    while (c > 0) {
        moves[count++] = MOVE_INS;
        c--;
    }
    /* Moves over the original sequences, left to right: a reversed sweep's walk already runs from the origin;
     * a forward one's runs from the corner. */
    Cigar w = {cigar, cap, 0, 0, 0, 0};
    int64_t ri = 0, qi = 0;
    for (int64_t i = 0; i < count; i++) {
        uint8_t move = left ? moves[i] : moves[count - 1 - i];
        if (move == MOVE_DEL) {
            cigar_add(&w, 'D');
            ri++;
        } else if (move == MOVE_INS) {
            cigar_add(&w, 'I');
            qi++;
        } else {
            cigar_add(&w, eqx ? (ref[ri] == qry[qi] ? '=' : 'X') : 'M');
            ri++;
            qi++;
        }
    }
    int64_t written = cigar_finish(&w);
    free(moves);
    free(mat.h);
    free(mat.flags);
    free(a);
    free(b);
    return written < 0 ? written : cost;
}

/* ---------------------------------------------------------------- Scoring (substitution tables) */

static int64_t gap_run(const orc_scoring *s, int64_t length) {
    /* gotoh.mojo:57-60 */
    return length == 0 ? 0 : s->open + (length - 1) * s->extend;
}

typedef struct {
    int64_t n, m, stride;
    int64_t *h, *d, *i;
} Layers;

/* gotoh.mojo:265-352 (fill_rows, forward, top OPENS): the three layers of the whole matrix. Local clamps every
 * cell at zero and its borders are zero. Unreached layers sit far below every score. */
static void scoring_fill(const uint8_t *ref, int64_t n, const uint8_t *qry, int64_t m, const orc_scoring *s,
                         int local, Layers *out) {
    int64_t stride = m + 1;
    size_t cells = (size_t)((n + 1) * stride);
    int64_t *h = malloc(sizeof(int64_t) * cells), *d = malloc(sizeof(int64_t) * cells);
    int64_t *in = malloc(sizeof(int64_t) * cells);
    int64_t opening = s->open + s->extend;
    for (int64_t c = 0; c <= m; c++) {
        int64_t border = local ? 0 : gap_run(s, c);
        h[c] = border;
        d[c] = border + opening; /* a deletion from the first row pays one more opening */
        in[c] = 0;
    }
    for (int64_t r = 1; r <= n; r++) {
        int64_t border = local ? 0 : gap_run(s, r);
        size_t first = (size_t)(r * stride);
        h[first] = border;
        d[first] = border;
        in[first] = border + opening;
        const int8_t *row = s->table + ref[r - 1] * s->letters;
        for (int64_t c = 1; c <= m; c++) {
            size_t at = first + (size_t)c, above = at - (size_t)stride;
            int64_t del = max64(h[above] + s->open, d[above] + s->extend);
            int64_t ins = max64(h[at - 1] + s->open, in[at - 1] + s->extend);
            int64_t score = max64(h[above - 1] + row[qry[c - 1]], max64(del, ins));
            if (local) score = max64(score, 0);
            h[at] = score;
            d[at] = del;
            in[at] = ins;
        }
    }
    out->n = n;
    out->m = m;
    out->stride = stride;
    out->h = h;
    out->d = d;
    out->i = in;
}

static void layers_free(Layers *l) {
    free(l->h);
    free(l->d);
    free(l->i);
}

int64_t orc_scoring_global_score(const uint8_t *ref, int64_t n, const uint8_t *qry, int64_t m,
                                 const orc_scoring *scoring) {
    /* scoring.mojo:381-394: against an empty side, one gap of every letter. */
    if (n == 0 || m == 0) return gap_run(scoring, n + m);
    Layers l;
    scoring_fill(ref, n, qry, m, scoring, 0, &l);
    int64_t out = l.h[n * l.stride + m];
    layers_free(&l);
    return out;
}

/* The Gotoh walk (gotoh.mojo:129-192 decide/advance, :355-389 walk, :392-442 reconstruct), global: the moves
 * right to left from the corner into `moves` (0 aligned, 1 deletion, 2 insertion). */
static int64_t scoring_walk(const Layers *l, const uint8_t *ref, const uint8_t *qry, const orc_scoring *s,
                            uint8_t *moves) {
    int64_t count = 0, r = l->n, c = l->m, stride = l->stride;
    int layer = 0; /* 0 aligning, 1 deleting, 2 inserting */
    while (r > 0 && c > 0) {
        size_t at = (size_t)(r * stride + c), above = at - (size_t)stride;
        int64_t aligned = l->h[above - 1] + s->table[ref[r - 1] * s->letters + qry[c - 1]];
        int del_extends = l->d[above] + s->extend > l->h[above] + s->open;
        int ins_extends = l->i[at - 1] + s->extend > l->h[at - 1] + s->open;
        int source;
        if (l->h[at] == aligned) source = 0;
        else if (l->h[at] == l->d[at]) source = 1;
        else source = 2;
        int state = layer == 0 ? source : layer;
        if (state == 1) {
            moves[count++] = MOVE_DEL;
            r--;
            layer = del_extends ? 1 : 0;
        } else if (state == 2) {
            moves[count++] = MOVE_INS;
            c--;
            layer = ins_extends ? 2 : 0;
        } else {
            moves[count++] = MOVE_ALIGNED;
            r--;
            c--;
            layer = 0;
        }
    }
    while (r > 0) {
        moves[count++] = MOVE_DEL;
        r--;
    }
    while (c > 0) {
        moves[count++] = MOVE_INS;
        c--;
    }
    return count;
}

/* GappedAlignment.cigar (gotoh.mojo:202-219): `=`/`X` by the letters' own bytes, or `M`. */
static int64_t moves_cigar(const uint8_t *moves, int64_t count, const uint8_t *ref_letters,
                           const uint8_t *qry_letters, int eqx, char *cigar, int64_t cap) {
    Cigar w = {cigar, cap, 0, 0, 0, 0};
    int64_t ri = 0, qi = 0;
    for (int64_t k = count - 1; k >= 0; k--) {
        if (moves[k] == MOVE_DEL) {
            cigar_add(&w, 'D');
            ri++;
        } else if (moves[k] == MOVE_INS) {
            cigar_add(&w, 'I');
            qi++;
        } else {
            cigar_add(&w, eqx ? (ref_letters[ri] == qry_letters[qi] ? '=' : 'X') : 'M');
            ri++;
            qi++;
        }
    }
    return cigar_finish(&w);
}

int64_t orc_scoring_global_align(const uint8_t *ref, int64_t n, const uint8_t *qry, int64_t m,
                                 const orc_scoring *scoring, const uint8_t *ref_letters, const uint8_t *qry_letters,
                                 int eqx, char *cigar, int64_t cap) {
    Layers l;
    scoring_fill(ref, n, qry, m, scoring, 0, &l);
    int64_t score = l.h[n * l.stride + m];
    uint8_t *moves = malloc((size_t)(n + m + 1));
    int64_t count = scoring_walk(&l, ref, qry, scoring, moves);
    int64_t written = moves_cigar(moves, count, ref_letters, qry_letters, eqx, cigar, cap);
    free(moves);
    layers_free(&l);
    return written < 0 ? written : score;
}

int64_t orc_scoring_local_end(const uint8_t *ref, int64_t n, const uint8_t *qry, int64_t m, const orc_scoring *scoring,
                              int64_t *end_row, int64_t *end_column) {
    /* scored.mojo:81-84 and the placement of swept_cells (:259-275): of the cells scoring the best, above zero,
     * the one on the latest anti-diagonal, then the furthest along the reference. */
    *end_row = 0;
    *end_column = 0;
    if (n == 0 || m == 0) return 0;
    Layers l;
    scoring_fill(ref, n, qry, m, scoring, 1, &l);
    int64_t best = 0, best_r = 0, best_c = 0;
    for (int64_t r = 1; r <= n; r++) {
        for (int64_t c = 1; c <= m; c++) {
            int64_t v = l.h[r * l.stride + c];
            if (v <= 0) continue;
            if (v > best || (v == best && (r + c > best_r + best_c || (r + c == best_r + best_c && r > best_r)))) {
                best = v;
                best_r = r;
                best_c = c;
            }
        }
    }
    layers_free(&l);
    *end_row = best_r;
    *end_column = best_c;
    return best;
}

void orc_scoring_reach_back(const uint8_t *ref, const uint8_t *qry, int64_t end_row, int64_t end_column,
                            const orc_scoring *scoring, int64_t target, int64_t *start_row, int64_t *start_column) {
    /* vector_score.mojo:37-162: a global sweep anchored at the end over both prefixes reversed; the start is the
     * first interior cell, by anti-diagonal then row, whose score reaches the target. */
    int64_t n = end_row, m = end_column;
    uint8_t *a = malloc((size_t)n + 1), *b = malloc((size_t)m + 1);
    for (int64_t i = 0; i < n; i++) a[i] = ref[end_row - 1 - i];
    for (int64_t i = 0; i < m; i++) b[i] = qry[end_column - 1 - i];
    Layers l;
    scoring_fill(a, n, b, m, scoring, 0, &l);
    int64_t found_r = -1, found_c = -1;
    for (int64_t diag = 2; diag <= n + m && found_r < 0; diag++) {
        for (int64_t r = max64(1, diag - m); r <= min64(n, diag - 1); r++) {
            if (l.h[r * l.stride + (diag - r)] >= target) {
                found_r = r;
                found_c = diag - r;
                break;
            }
        }
    }
    layers_free(&l);
    free(a);
    free(b);
    if (found_r < 0) {
        *start_row = 0;
        *start_column = 0;
        return;
    }
    *start_row = end_row - found_r;
    *start_column = end_column - found_c;
}

/* The last place `needle` occurs in `hay` (Mojo's String.rfind over bytes), or -1. */
static int64_t rfind(const uint8_t *hay, int64_t n, const uint8_t *needle, int64_t k) {
    for (int64_t at = n - k; at >= 0; at--)
        if (memcmp(hay + at, needle, (size_t)k) == 0) return at;
    return -1;
}

int64_t orc_scoring_local_align(const uint8_t *ref, int64_t n, const uint8_t *qry, int64_t m,
                                const orc_scoring *scoring, const uint8_t *ref_letters, const uint8_t *qry_letters,
                                int eqx, char *cigar, int64_t cap, int64_t spans[4]) {
    int64_t er, ec;
    int64_t best = orc_scoring_local_end(ref, n, qry, m, scoring, &er, &ec);
    spans[0] = spans[1] = spans[2] = spans[3] = 0;
    if (best <= 0) {
        if (cap < 1) return -2;
        cigar[0] = 0;
        return 0;
    }
    int64_t sr, sc;
    orc_scoring_reach_back(ref, qry, er, ec, scoring, best, &sr, &sc);
    int64_t rn = er - sr, qm = ec - sc;
    Layers l;
    scoring_fill(ref + sr, rn, qry + sc, qm, scoring, 0, &l);
    int64_t score = l.h[rn * l.stride + qm];
    uint8_t *moves = malloc((size_t)(rn + qm + 1));
    int64_t count = scoring_walk(&l, ref + sr, qry + sc, scoring, moves);
    int64_t written = moves_cigar(moves, count, ref_letters + sr, qry_letters + sc, eqx, cigar, cap);
    free(moves);
    layers_free(&l);
    if (written < 0) return written;
    /* as_alignment (scoring.mojo:724-745): the spans at the last place the aligned letters occur. */
    int64_t rs = rfind(ref_letters, n, ref_letters + sr, rn);
    int64_t qs = rfind(qry_letters, m, qry_letters + sc, qm);
    spans[0] = rs;
    spans[1] = rs + rn;
    spans[2] = qs;
    spans[3] = qs + qm;
    return score;
}

/* ---------------------------------------------------------------- gap costs over a proven band (long pairs) */

/* The least any path through library diagonal k costs (lanes.mojo:333-345 in library terms, k = reference
 * letters less query letters, the end on n - m). */
static int64_t beyond(const orc_costs *costs, int64_t n, int64_t m, int64_t k) {
    return orc_off_band(costs, -k, m - n);
}

/* costs_sweep's recurrence and flag priority (lanes.mojo:795-947, 905-923) over library diagonals lo..hi alone:
 * values in rolling rows, flags (when asked) in (n+1) rows of hi-lo+1 bytes, row r's cell c at r*width + c-r+hi.
 * The band must hold diagonals 0 and n-m. Returns the corner's cost. */
static int64_t band_sweep(const uint8_t *ref, int64_t n, const uint8_t *qry, int64_t m, const orc_costs *costs,
                          int64_t lo, int64_t hi, uint8_t *flags) {
    const int64_t width = hi - lo + 1;
    const size_t slots = (size_t)width + 2;
    int64_t *mem = malloc(sizeof(int64_t) * slots * 8);
    int64_t *ph = mem, *pd = mem + slots, *pd2 = mem + 2 * slots, *ch = mem + 3 * slots, *cd = mem + 4 * slots,
            *cd2 = mem + 5 * slots, *ci = mem + 6 * slots, *ci2 = mem + 7 * slots;
    for (size_t i = 0; i < slots * 8; i++) mem[i] = INF;
    int64_t corner = INF;
    for (int64_t r = 0; r <= n; r++) {
        /* Slot j+1 holds cell c = r - hi + j; the row above's cell c sits one slot right, its c-1 in slot j+1. */
        for (size_t i = 0; i < slots; i++) ch[i] = cd[i] = cd2[i] = ci[i] = ci2[i] = INF;
        const int64_t c_first = r - hi > 0 ? r - hi : 0, c_last = r - lo < m ? r - lo : m;
        for (int64_t c = c_first; c <= c_last; c++) {
            const int64_t j = c - r + hi;
            const size_t at = (size_t)j + 1;
            if (r == 0) {
                ch[at] = inserted(costs, c);
                continue;
            }
            if (c == 0) {
                ch[at] = deleted(costs, r);
                continue;
            }
            const int equal = ref[r - 1] == qry[c - 1];
            const int64_t substituted = ph[at] + (equal ? 0 : costs->mismatch);
            const int64_t opened_del = ph[at + 1] + costs->del_open + costs->del_ext;
            const int64_t extended_del = pd[at + 1] + costs->del_ext;
            const int64_t del = min64(opened_del, extended_del);
            const int64_t opened_ins = ch[at - 1] + costs->ins_open + costs->ins_ext;
            const int64_t extended_ins = ci[at - 1] + costs->ins_ext;
            const int64_t in = min64(opened_ins, extended_ins);
            int64_t score = min64(min64(substituted, del), in);
            int64_t opened_del2 = INF, extended_del2 = INF, del2 = INF;
            int64_t opened_ins2 = INF, extended_ins2 = INF, in2 = INF;
            if (costs->pieces == 2) {
                opened_del2 = ph[at + 1] + costs->del_open2 + costs->del_ext2;
                extended_del2 = pd2[at + 1] + costs->del_ext2;
                del2 = min64(opened_del2, extended_del2);
                opened_ins2 = ch[at - 1] + costs->ins_open2 + costs->ins_ext2;
                extended_ins2 = ci2[at - 1] + costs->ins_ext2;
                in2 = min64(opened_ins2, extended_ins2);
                score = min64(score, min64(del2, in2));
            }
            if (score >= INF) score = INF;
            ch[at] = score;
            cd[at] = min64(del, INF);
            ci[at] = min64(in, INF);
            cd2[at] = min64(del2, INF);
            ci2[at] = min64(in2, INF);
            if (flags) {
                uint8_t flag = in == score ? SRC_INS : SRC_ALIGNED;
                if (costs->pieces == 2 && in2 == score) flag = SRC_INS2;
                if (del == score) flag = SRC_DEL;
                if (costs->pieces == 2 && del2 == score) flag = SRC_DEL2;
                if (!equal && substituted == score) flag = SRC_ALIGNED;
                if (extended_del <= opened_del) flag |= EXT_BIT(SRC_DEL);
                if (extended_ins <= opened_ins) flag |= EXT_BIT(SRC_INS);
                if (costs->pieces == 2) {
                    if (extended_del2 <= opened_del2) flag |= EXT_BIT(SRC_DEL2);
                    if (extended_ins2 <= opened_ins2) flag |= EXT_BIT(SRC_INS2);
                }
                flags[(size_t)r * (size_t)width + (size_t)j] = flag;
            }
        }
        if (r == n && n - m >= lo && n - m <= hi) corner = ch[(size_t)(m - n + hi) + 1];
        int64_t *t;
        t = ph, ph = ch, ch = t;
        t = pd, pd = cd, cd = t;
        t = pd2, pd2 = cd2, cd2 = t;
    }
    free(mem);
    return corner;
}

/* The band every path dearer than nothing but `target` stays off: diagonals from 0 and n-m outwards while a path
 * through the next one could cost `target` or less (strict) or less than `target` (plain). */
static void band_for(const orc_costs *costs, int64_t n, int64_t m, int64_t target, int strict, int64_t *lo,
                     int64_t *hi) {
    const int64_t e = n - m;
    *lo = e < 0 ? e : 0;
    *hi = e > 0 ? e : 0;
    while (*hi < n && (strict ? beyond(costs, n, m, *hi + 1) <= target : beyond(costs, n, m, *hi + 1) < target))
        ++*hi;
    while (*lo > -m && (strict ? beyond(costs, n, m, *lo - 1) <= target : beyond(costs, n, m, *lo - 1) < target))
        --*lo;
}

int64_t orc_costs_distance_proven(const uint8_t *ref, int64_t n, const uint8_t *qry, int64_t m,
                                  const orc_costs *costs) {
    /* A first band 8 diagonals past 0 and n-m, then, if a cheaper path could leave it, the band such a path
     * needs (the plain certificate of lanes.mojo:1337-1392): the second sweep's cost is then the least. */
    const int64_t e = n - m;
    int64_t lo = (e < 0 ? e : 0) - 8, hi = (e > 0 ? e : 0) + 8;
    if (lo < -m) lo = -m;
    if (hi > n) hi = n;
    int64_t found = band_sweep(ref, n, qry, m, costs, lo, hi, NULL);
    const int64_t off = min64(lo > -m ? beyond(costs, n, m, lo - 1) : INF, hi < n ? beyond(costs, n, m, hi + 1) : INF);
    if (found <= off) return found;
    band_for(costs, n, m, found, 0, &lo, &hi);
    return band_sweep(ref, n, qry, m, costs, lo, hi, NULL);
}

int64_t orc_costs_align_proven(const uint8_t *ref, int64_t n, const uint8_t *qry, int64_t m, const orc_costs *costs,
                               int right_ties, int eqx, char *cigar, int64_t cap, int64_t *band_cells) {
    const int64_t cost = orc_costs_distance_proven(ref, n, qry, m, costs);
    /* Every path off the strict band costs more than the least, so the flags inside it are the whole matrix's on
     * every optimal path; the sweep is over both sequences reversed for Ties.LEFT, whose band is the same set
     * of diagonals (a bound on any path through a diagonal depends on the lengths alone). */
    int64_t lo, hi;
    band_for(costs, n, m, cost, 1, &lo, &hi);
    const int64_t width = hi - lo + 1;
    if (band_cells) *band_cells = width * (n + 1);
    const int left = !right_ties;
    uint8_t *a = malloc((size_t)n + 1), *b = malloc((size_t)m + 1);
    for (int64_t i = 0; i < n; i++) a[i] = left ? ref[n - 1 - i] : ref[i];
    for (int64_t i = 0; i < m; i++) b[i] = left ? qry[m - 1 - i] : qry[i];
    uint8_t *flags = calloc((size_t)(n + 1) * (size_t)width, 1);
    if (!flags) {
        free(a);
        free(b);
        return -2;
    }
    const int64_t again = band_sweep(a, n, b, m, costs, lo, hi, flags);
    free(a);
    free(b);
    if (again != cost) {
        free(flags);
        return -9; /* the band lost the least cost: the certificate is broken */
    }
    /* The walk of lanes.mojo:1410-1460, as orc_costs_align walks it. */
    uint8_t *moves = malloc((size_t)(n + m + 1));
    int64_t count = 0, r = n, c = m;
    int layer = SRC_ALIGNED;
    while (r > 0 && c > 0) {
        const uint8_t flag = flags[(size_t)r * (size_t)width + (size_t)(c - r + hi)];
        if (layer == SRC_ALIGNED) {
            layer = flag & 7;
            if (layer == SRC_ALIGNED) {
                moves[count++] = MOVE_ALIGNED;
                r--;
                c--;
            }
            continue;
        }
        const int extended = (flag & EXT_BIT(layer)) != 0;
        if (layer % 2 == 1) {
            moves[count++] = MOVE_DEL;
            r--;
        } else {
            moves[count++] = MOVE_INS;
            c--;
        }
        if (!extended) layer = SRC_ALIGNED;
    }
    while (r > 0) moves[count++] = MOVE_DEL, r--;
    while (c > 0) moves[count++] = MOVE_INS, c--;
    free(flags);
    Cigar w = {cigar, cap, 0, 0, 0, 0};
    int64_t ri = 0, qi = 0;
    for (int64_t i = 0; i < count; i++) {
        const uint8_t move = left ? moves[i] : moves[count - 1 - i];
        if (move == MOVE_DEL) {
            cigar_add(&w, 'D');
            ri++;
        } else if (move == MOVE_INS) {
            cigar_add(&w, 'I');
            qi++;
        } else {
            cigar_add(&w, eqx ? (ref[ri] == qry[qi] ? '=' : 'X') : 'M');
            ri++;
            qi++;
        }
    }
    free(moves);
    const int64_t written = cigar_finish(&w);
    return written < 0 ? written : cost;
}

/* ---------------------------------------------------------------- checking an alignment against its pair */

int64_t orc_cigar_cost(const uint8_t *ref, int64_t n, const uint8_t *qry, int64_t m, const orc_costs *costs,
                       const char *cigar, int64_t length) {
    int64_t r = 0, q = 0, total = 0, count = 0;
    int digits = 0;
    for (int64_t i = 0; i < length; i++) {
        const char ch = cigar[i];
        if (ch >= '0' && ch <= '9') {
            count = count * 10 + (ch - '0');
            digits = 1;
            continue;
        }
        if (!digits || count <= 0) return -1;
        if (ch == 'D') {
            if (r + count > n) return -2;
            total += deleted(costs, count);
            r += count;
        } else if (ch == 'I') {
            if (q + count > m) return -2;
            total += inserted(costs, count);
            q += count;
        } else if (ch == '=' || ch == 'X' || ch == 'M') {
            if (r + count > n || q + count > m) return -2;
            for (int64_t k = 0; k < count; k++) {
                const int same = ref[r + k] == qry[q + k];
                if ((ch == '=' && !same) || (ch == 'X' && same)) return -3;
                if (!same) total += costs->mismatch;
            }
            r += count;
            q += count;
        } else {
            return -1;
        }
        count = 0;
        digits = 0;
    }
    if (digits || r != n || q != m) return -2;
    return total;
}

int64_t orc_check_alignments(int64_t pairs, const uint8_t *const *refs, const int64_t *ref_lengths,
                             const uint8_t *const *qrys, const int64_t *qry_lengths, const orc_costs *costs,
                             const char *const *cigars, const int64_t *cigar_lengths, const int64_t *claimed,
                             int64_t *first_bad) {
    int64_t bad = 0;
    *first_bad = -1;
    for (int64_t i = 0; i < pairs; i++) {
        const int64_t priced = orc_cigar_cost(refs[i], ref_lengths[i], qrys[i], qry_lengths[i], costs, cigars[i],
                                              cigar_lengths[i]);
        if (priced < 0 || priced != claimed[i]) {
            if (*first_bad < 0) *first_bad = i;
            bad++;
        }
    }
    return bad;
}

/* ------------------------------------------------------------------ free ends */

/* Gotoh's recurrence over the whole matrix with free starts (reference tests/test_alignment.mojo:1944-1980, the
 * lanes' LaneEnds of lanes.mojo:150-185): a cell (r, 0) with r <= start_rows or (0, c) with c <= start_columns
 * starts an alignment for nothing, every other cell takes the recurrence (borders from their one neighbour).
 * Cells off diagonals band_low..band_high (reference letters less query letters) are reached by nothing. */
static int64_t *ends_matrix(const uint8_t *ref, int64_t n, const uint8_t *qry, int64_t m, const orc_costs *costs,
                            int64_t start_rows, int64_t start_columns, int64_t band_low, int64_t band_high) {
    int64_t stride = m + 1;
    size_t cells = (size_t)((n + 1) * stride);
    int64_t *h = malloc(sizeof(int64_t) * cells);
    int64_t *d = malloc(sizeof(int64_t) * cells), *in = malloc(sizeof(int64_t) * cells);
    int64_t *d2 = malloc(sizeof(int64_t) * cells), *in2 = malloc(sizeof(int64_t) * cells);
    for (int64_t r = 0; r <= n; r++) {
        for (int64_t c = 0; c <= m; c++) {
            size_t at = (size_t)(r * stride + c);
            d[at] = in[at] = d2[at] = in2[at] = INF;
            if (r - c < band_low || r - c > band_high) {
                h[at] = INF;
                continue;
            }
            if ((c == 0 && r <= start_rows) || (r == 0 && c <= start_columns)) {
                h[at] = 0;
                continue;
            }
            int64_t best = INF;
            if (r > 0) {
                size_t up = at - stride;
                d[at] = min64(h[up] + costs->del_open + costs->del_ext, d[up] + costs->del_ext);
                best = min64(best, d[at]);
                if (costs->pieces == 2) {
                    d2[at] = min64(h[up] + costs->del_open2 + costs->del_ext2, d2[up] + costs->del_ext2);
                    best = min64(best, d2[at]);
                }
            }
            if (c > 0) {
                size_t left = at - 1;
                in[at] = min64(h[left] + costs->ins_open + costs->ins_ext, in[left] + costs->ins_ext);
                best = min64(best, in[at]);
                if (costs->pieces == 2) {
                    in2[at] = min64(h[left] + costs->ins_open2 + costs->ins_ext2, in2[left] + costs->ins_ext2);
                    best = min64(best, in2[at]);
                }
            }
            if (r > 0 && c > 0) best = min64(best, h[at - stride - 1] + (ref[r - 1] == qry[c - 1] ? 0 : costs->mismatch));
            h[at] = min64(best, INF);
            d[at] = min64(d[at], INF);
            in[at] = min64(in[at], INF);
            d2[at] = min64(d2[at], INF);
            in2[at] = min64(in2[at], INF);
        }
    }
    free(d);
    free(in);
    free(d2);
    free(in2);
    return h;
}

int64_t orc_ends_cost(const uint8_t *ref, int64_t n, const uint8_t *qry, int64_t m, const orc_costs *costs,
                      const int64_t ends[4], int64_t band_low, int64_t band_high, int64_t *end_row,
                      int64_t *end_column) {
    /* Of the cells an alignment may end on, the least, ties to the highest diagonal (reference
     * tests/test_alignment.mojo:1983-2000 rule_span). */
    int64_t *h = ends_matrix(ref, n, qry, m, costs, ends[0], ends[2], band_low, band_high);
    int64_t stride = m + 1, least = INF, er = -1, ec = -1;
    for (int64_t r = 0; r <= n; r++)
        for (int64_t c = 0; c <= m; c++) {
            if (!((r == n && m - c <= ends[3]) || (c == m && n - r <= ends[1]))) continue;
            int64_t value = h[r * stride + c];
            if (value < least || (value == least && value < INF && r - c > er - ec)) {
                least = value;
                er = r;
                ec = c;
            }
        }
    free(h);
    if (end_row) *end_row = er;
    if (end_column) *end_column = ec;
    return least >= INF ? ORC_OUTSIDE : least;
}

/* The left tie rule's span and CIGAR (reference tests/test_alignment.mojo:1983-2020, 2094-2110): the end as above;
 * then, back from that end over both prefixes reversed with no free letters, of the starts an optimum leaves from,
 * the one on the highest diagonal; the letters between aligned globally by the same rule. */
static int64_t ends_align_left(const uint8_t *ref, int64_t n, const uint8_t *qry, int64_t m, const orc_costs *costs,
                               const int64_t ends[4], int64_t band_low, int64_t band_high, int eqx, char *cigar,
                               int64_t cap, int64_t spans[4]) {
    int64_t er, ec;
    int64_t least = orc_ends_cost(ref, n, qry, m, costs, ends, band_low, band_high, &er, &ec);
    if (least == ORC_OUTSIDE) return ORC_OUTSIDE;
    uint8_t *a = malloc((size_t)er + 1), *b = malloc((size_t)ec + 1);
    for (int64_t i = 0; i < er; i++) a[i] = ref[er - 1 - i];
    for (int64_t i = 0; i < ec; i++) b[i] = qry[ec - 1 - i];
    /* Diagonal k of the reversed prefixes is er - ec - k of the pair. */
    int64_t low = band_high >= ORC_NO_BAND_HIGH ? ORC_NO_BAND_LOW : (er - ec) - band_high;
    int64_t high = band_low <= ORC_NO_BAND_LOW ? ORC_NO_BAND_HIGH : (er - ec) - band_low;
    int64_t *back = ends_matrix(a, er, b, ec, costs, 0, 0, low, high);
    int64_t sr = -1, sc = -1, stride = ec + 1;
    for (int64_t r = 0; r <= er; r++)
        for (int64_t c = 0; c <= ec; c++) {
            if (!((c == 0 && r <= ends[0]) || (r == 0 && c <= ends[2]))) continue;
            if (back[(er - r) * stride + (ec - c)] != least) continue;
            if (sr < 0 || r - c > sr - sc) {
                sr = r;
                sc = c;
            }
        }
    free(back);
    free(a);
    free(b);
    if (sr < 0) return -9; /* no start reaches the end at the least cost: never, if the recurrences agree */
    int64_t inner_low = band_low <= ORC_NO_BAND_LOW ? ORC_NO_BAND_LOW : band_low - (sr - sc);
    int64_t inner_high = band_high >= ORC_NO_BAND_HIGH ? ORC_NO_BAND_HIGH : band_high - (sr - sc);
    int64_t cost = orc_costs_align(ref + sr, er - sr, qry + sc, ec - sc, costs, inner_low, inner_high, 0, eqx, cigar,
                                   cap);
    spans[0] = sr;
    spans[1] = er;
    spans[2] = sc;
    spans[3] = ec;
    return cost < 0 ? cost : (cost == least ? least : -10);
}

static void reverse_runs(char *cigar) {
    /* A CIGAR's runs in the other order. */
    size_t length = strlen(cigar);
    char *copy = malloc(length + 1);
    size_t out = 0;
    size_t end = length;
    while (end > 0) {
        size_t start = end - 1; /* the run's operation */
        while (start > 0 && cigar[start - 1] >= '0' && cigar[start - 1] <= '9') start--;
        memcpy(copy + out, cigar + start, end - start);
        out += end - start;
        end = start;
    }
    copy[out] = 0;
    memcpy(cigar, copy, out + 1);
    free(copy);
}

int64_t orc_ends_align(const uint8_t *ref, int64_t n, const uint8_t *qry, int64_t m, const orc_costs *costs,
                       const int64_t ends[4], int64_t band_low, int64_t band_high, int right_ties, int eqx, char *cigar,
                       int64_t cap, int64_t spans[4]) {
    if (!right_ties) return ends_align_left(ref, n, qry, m, costs, ends, band_low, band_high, eqx, cigar, cap, spans);
    /* Ties.RIGHT: the left rule over both sequences reversed, the free letters of each end swapped and the band
     * mirrored about the end diagonal, its span and CIGAR read back (reference tests/test_alignment.mojo:2054-2093). */
    uint8_t *a = malloc((size_t)n + 1), *b = malloc((size_t)m + 1);
    for (int64_t i = 0; i < n; i++) a[i] = ref[n - 1 - i];
    for (int64_t i = 0; i < m; i++) b[i] = qry[m - 1 - i];
    int64_t flipped[4] = {ends[1], ends[0], ends[3], ends[2]};
    int64_t low = band_high >= ORC_NO_BAND_HIGH ? ORC_NO_BAND_LOW : (n - m) - band_high;
    int64_t high = band_low <= ORC_NO_BAND_LOW ? ORC_NO_BAND_HIGH : (n - m) - band_low;
    int64_t mirrored[4];
    int64_t cost = ends_align_left(a, n, b, m, costs, flipped, low, high, eqx, cigar, cap, mirrored);
    free(a);
    free(b);
    if (cost < 0) return cost;
    reverse_runs(cigar);
    spans[0] = n - mirrored[1];
    spans[1] = n - mirrored[0];
    spans[2] = m - mirrored[3];
    spans[3] = m - mirrored[2];
    return cost;
}
