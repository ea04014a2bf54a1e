/*
 * The oracle: a plain-C restatement of dinara-align's CPU alignment path, full matrix, int64 arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product links, loads or calls this library: only tests/, the
 * smoke check in __graft_entry__.py and bench.py's cpu_baseline leg use it, and there only as the checker
 * (or, for the baseline, as the timed CPU restatement beside the GPU). The product's path is the HIP
 * library under dinara-align_amd/csrc, which fails loudly when it is missing.
 *
 * Conventions (reference modes.mojo:3-7): the first sequence is the reference, down the rows; the second the
 * query, across the columns. A CIGAR's `D` is a reference letter alone, `I` a query letter alone, `=`/`X`
 * (or `M` without eqx) two letters aligned. A diagonal in the library's terms is reference letters less
 * query letters (modes.mojo:429-437); band bounds below are in those terms.
 */
#ifndef DINARA_ORACLE_H
#define DINARA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Gap costs in the lanes' terms (reference lanes.mojo:187-259): a deletion (reference letters alone) of k
 * letters costs del_open + k*del_ext, an insertion ins_open + k*ins_ext; with pieces == 2 the less of that
 * and the second piece's. */
typedef struct {
    int64_t mismatch;
    int64_t del_open, del_ext, ins_open, ins_ext;
    int64_t pieces;
    int64_t del_open2, del_ext2, ins_open2, ins_ext2;
} orc_costs;

#define ORC_NO_BAND_LOW (-(INT64_C(1) << 60))
#define ORC_NO_BAND_HIGH (INT64_C(1) << 60)
#define ORC_OUTSIDE (-5) /* no path inside the band */

/* Levenshtein distance by the textbook dynamic program. */
int64_t orc_edit_distance(const uint8_t *ref, int64_t n, const uint8_t *qry, int64_t m);

/* Myers/Hyyro multi-word bit-vector distance, restating reference device_edit.mojo:29-77 (the shorter side
 * the pattern, carries in {-1, 0, +1} between 64-bit words, the score starting at the pattern's length). */
int64_t orc_myers_distance(const uint8_t *ref, int64_t n, const uint8_t *qry, int64_t m);

/* The least cost over paths on diagonals band_low..band_high (library terms), or ORC_OUTSIDE. */
int64_t orc_costs_cost(const uint8_t *ref, int64_t n, const uint8_t *qry, int64_t m, const orc_costs *costs,
                       int64_t band_low, int64_t band_high);

/* An optimal global alignment under `costs` inside the band, its CIGAR the one the reference's Ties rule
 * names: the DP-flag rule (lanes.mojo:905-923) walked back from the corner (lanes.mojo:1410-1460), over both
 * sequences reversed for Ties.LEFT (right_ties == 0), the moves turned around (lanes.mojo:1523-1525).
 * Returns the cost or ORC_OUTSIDE; writes a NUL-terminated CIGAR of at most cap bytes (-2 if it does not fit). */
int64_t orc_costs_align(const uint8_t *ref, int64_t n, const uint8_t *qry, int64_t m, const orc_costs *costs,
                        int64_t band_low, int64_t band_high, int right_ties, int eqx, char *cigar, int64_t cap);

/* orc_costs_cost and orc_costs_align for long pairs, global and unbanded, in memory proportional to a band: the
 * cost from a first band 8 diagonals wide plus, when the plain certificate (lanes.mojo:1337-1392) cannot prove it,
 * the band a cheaper path would need; then the flags of the strict proven band alone (every path off it costs
 * more, so on every optimal path they are the whole matrix's) walked as orc_costs_align walks them. *band_cells
 * (unless null) gets the traced band's stored cells. -9: the band lost the least cost (never, if the certificate
 * holds). */
int64_t orc_costs_distance_proven(const uint8_t *ref, int64_t n, const uint8_t *qry, int64_t m,
                                  const orc_costs *costs);
int64_t orc_costs_align_proven(const uint8_t *ref, int64_t n, const uint8_t *qry, int64_t m, const orc_costs *costs,
                               int right_ties, int eqx, char *cigar, int64_t cap, int64_t *band_cells);

/* What a CIGAR's edits cost under `costs` (each gap run at its cheaper piece), after checking that it spells an
 * alignment of the whole pair: -1 malformed, -2 it does not consume exactly both sequences, -3 an `=` over
 * unequal letters or an `X` over equal ones. */
int64_t orc_cigar_cost(const uint8_t *ref, int64_t n, const uint8_t *qry, int64_t m, const orc_costs *costs,
                       const char *cigar, int64_t length);

/* orc_cigar_cost over a batch against each pair's claimed cost: the pairs that fail, the first in *first_bad. */
int64_t orc_check_alignments(int64_t pairs, const uint8_t *const *refs, const int64_t *ref_lengths,
                             const uint8_t *const *qrys, const int64_t *qry_lengths, const orc_costs *costs,
                             const char *const *cigars, const int64_t *cigar_lengths, const int64_t *claimed,
                             int64_t *first_bad);

/* The least any path visiting lane diagonal `diagonal` (query less reference) costs, its end on lane
 * diagonal `end`, no free ends: reference lanes.mojo:333-345. */
int64_t orc_off_band(const orc_costs *costs, int64_t diagonal, int64_t end);

/* The strict (alignments) or plain (distances) band certificate of reference lanes.mojo:1337-1392, lane
 * diagonals low..high, global mode, no cap, no user band: returns 1 when `found` is proven, else 0 with the
 * band a cheaper (strict: no dearer) path needs in *wide_low, *wide_high. */
int orc_proof(const orc_costs *costs, int64_t rows, int64_t columns, int64_t found, int64_t low, int64_t high,
              int strict, int64_t *wide_low, int64_t *wide_high);

/* Free ends (reference gap_affine.mojo:421-448 EndsFree, tests/test_alignment.mojo:1944-2020): ends[0] reference
 * letters an alignment may leave out at the start for nothing, ends[1] at the end, ends[2] and ends[3] the query's.
 * orc_ends_cost: the least cost over every start and end so allowed, and the end the tie rule takes (of the optimal
 * ends, the one on the highest diagonal, reference letters less query letters), or ORC_OUTSIDE. */
int64_t orc_ends_cost(const uint8_t *ref, int64_t n, const uint8_t *qry, int64_t m, const orc_costs *costs,
                      const int64_t ends[4], int64_t band_low, int64_t band_high, int64_t *end_row,
                      int64_t *end_column);

/* The alignment the tie rule names under free ends: Ties.LEFT takes the end above, then the start on the highest
 * diagonal an optimum ending there leaves from, and the letters between aligned globally by the left rule;
 * Ties.RIGHT is that over both sequences reversed with each sequence's free letters swapped end for end, read back.
 * Returns the cost (or ORC_OUTSIDE), writes the CIGAR and spans = {reference start, end, query start, end}. */
int64_t orc_ends_align(const uint8_t *ref, int64_t n, const uint8_t *qry, int64_t m, const orc_costs *costs,
                       const int64_t ends[4], int64_t band_low, int64_t band_high, int right_ties, int eqx, char *cigar,
                       int64_t cap, int64_t spans[4]);

/* Scoring: a substitution table `table[a * letters + b]` (a the reference's code), a gap of k letters scoring
 * `open + (k - 1) * extend` (both <= 0; the library's Scoring(opening, extension) has open = opening +
 * extension, extend = extension, reference scoring.mojo:202-214). Sequences are codes 0..letters-1. */
typedef struct {
    const int8_t *table;
    int64_t letters;
    int64_t open;
    int64_t extend;
} orc_scoring;

/* Best global score, reference gotoh.mojo:300-352 / scoring.mojo:381-394. */
int64_t orc_scoring_global_score(const uint8_t *ref, int64_t n, const uint8_t *qry, int64_t m,
                                 const orc_scoring *scoring);

/* An optimal global alignment by Gotoh's walk (decide/advance, gotoh.mojo:161-192; reconstruct :392-442).
 * `ref_letters`/`qry_letters` are the original bytes, for `=`/`X`. Returns the score; CIGAR into `cigar`. */
int64_t orc_scoring_global_align(const uint8_t *ref, int64_t n, const uint8_t *qry, int64_t m,
                                 const orc_scoring *scoring, const uint8_t *ref_letters, const uint8_t *qry_letters,
                                 int eqx, char *cigar, int64_t cap);

/* Best local score, floored at zero (scoring.mojo:381-389), and its end by the host rule (scored.mojo:81-84):
 * the highest score, then the furthest along both sequences together, then along the reference. */
int64_t orc_scoring_local_end(const uint8_t *ref, int64_t n, const uint8_t *qry, int64_t m, const orc_scoring *scoring,
                              int64_t *end_row, int64_t *end_column);

/* Where the local alignment ending at (end_row, end_column) and scoring `target` starts: the first cell by
 * anti-diagonal back from the end, then by row, whose anchored global score reaches it (vector_score.mojo:37-162). */
void orc_scoring_reach_back(const uint8_t *ref, const uint8_t *qry, int64_t end_row, int64_t end_column,
                            const orc_scoring *scoring, int64_t target, int64_t *start_row, int64_t *start_column);

/* The host local alignment (scoring.mojo:544-560 then as_alignment :724-745): returns the score, writes the
 * CIGAR and the spans, placed at rfind of the aligned letters in the original bytes. */
int64_t orc_scoring_local_align(const uint8_t *ref, int64_t n, const uint8_t *qry, int64_t m,
                                const orc_scoring *scoring, const uint8_t *ref_letters, const uint8_t *qry_letters,
                                int eqx, char *cigar, int64_t cap, int64_t spans[4]);

#ifdef __cplusplus
}
#endif

#endif
