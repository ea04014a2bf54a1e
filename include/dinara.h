/*
 * dinara-align on MI355X: the C ABI of libdinara_hip.so.
 *
 * The batch entry points run on an AMD Instinct MI355X (gfx950) through hand-written HIP kernels. Their
 * struct layouts, argument meanings and result codes are those of the reference library's C header, so a
 * caller of the reference's batch calls links against this library unchanged; the device entry points
 * (dinara_gpu_*, dinara_batch_*) are new symbols, so the header only grows.
 *
 * Sequences are bytes, each matching only itself, save 0xFE and 0xFF, which are refused. The first
 * sequence of a pair is the reference, the second the query: a CIGAR's `=` is a match, `X` a
 * substitution (`M` for either without eqx), `D` a reference letter alone, `I` a query letter alone.
 *
 * What the device serves: global alignment under unit, linear, gap-affine or two-piece costs, deletions
 * optionally priced apart from insertions, any band and any cost cap; free ends (ends_free and its infix,
 * prefix and suffix presets, costs minimized, no user band); and, under a substitution table, global and local
 * scores and alignments. Other modes (extension, local or free ends under Costs with a match score, free ends
 * under a user band) return DINARA_INVALID_MODE for the whole batch: this library holds no CPU fallback.
 * A global pair no band settles (its band wider than the widest, or its letters more than the band kernels
 * stage) is swept whole after the banded run; it reports DINARA_OUT_OF_MEMORY only when its decisions would
 * pass 16 GiB, or under free ends or a user band.
 */
#ifndef DINARA_H
#define DINARA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Per-pair and batch-wide result codes: reference c/dinara.h:35-49. */
#define DINARA_UNSUPPORTED_SYMBOLS (-1) /* a 0xFE or 0xFF byte, or a letter outside a Scoring's alphabet */
#define DINARA_OUT_OF_MEMORY (-2)       /* device or host memory ran out */
#define DINARA_INVALID_COSTS (-3)       /* costs no alignment can be searched by */
#define DINARA_ABOVE_MAX (-4)           /* every alignment costs more than max_cost */
#define DINARA_OUTSIDE_BAND (-5)        /* no alignment stays inside the band */
#define DINARA_INVALID_MODE (-6)        /* a mode this library does not serve */
#define DINARA_INVALID_LENGTH (-7)      /* a sequence's length below zero */
#define DINARA_NO_DEVICE (-8)           /* new: no usable GPU, or the GPU id is out of range */
#define DINARA_SCRATCH_FULL (-9)        /* new: the batch needs more than dinara_gpu_distances_static's scratch */

/* Reference c/dinara.h:59-69. A null pointer is unit costs {1, 0, 1, -1, 0}. */
typedef struct {
    int64_t mismatch;            /* a substitution's cost */
    int64_t opening;             /* a gap's cost before its letters */
    int64_t extension;           /* each gapped letter's cost */
    int64_t opening2;            /* the second piece's opening; negative for one piece */
    int64_t extension2;          /* the second piece's extension */
    int64_t deletion_opening;    /* a deletion's opening, with deletion_extension above zero */
    int64_t deletion_extension;  /* a deletion's extension; zero for deletions costing as insertions */
    int64_t deletion_opening2;   /* a deletion's second opening; negative for one piece */
    int64_t deletion_extension2; /* a deletion's second extension */
} dinara_costs;

#define DINARA_ENDS_FREE 0
#define DINARA_EXTENSION 1
#define DINARA_LOCAL 2
#define DINARA_ALL INT64_MAX

/* Reference c/dinara.h:91-101. A null pointer is a global alignment. */
typedef struct {
    int64_t kind;
    int64_t reference_start;
    int64_t reference_end;
    int64_t query_start;
    int64_t query_end;
    int64_t match_score;
    int64_t anchor;
    int64_t zdrop;
    int64_t end_bonus;
} dinara_mode;

/* Reference c/dinara.h:115-122. A band's diagonal is reference letters less query letters; INT64_MIN and
 * INT64_MAX for no band. A null pointer is no band, no cap, eqx, indels left, default memory. */
typedef struct {
    int64_t band_low;
    int64_t band_high;
    int64_t max_cost;   /* the cap when zero or more; negative for none */
    int64_t eqx;        /* nonzero writes `=` and `X`, zero `M` */
    int64_t right_ties; /* nonzero places tied indels right (WFA2-lib's rule), zero left */
    int64_t max_memory; /* accepted for compatibility; the device traces every pair whole */
} dinara_options;

/* Reference c/dinara.h:129-138. */
typedef struct {
    int64_t cost;
    int64_t score;
    int64_t reference_start;
    int64_t reference_end;
    int64_t query_start;
    int64_t query_end;
    char *cigar;          /* NUL-terminated; free with dinara_free */
    int64_t cigar_length; /* its bytes, the NUL not counted */
} dinara_alignment;

/* Frees a CIGAR a batch returned; null is nothing to free. Replaces reference c/dinara.h:207. */
void dinara_free(char *cigar);

/*
 * A batch on GPU 0: pair i is references[i] (reference_lengths[i] bytes) against queries[i]. Zero, or a
 * code failing the whole batch (costs or a mode no pair can take), and then nothing written. `threads`
 * above zero packs the batch on that many host threads; zero or less packs on the caller's thread.
 * Replace reference c/dinara.h:219-222 and :225-228 (same signatures, same result codes).
 */
int64_t dinara_distances(int64_t pairs, const char *const *references, const int64_t *reference_lengths,
                         const char *const *queries, const int64_t *query_lengths, const dinara_costs *costs,
                         const dinara_mode *mode, const dinara_options *options, int64_t threads, int64_t *results);

int64_t dinara_alignments(int64_t pairs, const char *const *references, const int64_t *reference_lengths,
                          const char *const *queries, const int64_t *query_lengths, const dinara_costs *costs,
                          const dinara_mode *mode, const dinara_options *options, int64_t threads,
                          dinara_alignment *alignments, int64_t *statuses);

/* New: how many GPUs the library sees. No counterpart in the reference (its Mojo Placement.on_gpu,
 * reference common.mojo:130-160, names the device instead). */
int64_t dinara_device_count(void);

/* New: dinara_distances / dinara_alignments on GPU `gpu`, mirroring the reference's Mojo
 * distances/alignments(..., placement=Placement.on_gpu(gpu, threads)), reference api.mojo:895-903,1216-1225. */
int64_t dinara_gpu_distances(int64_t gpu, int64_t pairs, const char *const *references,
                             const int64_t *reference_lengths, const char *const *queries,
                             const int64_t *query_lengths, const dinara_costs *costs, const dinara_mode *mode,
                             const dinara_options *options, int64_t threads, int64_t *results);

int64_t dinara_gpu_alignments(int64_t gpu, int64_t pairs, const char *const *references,
                              const int64_t *reference_lengths, const char *const *queries,
                              const int64_t *query_lengths, const dinara_costs *costs, const dinara_mode *mode,
                              const dinara_options *options, int64_t threads, dinara_alignment *alignments,
                              int64_t *statuses);

/*
 * New, with no counterpart in the reference: dinara_gpu_distances for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). Same arguments and codes,
 * for unit costs (times a scale), global, with no band; other costs return DINARA_INVALID_MODE. Every device
 * buffer the batch needs (packed letters, queues, counters, results) comes from one fixed array of
 * DINARA_STATIC_SCRATCH_BYTES in the library, of which each GPU holds a copy from its first call until the process
 * ends; a batch needing more returns DINARA_SCRATCH_FULL before anything runs (letters at 2 bits and about 444 bytes
 * a pair must fit; 8-bit letters or ragged groups may need more, which is refused the same way). The call allocates
 * no device memory and makes no graph, stream or event (beyond the device's own stream and events, made once by the
 * library's first call on that GPU). Calls on one GPU take turns on a lock; the array keeps the last call's
 * contents, and nothing in it is read before the call writes it.
 */
#define DINARA_STATIC_SCRATCH_BYTES (64ll << 20)
int64_t dinara_gpu_distances_static(int64_t gpu, int64_t pairs, const char *const *references,
                                    const int64_t *reference_lengths, const char *const *queries,
                                    const int64_t *query_lengths, const dinara_costs *costs, const dinara_mode *mode,
                                    const dinara_options *options, int64_t threads, int64_t *results);

/* New: a substitution table, as the reference's Mojo/Python `Scoring` (reference scoring.mojo:126-180,
 * python/dinara_align/__init__.py:272-309): `letters` letters of `alphabet`, `table[a * letters + b]` what
 * reference letter a against query letter b scores (-128..127), a gap of k letters `opening + k * extension`,
 * both zero or less. At most 32 letters. */
typedef struct {
    const char *alphabet;
    int64_t letters;
    const int8_t *table;
    int64_t opening;
    int64_t extension;
} dinara_scoring;

/* New: every pair's best score under `scoring` (global, or DINARA_LOCAL with match_score zero), mirroring
 * the reference's Mojo api.scores (reference api.mojo:843-892). Scores may be below zero, so a pair that cannot
 * be scored fails the whole call with its code (the first such pair's in the batch's order, as the reference's
 * batch raises), and nothing is written. */
int64_t dinara_gpu_scores(int64_t gpu, int64_t pairs, const char *const *references,
                          const int64_t *reference_lengths, const char *const *queries, const int64_t *query_lengths,
                          const dinara_scoring *scoring, const dinara_mode *mode, int64_t threads, int64_t *scores);

/* New: every pair's alignment under `scoring`, mirroring the reference's Mojo api.alignments with a Scoring
 * (reference api.mojo:895-1001): Gotoh's walk picks among equal alignments; a local alignment's spans are
 * placed at the last place its letters occur (reference scoring.mojo:724-745). cost = -score. */
int64_t dinara_gpu_scoring_alignments(int64_t gpu, int64_t pairs, const char *const *references,
                                      const int64_t *reference_lengths, const char *const *queries,
                                      const int64_t *query_lengths, const dinara_scoring *scoring,
                                      const dinara_mode *mode, int64_t eqx, int64_t threads,
                                      dinara_alignment *alignments, int64_t *statuses);

/*
 * New: a Scoring batch kept resident on the device, as dinara_batch below is for costs: dinara_scoring_batch_new
 * validates, packs and uploads (kind DINARA_BATCH_DISTANCES for scores, DINARA_BATCH_ALIGNMENTS for alignments);
 * dinara_scoring_batch_run runs every pass on the device; _scores / _alignments copy the last run's results out in
 * the caller's order (a local alignment's spans are found in `references`/`queries`, which must be the batch's own
 * letters). A pair whose scores could pass 32 bits (reference scoring.mojo:216-238) reports DINARA_INVALID_COSTS;
 * one whose reference passes what a block's LDS holds, DINARA_OUT_OF_MEMORY. dinara_scoring_batch_stat: `what` 0
 * the run (ms), 2 the launches, 10-13 the global, local-end, reach and trace passes (ms). No counterpart in the
 * reference.
 */
typedef struct dinara_scoring_batch dinara_scoring_batch;

dinara_scoring_batch *dinara_scoring_batch_new(int64_t gpu, int64_t kind, int64_t pairs,
                                               const char *const *references, const int64_t *reference_lengths,
                                               const char *const *queries, const int64_t *query_lengths,
                                               const dinara_scoring *scoring, const dinara_mode *mode, int64_t eqx,
                                               int64_t threads, int64_t *status);
int64_t dinara_scoring_batch_run(dinara_scoring_batch *batch);
int64_t dinara_scoring_batch_scores(dinara_scoring_batch *batch, int64_t *scores);
int64_t dinara_scoring_batch_alignments(dinara_scoring_batch *batch, const char *const *references,
                                        const char *const *queries, dinara_alignment *alignments,
                                        int64_t *statuses);
void dinara_scoring_batch_free(dinara_scoring_batch *batch);
double dinara_scoring_batch_stat(dinara_scoring_batch *batch, int64_t what);

/*
 * New: a batch kept resident on the device, for callers that run the same pairs more than once (and for
 * timing the device path alone). dinara_batch_new validates, packs and uploads; dinara_batch_run runs the
 * kernels with every input already in HBM and waits for them; dinara_batch_distances / _alignments copy the
 * last run's results out, in the caller's order, as dinara_distances / dinara_alignments write them.
 * `kind` is DINARA_BATCH_DISTANCES or DINARA_BATCH_ALIGNMENTS. No counterpart in the reference.
 */
typedef struct dinara_batch dinara_batch;
#define DINARA_BATCH_DISTANCES 0
#define DINARA_BATCH_ALIGNMENTS 1

dinara_batch *dinara_batch_new(int64_t gpu, int64_t kind, int64_t pairs, const char *const *references,
                               const int64_t *reference_lengths, const char *const *queries,
                               const int64_t *query_lengths, const dinara_costs *costs, const dinara_mode *mode,
                               const dinara_options *options, int64_t threads, int64_t *status);
int64_t dinara_batch_run(dinara_batch *batch);
int64_t dinara_batch_distances(dinara_batch *batch, int64_t *results);
int64_t dinara_batch_alignments(dinara_batch *batch, dinara_alignment *alignments, int64_t *statuses);
void dinara_batch_free(dinara_batch *batch);

/* New: what the last dinara_batch_run spent, read from HIP events on the library's stream: `what` 0 the
 * whole run (ms), 1 the banded sweeps (ms), 2 the number of sweep launches, 4 the pairs the first band proved,
 * 5 the walk-back kernel (ms); for band class c 100+c its kernels' time (ms), 200+c its launches, 400+c the pairs
 * it swept. Classes: 0-3 sixteen lanes a pair (32/64/128/256 diagonals), 4-9 one wave a pair (64..2048, every
 * pair under a user band), 10-21 one pair a block of waves (512, 1024, 1536, 2048, 3072, 4096..8192 in steps of
 * 1024, 12288, 16384; the last two distances only). Unit-cost batches (the bit-parallel path): 6 is 1 when the
 * batch took it, 7 is 1 when the run replayed its captured hipGraph, 8 the further tracing passes the arena
 * needed, 9 the pairs proven by a sweep that kept its decisions (walked without a second sweep), 12 the number of
 * bins, 13 the first pass's bin set (dinara_batch_set_bins); 14 above zero when the gap-cost group sweeps hold
 * packed 16-bit cells (shifted costs, mismatch at most two extensions, pairs short enough), 0 for 32-bit ones; for
 * bin b (widths 64, 128, 192, 256, 384, 512, 768, 1024, 1344, 2048, 4096, 8192, 16384 diagonals) 1200+b
 * its width, 600+b its sweep's time (ms, runs without the graph), 700+b the sweeps it began, 800+b its tracing
 * sweep's time, 900+b the traces it began, 1000+b and 1100+b the word-columns (a pair's columns times its band's
 * 64-bit words) its sweeps and traces covered. Every
 * batch: 10 the pairs no band settled that the whole-matrix path swept after the run, 11 that path's time (ms,
 * host clock, included in no other figure). For roofline reporting. No counterpart in the reference. */
double dinara_batch_stat(dinara_batch *batch, int64_t what);

/*
 * New: the learned band proposal. No counterpart in the reference. A small network (288 inputs: 256 4-mer count
 * differences, reference less query, 2 log-lengths and zero padding; two ReLU layers of 256; one output, log(1 +
 * cost)) trained and run on the device by bf16 MFMA kernels with fp32 accumulation, whose predicted cost seeds a
 * batch's first bands. The band certificate still proves every result: a seed only moves work, never an answer.
 */
#define DINARA_FEATURE_STRIDE 288

/* Every pair's features, DINARA_FEATURE_STRIDE floats a pair in the caller's order, computed on the device from the
 * batch's packed letters. */
int64_t dinara_batch_features(dinara_batch *batch, float *features);

/* Seeds each pair's first band from predicted[i], its guessed cost (below zero: no guess): the narrowest band every
 * path outside which costs more than the guess, so a pair costing no more is proven in one sweep. Call before
 * dinara_batch_run; the results do not change, only the work. */
int64_t dinara_batch_seed(dinara_batch *batch, const int64_t *predicted);

typedef struct dinara_mlp dinara_mlp;

/* The same seeds from a trained network without leaving the device: the batch's features, the network's
 * predictions, and each pair's guess e^(prediction + margin) - 1. */
int64_t dinara_batch_seed_learned(dinara_batch *batch, dinara_mlp *net, double margin);
/* The same with a margin that grows with the pair's longer side: e^(prediction + margin + slope ln(max(longer, 1000)
 * / 1000)) - 1, since a long pair whose first band fails costs a long second sweep. */
int64_t dinara_batch_seed_learned_by_length(dinara_batch *batch, dinara_mlp *net, double margin, double slope);

/* How an unproven pair's next band is chosen: 0 (the default) the band its found cost proves, 1 plain band doubling,
 * the same band twice as wide (the policy the learned band is measured against). */
int64_t dinara_batch_set_retry(dinara_batch *batch, int64_t doubling);

/*
 * New: the unit-cost path. A batch under unit costs (mismatch = extension, no opening, deletions priced the same,
 * or all of it times one scale) with no user band is served by a banded bit-parallel sweep: pairs binned by band
 * width (64 diagonals a lane; lanes a pair 1 to 64, some between the powers of two, then 64 lanes of two and four
 * words), each bin one persistent launch in which groups of lanes take pairs from a queue as they finish, a pair
 * the band cannot prove filed into a wider bin; alignments sweep each settled pair's band again
 * keeping 2 bits a cell, then walk. No counterpart in the reference (its device edit distance is
 * device_edit.mojo:29-207, whole matrices, distances only).
 */
#define DINARA_EDIT_LENGTH_CLASSES 5

/* Runs replay one captured hipGraph of the run's launches (1, the default) or make them one by one with an event
 * pair around each, for per-kernel times (0). */
int64_t dinara_batch_set_graph(dinara_batch *batch, int64_t on);
/* Alignments: every sweep keeps its decisions and a pair its band proves walks them at once (1), or the sweeps
 * only prove bands and each settled pair's narrowest proven band is swept again keeping them (0, the default).
 * With good first bands (learned seeds) the first saves the second sweep; with poor ones it keeps decisions of
 * bands that fail. */
int64_t dinara_batch_set_keep_first(dinara_batch *batch, int64_t on);
/* Lane groups refill from their bin's queue (1, the default) or sweep one pair each (0: a launch as wide as the
 * bin's pairs, for comparison). */
int64_t dinara_batch_set_persistent(dinara_batch *batch, int64_t on);
/* Every pair of the batch (any costs, global, no user band) takes the whole-matrix path that serves the pairs no
 * band settles (1), or the band sweeps first (0, the default): for checking the one against the other.
 * DINARA_INVALID_MODE for a batch under free ends or a user band. No counterpart in the reference (the whole
 * matrix without a band is what its CPU path fills, lanes.mojo:795-947). */
int64_t dinara_batch_set_whole(dinara_batch *batch, int64_t on);
/* The launch shape: persistent workgroups a compute unit (1..64), and for each class of the pairs' longer side
 * (below 256, 1024, 4096, 16384 letters, and longer) the least bin its first band takes (0..12, in the widths' order
 * given for dinara_batch_stat); null keeps them. */
int64_t dinara_batch_set_tuning(dinara_batch *batch, int64_t grid_per_cu, const int64_t *start_bins);
/* The first pass's bin edges: the bins a first band may take, bit b for bin b (in the widths' order given for
 * dinara_batch_stat; every bin, 8191, by default). A pair's first band takes the narrowest bin of the set that holds
 * it and that its launches reach, else the narrowest bin of all that holds it; retries take every bin, and plain
 * band doubling keeps to the powers of two. DINARA_INVALID_MODE for an empty set or bits past bin 12. Results do
 * not change, only the work. */
int64_t dinara_batch_set_bins(dinara_batch *batch, int64_t bins);
/* Each bin's first queue, and each bin's tracing sweep, on a stream of its own so the bins run side by side (1, the
 * default), or the bins one after another (0). */
int64_t dinara_batch_set_side_by_side(dinara_batch *batch, int64_t on);
/* Chooses that shape by timing the batch's own first pass (workgroups a compute unit, side by side or not, then
 * each class's start bin, which stays as it is after learned seeding) and keeps it; writes the choice to `chosen`
 * (workgroups, the DINARA_EDIT_LENGTH_CLASSES start bins, side by side) unless null. */
int64_t dinara_batch_autotune(dinara_batch *batch, int64_t *chosen);

dinara_mlp *dinara_mlp_new(int64_t gpu, uint64_t seed);
void dinara_mlp_free(dinara_mlp *net);
/* The fp32 master weights, flat: w1 (288 x 256, row k for input k), b1 (256), w2 (256 x 256), b2 (256), w3 (256),
 * b3 (1). */
int64_t dinara_mlp_parameter_count(dinara_mlp *net);
int64_t dinara_mlp_get_weights(dinara_mlp *net, float *weights);
int64_t dinara_mlp_set_weights(dinara_mlp *net, const float *weights);
/* The mean squared error's gradients over `rows` rows at the current weights (flat, as the weights). */
int64_t dinara_mlp_gradients(dinara_mlp *net, int64_t rows, const float *features, const float *targets,
                             float *gradients, double *loss);
/* Adam over `epochs` shuffled passes in batches of `batch` rows (a multiple of 32); the last epoch's mean loss. */
int64_t dinara_mlp_train(dinara_mlp *net, int64_t rows, const float *features, const float *targets, int64_t epochs,
                         int64_t batch, double lr, uint64_t seed, double *final_loss);
/* Predictions, log(1 + cost), for `rows` rows. */
int64_t dinara_mlp_predict(dinara_mlp *net, int64_t rows, const float *features, float *out);

#ifdef __cplusplus
}
#endif

#endif
