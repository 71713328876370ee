/*
 * episdet.h -- C ABI of the MI355X (gfx950) exhaustive epistasis search.
 *
 * Drop-in boundary for hiperbio/tensor-episdet's search entry point
 *
 *   cudaError_t EpistasisDetectionSearch(unsigned* cases, unsigned* controls,
 *       int numSNPs, int numCases, int numControls, uint numSNPsWithPadding,
 *       int numCasesWithPadding, int numControlsWithPadding, int* roundsCounter,
 *       double* searchTime, float* outputFromGpu,
 *       unsigned long long* packedIndices);
 *                       (include/search-pairs.hpp:4, include/search-triplets.hpp:4)
 *
 * which the reference selects at build time (pairs/triplets x K2/MI,
 * Makefile:8-22).  Here the interaction order and the score are arguments.
 *
 * Dataset layout (unchanged from the reference, tensor-episdet.cu:111-148):
 * per class, u32 words [snp][plane][W]; plane 0 marks genotype 0, plane 1
 * genotype 1, genotype 2 is neither; sample s is bit 31-(s%32) of word s/32;
 * bits past the last sample are zero.  W is the padded word count per plane
 * (the reference pads samples to a multiple of 1024, i.e. W = 32*ceil(N/1024)).
 *
 * Every function returns EPISDET_OK (0) or a positive episdet_status code and
 * never throws.  The caller owns all host memory; the library owns all device
 * memory.  A dataset handle may be used by one host thread at a time.
 */
#ifndef EPISDET_H_
#define EPISDET_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum episdet_status {
    EPISDET_OK = 0,
    EPISDET_ERR_INVALID = 1,      /* bad argument or unsupported shape              */
    EPISDET_ERR_NO_DEVICE = 2,    /* no gfx950 device visible                       */
    EPISDET_ERR_HIP = 3,          /* a HIP runtime call failed                      */
    EPISDET_ERR_ALLOC = 4,        /* device or host allocation failed               */
    EPISDET_ERR_CANDIDATES = 5,   /* candidate buffer overflow that could not be
                                     resolved (not returned by this version: an
                                     overflow reruns the search in windows)     */
    EPISDET_ERR_INTERNAL = 6,
    EPISDET_ERR_GENOTYPES = 7,    /* a sample is set in both planes of a SNP: the
                                     input encodes no genotype for it             */
    EPISDET_ERR_SCRATCH = 8       /* episdet_run_noalloc: the search needs more
                                     device memory than its static scratch holds */
};

enum episdet_score { EPISDET_SCORE_K2 = 0, EPISDET_SCORE_MI = 1 };

enum episdet_path {
    EPISDET_PATH_AUTO = 0,        /* library's choice                               */
    EPISDET_PATH_FUSED_VALU = 1,  /* persistent count+score+reduce, AND+v_bcnt      */
    EPISDET_PATH_FUSED_MFMA = 2,  /* persistent count+score+reduce, FP4 MFMA        */
    EPISDET_PATH_PLAIN = 3        /* reference-shaped: host round loop, count kernel
                                     writes int32 tables, FP64 objective kernel   */
};

/* Reference search, same twelve parameters and meaning as
 * EpistasisDetectionSearch (search-pairs.hpp:4 / search-triplets.hpp:4),
 * preceded by the interaction order (2 or 3) and the score.
 *   cases, controls      host planes, rows [numSNPsWithPadding][2][W] with
 *                        W = numCasesWithPadding/32 (resp. controls); rows at
 *                        or beyond numSNPs are ignored.
 *   roundsCounter        rounds the reference's host loop would run at
 *                        BLOCK_SIZE=256 (search-triplets.cu:162-200), so its
 *                        "Tensor TOPS" and "Ratio" lines stay comparable.
 *   searchTime           wall-clock seconds including device allocation and
 *                        host-to-device copies (search-triplets.cu:11-12,287-289).
 *   outputFromGpu        best score (K2 minimised, MI maximised), FP64 value
 *                        rounded to float.
 *   packedIndices        X | Y<<21 | Z<<42 (3-way) or A | B<<32 (2-way),
 *                        X < Y < Z; ties go to the lowest packed index.
 * Like the reference, it prints "Outer loop iteration i out of nb" to stdout,
 * one line per block of 256 SNPs of numSNPsWithPadding (search-pairs.cu:139,
 * search-triplets.cu:164); the lines come before the search, which is one
 * launch. */
int episdet_search(int order, int score, const unsigned *cases, const unsigned *controls, int numSNPs,
                   int numCases, int numControls, unsigned numSNPsWithPadding, int numCasesWithPadding,
                   int numControlsWithPadding, int *roundsCounter, double *searchTime, float *outputFromGpu,
                   unsigned long long *packedIndices);

/* ---- device-resident datasets (extended interface, no reference counterpart
 * beyond the upload the reference does inside EpistasisDetectionSearch,
 * search-triplets.cu:17-55) ---- */

typedef struct episdet_dataset episdet_dataset;

/* Uploads both classes to device 0 and keeps them resident.
 * casesWords / controlsWords: W of the given host layout (>= ceil(N/32)).
 * numCases + numControls must stay below 2^30 (EPISDET_ERR_INVALID). */
int episdet_dataset_create(const unsigned *cases, const unsigned *controls, int numSNPs, int numCases,
                           int numControls, int casesWords, int controlsWords, episdet_dataset **out);
void episdet_dataset_destroy(episdet_dataset *ds);

typedef struct {
    int order;            /* 2 or 3                                              */
    int score;            /* episdet_score                                       */
    int path;             /* episdet_path                                        */
    int tile;             /* kernel tile id of the chosen path (ignored by AUTO):
                             FUSED_MFMA 3-way 0-5 (0: X x 64 x 64 units; 1: X x
                             32 x 64; 2: X x 64 x 32; 3 and 4: two X x 32 x 32,
                             512- or 1024-sample chunks; 5: as 4, classes in
                             sequence), FUSED_MFMA 2-way 0-1
                             (0: 128 x 64 units; 1: 64 x 64, classes in
                             sequence), FUSED_VALU 3-way 0-3, FUSED_VALU 2-way
                             0-2; PLAIN: the count kernel of its rounds, 0 =
                             AND + v_bcnt on the VALU, 1 = FP4 MFMA (3-way: the
                             64 x 64 kernel of FUSED_MFMA tile 0 writing its
                             counts, 16 rounds per launch; 2-way: the 128 x 64
                             kernel of 2-way tile 0, 32 rounds per launch)       */
    int plain_rounds;     /* PLAIN only: stop after this many rounds (0 = all);
                             used to time a contiguous slice of the round loop   */
    int candidate_cap;    /* FP32 candidates kept per launch for FP64 re-scoring;
                             0 = default (1<<20).  Values below one work unit's
                             sets (at most 8192) are raised to it.  When more
                             sets fall within the FP32 margin (heavy ties), the
                             search reruns in windows of work units that fit,
                             with the same exact answer (result.fallback = 1) */
    int rebuild;          /* 1: recompute the single and pair count tables inside
                             this call, so their cost is part of it (the
                             reference counts them inside its search,
                             search-triplets.cu:113-117,173-197)                 */
    int topk;             /* > 1: keep the topk best sets (episdet_run_topk);
                             0 or 1: the single best set                         */
    int autotune;         /* AUTO only: 1 = time the candidate kernels for this
                             shape first if it has no cached choice
                             (episdet_autotune); 0 = cached choice or the
                             default rule                                        */
    int reserved[5];
} episdet_options;

typedef struct {
    double best_score;          /* FP64 score of the winner                      */
    float best_score_f32;       /* FP32 score the fused kernel gave the winner
                                   (fused paths only; else equal to best_score)  */
    int pad0;
    unsigned long long packed;  /* winner, packed as above                       */
    int idx[3];                 /* winner indices (idx[2] = -1 for pairs)        */
    int path;                   /* path that ran                                 */
    long long sets;             /* sets evaluated (C(n,k), or the slice's share) */
    long long rounds;           /* reference-equivalent rounds                   */
    long long rounds_run;       /* PLAIN: rounds actually launched               */
    long long candidates;       /* fused: candidates re-scored in FP64           */
    double kernel_ms;           /* device time of the main search kernel(s),
                                   hipEvents on the launching stream             */
    double rescore_ms;          /* device time of the FP64 re-scoring            */
    double prep_ms;             /* host time of single/pair table builds in this
                                   call (0 when cached); the single counts are
                                   recomputed without a wait of their own, so
                                   their device time falls in the search's      */
    double wall_s;              /* host wall-clock of this call                  */
    double fp32_max_abs_err;    /* max |FP32 - FP64| over re-scored candidates   */
    double fp32_max_rel_err;    /* same, relative to |FP64|                      */
    double margin;              /* FP32 candidate margin used (2 x error bound)  */
    int fallback;               /* 1 if a candidate overflow made the fused path
                                   rerun the search in windows                  */
    int tile;                   /* kernel tile id used                           */
    int seeded;                 /* 3-way fused: 1 if the search started from a
                                   bound found through the best pairs          */
    int reserved[5];
} episdet_result;

int episdet_run(episdet_dataset *ds, const episdet_options *opt, episdet_result *res);

/* episdet_run without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same options, result and answer as episdet_run on the same
 * dataset, except:
 *  - its per-search device memory (control block, argument block, 2^18
 *    candidate slots, work-unit groups, and the FP64 score table and 3-way
 *    pair table when the dataset has not built them in an earlier
 *    episdet_run) is a static 64 MiB array that every device the library
 *    runs on holds until the process ends; a search that needs more returns
 *    EPISDET_ERR_SCRATCH before touching the device (without the dataset's
 *    own tables: 3-way up to about 1400 SNPs, any order up to about 7.5M
 *    samples);
 *  - PLAIN, rebuild and autotune are refused (EPISDET_ERR_INVALID): they
 *    allocate buffers of their own;
 *  - no seeding pre-pass (result.seeded = 0): it runs a search of its own.
 * Calls share the array of their device, which keeps the last call's
 * contents, so calls that overlapped on one device would race: the library
 * runs the calls of a process one at a time (one lock).  Dataset creation
 * itself allocates, as for episdet_run. */
int episdet_run_noalloc(episdet_dataset *ds, const episdet_options *opt, episdet_result *res);

/* ---- autotuner (no reference counterpart: the reference's BLOCK_SIZE and
 * NUM_STREAMS are compile-time constants, search-triplets.cu:4-5) ----
 *
 * Times every kernel configuration that can run this dataset's shape
 * (FP4-MFMA tiles, VALU tiles) on a window of work units from the middle of
 * its work queue, predicts each one's whole-search kernel time from the exact
 * number of sets in the window, and caches the fastest for the shape
 * (order, score, n, words per class) for the rest of the process: AUTO then
 * runs it, in episdet_run and in episdet_run_batch.  entries: up to
 * capacity of them, in the order timed; *count receives the number of
 * configurations timed, *chosen the index of the one cached. */
typedef struct {
    int path;               /* EPISDET_PATH_FUSED_VALU or _FUSED_MFMA           */
    int tile;               /* tile id                                          */
    double est_ms;          /* predicted kernel time of the whole search        */
    double sample_ms;       /* measured kernel time of the window               */
    long long sample_sets;  /* sets in the window                               */
} episdet_tune_entry;

int episdet_autotune(episdet_dataset *ds, int order, int score, episdet_tune_entry *entries, int capacity,
                     int *count, int *chosen);

/* The k best sets of one search, by exact FP64 score: K2 ascending or MI
 * descending, ties to the lowest packed index.  The reference keeps only the
 * single best set (helper.cu:552-583, reduction.hpp:64-122); this widens the
 * same reduction to k.  One fused launch keeps every set whose FP32 score is
 * within the error margin of a running bound on the k-th best; those are
 * re-scored in FP64 and ranked.  packed / scores: capacity k, best first;
 * *count receives min(k, C(n, order)); res as for episdet_run (the best set).
 * 1 <= k <= EPISDET_TOPK_MAX.  Fused paths only (AUTO, FUSED_VALU,
 * FUSED_MFMA): PLAIN returns EPISDET_ERR_INVALID. */
#define EPISDET_TOPK_MAX 1024
int episdet_run_topk(episdet_dataset *ds, const episdet_options *opt, int k, unsigned long long *packed,
                     double *scores, int *count, episdet_result *res);

/* ---- batched searches (config "many independent jobs", no reference
 * counterpart: the reference runs one search per process) ---- */

typedef struct {
    int order;                        /* 2 or 3                                  */
    int score;                        /* episdet_score                           */
    const unsigned *cases;            /* host planes as for episdet_dataset_create */
    const unsigned *controls;
    int numSNPs, numCases, numControls;
    int casesWords, controlsWords;
} episdet_job;

/* Runs many independent searches in one hipGraph launch: the work units of
 * every job go through one persistent FP4-MFMA kernel launch per kernel tile
 * (AUTO jobs of one order share one tile), side by side in the graph,
 * followed by one FP64 re-scoring kernel over all jobs' candidates.  Jobs the
 * batch kernels cannot take (PLAIN and VALU paths, >= 2^24 samples per class,
 * a second job on the same dataset, candidate overflow) run on their own
 * afterwards.  opts[j] as for episdet_run (order and score per job);
 * batch_ms, if not NULL, receives the device time of the graph.  Results are
 * filled as episdet_run fills them; for jobs run in the graph, tile is the
 * MFMA tile that ran, kernel_ms the graph's device time (re-scoring
 * included) and rescore_ms 0. */
int episdet_run_batch(episdet_dataset *const *ds, const episdet_options *opts, int njobs, episdet_result *results,
                      double *batch_ms);

/* Same, from host buffers: uploads every job, runs the batch, frees. */
int episdet_search_batch(const episdet_job *jobs, int njobs, const episdet_options *opt, episdet_result *results,
                         double *batch_ms);

/* Tables of explicit sets, counted on the device by the same counting and
 * derivation code as the fused kernels.  sets: m x order indices, each row
 * strictly increasing.  out: m x 2 x 3^order u32 (cases cells, then controls
 * cells, cell index x*9+y*3+z or a*3+b). */
int episdet_tables(episdet_dataset *ds, int order, const int *sets, int m, unsigned *out);

/* Every set's table from the fused kernel itself (small n only):
 * out[rank][2][3^order] in colex rank order, rank = C(Z,3)+C(Y,2)+X or
 * C(B,2)+A; capacity in elements.  path: FUSED_VALU or FUSED_MFMA; tile: the
 * kernel tile id of that path (as in episdet_options). */
int episdet_dump_tables(episdet_dataset *ds, int order, int path, int tile, unsigned *out, size_t capacity);

/* Order-independent 64-bit checksum over every set's table computed by the
 * fused kernel (sum over sets of a hash of rank and the 2 x 3^k cells). */
int episdet_checksum(episdet_dataset *ds, int order, int path, int tile, unsigned long long *checksum);

/* Exact binomial C(n,k) (helper.cu:758-768). */
unsigned long long episdet_n_choose_k(unsigned n, unsigned k);

/* Rounds the reference's host loop runs for n SNPs at block size 256. */
long long episdet_reference_rounds(int order, int numSNPs);

const char *episdet_status_string(int status);

/* Name, compute units and clock of device 0; EPISDET_ERR_NO_DEVICE if none. */
int episdet_device_info(char *name, int name_len, int *compute_units, int *clock_khz);

/* Waits for all work on the current device (hipDeviceSynchronize).  No
 * counterpart in the reference; host drivers use it to bracket timed regions.
 * Every search call above already returns only after its own work is done. */
int episdet_device_synchronize(void);

#ifdef __cplusplus
}
#endif

#endif /* EPISDET_H_ */
