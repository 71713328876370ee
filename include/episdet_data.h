/*
 * episdet_data.h -- host-side dataset format and synthetic data generator.
 *
 * File format (read by the reference's main, tensor-episdet.cu:58-148):
 *   dataset.txt, exactly five lines:
 *       numSNPs
 *       controls .bin file name (relative to the .txt's directory)
 *       numControls
 *       cases .bin file name
 *       numCases
 *   each .bin: raw little-endian u32, numSNPs x 2 x W words, no header,
 *   W = 32 * ceil(N / 1024) (samples padded to a multiple of 1024 bits).
 *
 * Plane layout in memory is the same [snp][plane][W] (see episdet.h).
 * Pure host code: no device is touched.
 */
#ifndef EPISDET_DATA_H_
#define EPISDET_DATA_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum episdet_maf_dist {
    EPISDET_MAF_UNIFORM = 0,   /* MAF ~ U[maf_lo, maf_hi]                           */
    EPISDET_MAF_BETA = 1       /* MAF ~ Beta(beta_a, beta_b) clipped to [maf_lo, maf_hi] */
};

typedef struct {
    int num_snps;
    int num_cases;
    int num_controls;
    int cases_words;        /* W per plane for cases; 0 = 32*ceil(N/1024)          */
    int controls_words;     /* same for controls                                    */
    unsigned long long seed;
    int maf_dist;           /* episdet_maf_dist                                     */
    double maf_lo, maf_hi;
    double beta_a, beta_b;
    int planted_sets;       /* 0, 1 or 2 planted interacting sets                  */
    int planted_order;      /* 2 or 3                                              */
    double baseline_risk;   /* penetrance of non-risk cells (e.g. 0.1)             */
    double peak_risk;       /* highest penetrance (e.g. 0.4)                       */
    double planted_maf_lo;  /* MAF range of planted SNPs                           */
    double planted_maf_hi;
    int threads;            /* 0 = OpenMP default                                  */
} episdet_gen_params;

/* Fills gp with the defaults above (seed 1, MAF U[0.05,0.5], no planted set,
 * baseline 0.1, peak 0.4, planted MAF U[0.25,0.5]). */
void episdet_gen_defaults(episdet_gen_params *gp);

/* Generates HWE genotypes (P = (1-p)^2, 2p(1-p), p^2 for genotypes 0,1,2 with
 * p the per-SNP minor allele frequency) and packs them MSB-first with zero
 * padding.  With planted sets, individuals are drawn one at a time, their
 * planted genotypes set the disease risk through a seeded penetrance table, and
 * they are kept until exactly num_cases cases and num_controls controls exist;
 * every other SNP is independent of status.  Without planted sets the labels
 * are fixed (the first num_cases individuals are cases).
 *   cases:    num_snps x 2 x cases_words u32 (caller allocated)
 *   controls: num_snps x 2 x controls_words u32
 *   planted:  planted_sets x 3 ints (sorted SNP indices, -1 padded), may be NULL
 *   penetrance: planted_sets x 27 doubles (cell x*9+y*3+z, or a*3+b), may be NULL
 * Deterministic for a given seed and independent of the thread count. */
int episdet_generate(const episdet_gen_params *gp, uint32_t *cases, uint32_t *controls, int *planted,
                     double *penetrance);

/* Reference padding: 32 * ceil(N / 1024). */
int episdet_padded_words(int num_samples);

/* Writes dir/<name>.txt, dir/<name>.controls.bin, dir/<name>.cases.bin in the
 * reference's format.  words must be the reference padding. */
int episdet_write_dataset(const char *dir, const char *name, int num_snps, int num_cases, int num_controls,
                          const uint32_t *cases, const uint32_t *controls);

/* Reads the .txt header: sizes and absolute .bin paths (buffers of path_len). */
int episdet_read_header(const char *txt_path, int *num_snps, int *num_cases, int *num_controls,
                        char *cases_path, char *controls_path, int path_len);

/* Reads both .bin files into caller buffers of num_snps_alloc x 2 x W words
 * (W = reference padding); rows >= num_snps are zero-filled.  Returns 1 if a
 * file is missing or shorter than num_snps x 2 x W words. */
int episdet_read_planes(const char *path, int num_snps, int num_samples, int num_snps_alloc, uint32_t *planes);

#ifdef __cplusplus
}
#endif

#endif /* EPISDET_DATA_H_ */
