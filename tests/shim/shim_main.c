/* shim_main.c -- a C caller written against the reference's entry point
 * EpistasisDetectionSearch (search-triplets.hpp:4), linked with
 * episdet_shim.cpp instead of the reference's CUDA objects: the check that
 * a reference-side program keeps its stdout, the "Outer loop iteration"
 * lines included, when only the link line changes.
 *
 *   shim_main.<variant> numSNPs numCases numControls seed
 *
 * The dataset comes from the seeded generator (one planted set of the
 * variant's order), in the reference's layout: rows padded to a multiple of
 * 256 SNPs, planes padded to 1024 samples. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "episdet_data.h"

int EpistasisDetectionSearch(unsigned *cases, unsigned *controls, int numSNPs, int numCases, int numControls,
                             unsigned numSNPsWithPadding, int numCasesWithPadding, int numControlsWithPadding,
                             int *roundsCounter, double *searchTime, float *outputFromGpu,
                             unsigned long long *packedIndices);

#ifdef TRIPLETS
#define ORDER 3
#else
#define ORDER 2
#endif

int main(int argc, char **argv)
{
    if (argc != 5) {
        fprintf(stderr, "usage: %s numSNPs numCases numControls seed\n", argv[0]);
        return 1;
    }
    const int n = atoi(argv[1]), nc = atoi(argv[2]), nt = atoi(argv[3]);
    episdet_gen_params gp;
    episdet_gen_defaults(&gp);
    gp.num_snps = n;
    gp.num_cases = nc;
    gp.num_controls = nt;
    gp.seed = strtoull(argv[4], NULL, 10);
    gp.planted_sets = 1;
    gp.planted_order = ORDER;
    const int wc = episdet_padded_words(nc), wt = episdet_padded_words(nt);
    const unsigned npad = (unsigned)((n + 255) / 256 * 256);
    uint32_t *cases = calloc((size_t)npad * 2 * wc, sizeof(uint32_t));
    uint32_t *controls = calloc((size_t)npad * 2 * wt, sizeof(uint32_t));
    int planted[3];
    if (!cases || !controls || episdet_generate(&gp, cases, controls, planted, NULL) != 0) {
        fprintf(stderr, "generation failed\n");
        return 1;
    }
    int rounds = 0;
    double secs = 0;
    float best = 0.f;
    unsigned long long packed = 0;
    const int rc = EpistasisDetectionSearch(cases, controls, n, nc, nt, npad, wc * 32, wt * 32, &rounds, &secs, &best,
                                            &packed);
    if (rc != 0) {
        fprintf(stderr, "EpistasisDetectionSearch failed: %d\n", rc);
        return 1;
    }
#ifdef TRIPLETS
    printf("{SNP_X_i, SNP_Y_i, SNP_Z_i}: SCORE\t->\t{%llu, %llu, %llu}: %.6f\n", packed & 0x1FFFFF,
           (packed >> 21) & 0x1FFFFF, (packed >> 42) & 0x1FFFFF, best);
#else
    printf("{SNP_X_i, SNP_Y_i}: SCORE\t->\t{%llu, %llu}: %.6f\n", packed & 0xFFFFFFFFull, packed >> 32, best);
#endif
    printf("Num. of rounds processed: %d\n", rounds);
    printf("planted:");
    for (int i = 0; i < ORDER; i++) printf(" %d", planted[i]);
    printf("\n");
    free(cases);
    free(controls);
    return 0;
}
