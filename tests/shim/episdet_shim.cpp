// episdet_shim.cpp -- the reference's entry point on top of the C ABI.
//
// hiperbio/tensor-episdet links EpistasisDetectionSearch (search-pairs.hpp:4,
// search-triplets.hpp:4) from one of four CUDA builds, chosen by Makefile
// target (Makefile:8-22).  This shim gives a caller written against that
// entry point the same symbol, compiled once per variant with -DTRIPLETS and
// -DMI_SCORE exactly as the reference's targets select them, on top of
// episdet_search.  The return type is int instead of cudaError_t; 0 is
// success in both.
#include "episdet.h"

extern "C" int EpistasisDetectionSearch(unsigned *cases, unsigned *controls, int numSNPs, int numCases,
                                        int numControls, unsigned numSNPsWithPadding, int numCasesWithPadding,
                                        int numControlsWithPadding, int *roundsCounter, double *searchTime,
                                        float *outputFromGpu, unsigned long long *packedIndices)
{
#if defined(TRIPLETS)
    const int order = 3;
#else
    const int order = 2;
#endif
#if defined(MI_SCORE)
    const int score = EPISDET_SCORE_MI;
#else
    const int score = EPISDET_SCORE_K2;
#endif
    return episdet_search(order, score, cases, controls, numSNPs, numCases, numControls, numSNPsWithPadding,
                          numCasesWithPadding, numControlsWithPadding, roundsCounter, searchTime, outputFromGpu,
                          packedIndices);
}
