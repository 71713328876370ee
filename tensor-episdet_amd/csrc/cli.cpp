// cli.cpp -- tensor-episdet.{pairs,triplets}.{k2,mi}.bin <dataset.txt>
//
// Drop-in for the reference binaries (tensor-episdet.cu:32-206, Makefile:8-22):
// same single argument, same file handling, same stdout lines.  Built four
// times with EPISDET_ORDER (2|3) and EPISDET_SCORE (0 = K2, 1 = MI).
//
// Deviation: a missing or short .bin file is an error (exit 1) instead of a
// message followed by a search over uninitialised memory.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "episdet.h"
#include "episdet_data.h"

#ifndef EPISDET_ORDER
#define EPISDET_ORDER 3
#endif
#ifndef EPISDET_SCORE
#define EPISDET_SCORE 0
#endif

static const int BLOCK_SIZE = 256;   // reference Makefile:2, used for the reported metrics
static const int COMB_CALC = EPISDET_ORDER == 3 ? 8 : 4;

int main(int argc, const char *argv[])
{
    char devname[256];
    int cu = 0, clk = 0;
    if (episdet_device_info(devname, sizeof devname, &cu, &clk) != EPISDET_OK) {
        std::cerr << "Could not get device properties: no device" << std::endl;
        return 1;
    }
    if (std::string(devname).find("gfx950") == std::string::npos) {
        std::cerr << "A gfx950 (MI355X) device is required." << std::endl;
        return 1;
    }
    if (argc < 2) {
        std::cerr << "Usage: tensor-episdet dataset.txt" << std::endl;
        return 1;
    }
    int numSNPs = 0, numCases = 0, numControls = 0;
    char casesPath[4096], controlsPath[4096];
    {
        FILE *f = std::fopen(argv[1], "r");
        if (!f) {
            std::cerr << "File '" << argv[1] << "' does not exist!" << std::endl;
            return 1;
        }
        std::fclose(f);
    }
    if (episdet_read_header(argv[1], &numSNPs, &numCases, &numControls, casesPath, controlsPath,
                            (int)sizeof casesPath) != 0) {
        std::cerr << "Malformed dataset description '" << argv[1] << "'" << std::endl;
        return 1;
    }
    const unsigned numBlocks = (unsigned)std::ceil((float)numSNPs / (float)BLOCK_SIZE);
    const unsigned numSNPsWithPadding = numBlocks * BLOCK_SIZE;
    const int numCasesWithPadding = episdet_padded_words(numCases) * 32;
    const int numControlsWithPadding = episdet_padded_words(numControls) * 32;

    std::cout << "Num. SNPs: " << numSNPs << std::endl;
    std::cout << "Num. Blocks of SNPs: " << numBlocks << std::endl;
    std::cout << "Num. Cases: " << numCases << std::endl;
    std::cout << "Num. Controls: " << numControls << std::endl;

    std::vector<uint32_t> cases((size_t)numSNPsWithPadding * 2 * (numCasesWithPadding / 32));
    std::vector<uint32_t> controls((size_t)numSNPsWithPadding * 2 * (numControlsWithPadding / 32));
    if (episdet_read_planes(casesPath, numSNPs, numCases, (int)numSNPsWithPadding, cases.data()) != 0) {
        std::cerr << "Problem loading cases from storage device" << std::endl;
        return 1;
    }
    if (episdet_read_planes(controlsPath, numSNPs, numControls, (int)numSNPsWithPadding, controls.data()) != 0) {
        std::cerr << "Problem loading controls from storage device" << std::endl;
        return 1;
    }
    std::cout << "-------------------------------" << std::endl;
    // The "Outer loop iteration" lines come from episdet_search itself, as
    // they come from EpistasisDetectionSearch in the reference.

    int roundsCounter = 0;
    double searchTime = 0;
    float best = 0.f;
    unsigned long long packed = 0;
    const int rc = episdet_search(EPISDET_ORDER, EPISDET_SCORE, cases.data(), controls.data(), numSNPs, numCases,
                                  numControls, numSNPsWithPadding, numCasesWithPadding, numControlsWithPadding,
                                  &roundsCounter, &searchTime, &best, &packed);
    if (rc != EPISDET_OK) std::cerr << "Epistasis detection search failed." << std::endl;

#if EPISDET_ORDER == 3
    std::cout << "-------------------------------" << std::endl
              << "{SNP_X_i, SNP_Y_i, SNP_Z_i}: SCORE\t->\t{" << ((packed >> 0) & 0x1FFFFF) << ", "
              << ((packed >> 21) & 0x1FFFFF) << ", " << ((packed >> 42) & 0x1FFFFF) << "}: " << std::fixed
              << std::setprecision(6) << best << std::endl;
#else
    std::cout << "-------------------------------" << std::endl
              << "{SNP_X_i, SNP_Y_i}: SCORE\t->\t{" << ((packed >> 0) & 0xFFFFFFFF) << ", "
              << ((packed >> 32) & 0xFFFFFFFF) << "}: " << std::fixed << std::setprecision(6) << best << std::endl;
#endif
    const unsigned long long numCombinations = episdet_n_choose_k((unsigned)numSNPs, EPISDET_ORDER);
    std::cout << "Num. of rounds processed: " << roundsCounter << std::endl;
    std::cout << "Wall-clock time:\t" << std::fixed << std::setprecision(3) << searchTime << " seconds" << std::endl;
    std::cout << "Tensor TOPS: " << std::fixed << std::setprecision(3)
              << ((((double)BLOCK_SIZE * (double)BLOCK_SIZE * (double)(numCasesWithPadding + numControlsWithPadding) *
                    2 * COMB_CALC * (double)roundsCounter) /
                   searchTime) /
                  1e12)
              << std::endl;
    std::cout << "Num. unique sets per sec. (scaled to sample size): " << std::fixed << std::setprecision(3)
              << (((double)numCombinations * (double)(numCases + numControls) / searchTime) / 1e12) << " × 10^12"
              << std::endl;
    std::cout << "Unique sets of SNPs evaluated (k=" << EPISDET_ORDER << "): " << numCombinations << std::endl;
    std::cout << "Ratio of unique sets of SNPs: "
              << (((double)numCombinations) / (((double)roundsCounter) * (BLOCK_SIZE * BLOCK_SIZE))) << std::endl;
    return rc == EPISDET_OK ? 0 : 1;
}
