// Host-side check of the work-unit orders of the MFMA pair kernels, built
// from the library's own headers: every unit of the triangular (tile 1) and
// B-tile-major (tile 0) enumerations appears exactly once in the XCD-blocked
// orders that replace them (blocked_unrank, p2_blocked_unrank), and the unit
// counts (blocked_units, p2_blocked_units) cover every block.  Prints
// "ok <checked>" or the first mismatches, and exits non-zero on a mismatch.
#include <cstdio>
#include <vector>
#include "kernels_mfma_pairs.hpp"

using namespace episdet;

int main()
{
    long long bad = 0, checked = 0;
    // tile 1: T x T triangle of 64-SNP tiles (ta <= tb)
    for (int T = 1; T <= 400; T++) {
        std::vector<int> seen((size_t)T * T, 0);
        for (unsigned long long u = 0; u < blocked_units(T); u++) {
            int a, b;
            if (blocked_unrank(u, T, a, b)) {
                if (a < 0 || b < 0 || a >= T || b >= T) {
                    bad++;
                    continue;
                }
                seen[(size_t)a * T + b]++;
            }
        }
        for (int a = 0; a < T; a++)
            for (int b = 0; b < T; b++) {
                checked++;
                if (seen[(size_t)a * T + b] != (a <= b ? 1 : 0) && bad++ < 5)
                    std::printf("tile 1, T=%d: unit (%d, %d) seen %d times\n", T, a, b, seen[(size_t)a * T + b]);
            }
    }
    // tile 0: 128-SNP A tiles x 64-SNP B tiles with some A < B (ta <= tb / 2)
    for (int n = 2; n <= 20000; n += n < 2000 ? 1 : 97) {
        const int ntb = (n + P2_TB - 1) / P2_TB, nta = (n + P2_TA - 1) / P2_TA;
        std::vector<int> seen((size_t)nta * ntb, 0);
        for (unsigned long long u = 0; u < p2_blocked_units(n); u++) {
            int a, b;
            if (p2_blocked_unrank(u, n, a, b)) {
                if (a < 0 || b < 0 || a >= nta || b >= ntb) {
                    bad++;
                    continue;
                }
                seen[(size_t)a * ntb + b]++;
            }
        }
        for (int a = 0; a < nta; a++)
            for (int b = 0; b < ntb; b++) {
                checked++;
                if (seen[(size_t)a * ntb + b] != (a <= b / 2 ? 1 : 0) && bad++ < 5)
                    std::printf("tile 0, n=%d: unit (%d, %d) seen %d times\n", n, a, b, seen[(size_t)a * ntb + b]);
            }
    }
    if (bad) {
        std::printf("%lld mismatches\n", bad);
        return 1;
    }
    std::printf("ok %lld\n", checked);
    return 0;
}
