/*
 * episdet_oracle.c -- CPU oracle for exhaustive 2-way / 3-way epistasis search.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is loaded by the
 * tests, by __graft_entry__.smoke() and by the CPU-baseline leg of bench.py,
 * and by nothing else.  The product (the HIP library and the CLI) never links
 * or calls it.
 *
 * It restates, in plain C with OpenMP, the counting, cell derivation, scoring
 * and best-set rules of hiperbio/tensor-episdet:
 *
 *   - dataset layout: per class, u32 words [snp][plane][W]; plane 0 marks
 *     genotype 0, plane 1 genotype 1, genotype 2 is "neither" (helper.hpp:10-11,
 *     helper.cu:40,92).  Sample s sits in word s/32 at bit 31-(s%32): the last
 *     partial word keeps its valid samples in the high bits, because the mask
 *     the reference applies is ~0u << (32*ceil(N/32) - N) (helper.cu:72-73).
 *   - single-SNP counts c0, c1, c2 = N - c0 - c1            (helper.cu:13-52)
 *   - 2^k AND-counts plus derived cells: 2-way 4 + 5 cells   (helper.cu:634-662)
 *                                        3-way 8 + 19 cells  (helper.cu:480-548,
 *                                                             helper.hpp:34-51)
 *   - K2 score |sum_c lnG(ctl+1) + lnG(cas+1) - lnG(ctl+cas+2)|, minimised
 *                                                            (scoring.hpp:5-7,41-46,78-85;
 *                                                             search-triplets.cu:72-76)
 *   - MI score H(G) + H(Y) - H(G,Y) in bits, maximised      (scoring.hpp:9-16,22-37,57-74;
 *                                                             search-triplets.cu:63-69)
 *   - candidates only with X < Y < Z (A < B)                 (helper.cu:553-555,667-669)
 *   - packed index X | Y<<21 | Z<<42, A | B<<32              (helper.cu:576,690)
 *   - n_choose_k recurrence                                  (helper.cu:758-768)
 *
 * Intentional deviations from the reference, on purpose and documented:
 *   - the K2 log-factorial table covers every index the formula reaches
 *     (0 .. Nc+Nt+1); the reference's table stops at max(Nc,Nt) and reads past
 *     its end for large cells (search-triplets.cu:72-76);
 *   - the MI table is log2(m/N) in FP64 with the exact 1/N (the reference
 *     rounds 1/N to float first, search-triplets.cu:65);
 *   - ties in the best score are broken by the lowest packed index, so the
 *     winner is deterministic (the reference publishes whichever thread wins
 *     an atomic race, reduction.hpp:96-122);
 *   - any padding W >= ceil(N/32) is accepted (the reference's 3-way kernels
 *     need W % 128 == 0, helper.cu:237,260-287).
 *
 * FP64 scoring is written as a fixed sequence of IEEE operations (no FMA
 * contraction: build with -ffp-contract=off) so that the GPU's FP64
 * re-scoring, which follows the same sequence over the same tables, produces
 * bit-identical scores.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define ORC_K2 0
#define ORC_MI 1

/* ------------------------------------------------------------------ utils */

/* Exact binomial by the reference's recurrence (helper.cu:758-768). */
uint64_t orc_n_choose_k(unsigned n, unsigned k)
{
    uint64_t r = 1;
    for (unsigned i = 1; i <= k; i++) {
        r = r * n / i;
        n = n - 1;
    }
    return r;
}

static inline int bit_of(const uint32_t *row, int s)
{
    return (int)((row[s >> 5] >> (31 - (s & 31))) & 1u);
}

/* Genotype of every (snp, sample): 0, 1, 2, or 3 when both plane bits are set
 * (invalid input).  g has n*N bytes. */
void orc_decode(const uint32_t *planes, int n, int N, int W, uint8_t *g)
{
    for (int v = 0; v < n; v++) {
        const uint32_t *p0 = planes + (size_t)v * 2 * W;
        const uint32_t *p1 = p0 + W;
        for (int s = 0; s < N; s++) {
            int b0 = bit_of(p0, s), b1 = bit_of(p1, s);
            g[(size_t)v * N + s] = (uint8_t)(b0 && b1 ? 3 : (b0 ? 0 : (b1 ? 1 : 2)));
        }
    }
}

/* Number of format violations: samples with both plane bits set, and set
 * bits in the padding (positions >= N). */
int64_t orc_check_planes(const uint32_t *planes, int n, int N, int W)
{
    int64_t bad = 0;
    for (int v = 0; v < n; v++) {
        const uint32_t *p0 = planes + (size_t)v * 2 * W;
        const uint32_t *p1 = p0 + W;
        for (int w = 0; w < W; w++) {
            bad += __builtin_popcount(p0[w] & p1[w]);
            uint32_t valid;
            int lo = w * 32;
            if (lo + 32 <= N) valid = 0xFFFFFFFFu;
            else if (lo >= N) valid = 0u;
            else valid = 0xFFFFFFFFu << (32 - (N - lo));
            bad += __builtin_popcount((p0[w] | p1[w]) & ~valid);
        }
    }
    return bad;
}

/* Single-SNP counts, out[v*3+g]  (helper.cu:13-52). */
void orc_singles(const uint32_t *planes, int n, int N, int W, uint32_t *out)
{
    for (int v = 0; v < n; v++) {
        const uint32_t *p0 = planes + (size_t)v * 2 * W;
        const uint32_t *p1 = p0 + W;
        uint32_t c0 = 0, c1 = 0;
        for (int w = 0; w < W; w++) {
            c0 += (uint32_t)__builtin_popcount(p0[w]);
            c1 += (uint32_t)__builtin_popcount(p1[w]);
        }
        out[v * 3 + 0] = c0;
        out[v * 3 + 1] = c1;
        out[v * 3 + 2] = (uint32_t)N - c0 - c1;
    }
}

/* Brute-force contingency table from decoded genotypes: one pass over the
 * samples, cell index x*9+y*3+z (3-way) or a*3+b (2-way). */
void orc_table_direct(const uint8_t *g, int N, int order, const int *snps, uint32_t *cells)
{
    int nc = order == 3 ? 27 : 9;
    memset(cells, 0, sizeof(uint32_t) * nc);
    for (int s = 0; s < N; s++) {
        int c = 0;
        for (int o = 0; o < order; o++) c = c * 3 + g[(size_t)snps[o] * N + s];
        cells[c]++;
    }
}

static uint32_t and_count2(const uint32_t *a, const uint32_t *b, int W)
{
    uint32_t c = 0;
    for (int w = 0; w < W; w++) c += (uint32_t)__builtin_popcount(a[w] & b[w]);
    return c;
}

static uint32_t and_count3(const uint32_t *a, const uint32_t *b, const uint32_t *c_, int W)
{
    uint32_t c = 0;
    for (int w = 0; w < W; w++) c += (uint32_t)__builtin_popcount(a[w] & b[w] & c_[w]);
    return c;
}

/* 3x3 table of a pair from its four AND-counts and the singles
 * (the 5 derived cells of helper.cu:644-662). */
static void pair9_from4(const uint32_t and4[4], const uint32_t sa[3], const uint32_t sb[3], uint32_t t[9])
{
    t[0] = and4[0]; t[1] = and4[1];
    t[3] = and4[2]; t[4] = and4[3];
    t[2] = sa[0] - (t[0] + t[1]);
    t[5] = sa[1] - (t[3] + t[4]);
    t[6] = sb[0] - (t[0] + t[3]);
    t[7] = sb[1] - (t[1] + t[4]);
    t[8] = sa[2] - (t[6] + t[7]);
}

#define C3(x, y, z) ((x) * 9 + (y) * 3 + (z))

/* 27-cell table from the eight AND-counts and the three 3x3 pair tables, in
 * the derivation order of helper.cu:493-548 (each derived cell is a pair
 * marginal minus two cells that are already known). */
static void triplet27_from8(const uint32_t n8[8], const uint32_t xy[9], const uint32_t xz[9],
                            const uint32_t yz[9], uint32_t t[27])
{
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++)
            for (int k = 0; k < 2; k++) t[C3(i, j, k)] = n8[i * 4 + j * 2 + k];
    t[C3(0, 0, 2)] = xy[0 * 3 + 0] - (t[C3(0, 0, 0)] + t[C3(0, 0, 1)]);
    t[C3(0, 1, 2)] = xy[0 * 3 + 1] - (t[C3(0, 1, 0)] + t[C3(0, 1, 1)]);
    t[C3(0, 2, 0)] = xz[0 * 3 + 0] - (t[C3(0, 0, 0)] + t[C3(0, 1, 0)]);
    t[C3(0, 2, 1)] = xz[0 * 3 + 1] - (t[C3(0, 0, 1)] + t[C3(0, 1, 1)]);
    t[C3(0, 2, 2)] = xz[0 * 3 + 2] - (t[C3(0, 0, 2)] + t[C3(0, 1, 2)]);
    t[C3(1, 0, 2)] = xy[1 * 3 + 0] - (t[C3(1, 0, 0)] + t[C3(1, 0, 1)]);
    t[C3(1, 1, 2)] = xy[1 * 3 + 1] - (t[C3(1, 1, 0)] + t[C3(1, 1, 1)]);
    t[C3(1, 2, 0)] = xz[1 * 3 + 0] - (t[C3(1, 0, 0)] + t[C3(1, 1, 0)]);
    t[C3(1, 2, 1)] = xz[1 * 3 + 1] - (t[C3(1, 0, 1)] + t[C3(1, 1, 1)]);
    t[C3(1, 2, 2)] = xz[1 * 3 + 2] - (t[C3(1, 0, 2)] + t[C3(1, 1, 2)]);
    t[C3(2, 0, 0)] = yz[0 * 3 + 0] - (t[C3(0, 0, 0)] + t[C3(1, 0, 0)]);
    t[C3(2, 0, 1)] = yz[0 * 3 + 1] - (t[C3(0, 0, 1)] + t[C3(1, 0, 1)]);
    t[C3(2, 0, 2)] = xy[2 * 3 + 0] - (t[C3(2, 0, 0)] + t[C3(2, 0, 1)]);
    t[C3(2, 1, 0)] = yz[1 * 3 + 0] - (t[C3(0, 1, 0)] + t[C3(1, 1, 0)]);
    t[C3(2, 1, 1)] = yz[1 * 3 + 1] - (t[C3(0, 1, 1)] + t[C3(1, 1, 1)]);
    t[C3(2, 1, 2)] = xy[2 * 3 + 1] - (t[C3(2, 1, 0)] + t[C3(2, 1, 1)]);
    t[C3(2, 2, 0)] = xz[2 * 3 + 0] - (t[C3(2, 0, 0)] + t[C3(2, 1, 0)]);
    t[C3(2, 2, 1)] = xz[2 * 3 + 1] - (t[C3(2, 0, 1)] + t[C3(2, 1, 1)]);
    t[C3(2, 2, 2)] = xz[2 * 3 + 2] - (t[C3(2, 0, 2)] + t[C3(2, 1, 2)]);
}

static void pair9_bits(const uint32_t *planes, int N, int W, int a, int b, uint32_t t[9])
{
    const uint32_t *A = planes + (size_t)a * 2 * W, *B = planes + (size_t)b * 2 * W;
    uint32_t and4[4], sa[3], sb[3];
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) and4[i * 2 + j] = and_count2(A + i * W, B + j * W, W);
    orc_singles(A, 1, N, W, sa);
    orc_singles(B, 1, N, W, sb);
    pair9_from4(and4, sa, sb, t);
}

/* Contingency table by the bit-plane method: 2^k AND+popcount cells plus
 * 3^k - 2^k derived cells.  snps must be given in increasing order. */
void orc_table_bits(const uint32_t *planes, int N, int W, int order, const int *snps, uint32_t *cells)
{
    if (order == 2) {
        pair9_bits(planes, N, W, snps[0], snps[1], cells);
        return;
    }
    const uint32_t *X = planes + (size_t)snps[0] * 2 * W;
    const uint32_t *Y = planes + (size_t)snps[1] * 2 * W;
    const uint32_t *Z = planes + (size_t)snps[2] * 2 * W;
    uint32_t n8[8], xy[9], xz[9], yz[9];
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++)
            for (int k = 0; k < 2; k++) n8[i * 4 + j * 2 + k] = and_count3(X + i * W, Y + j * W, Z + k * W, W);
    pair9_bits(planes, N, W, snps[0], snps[1], xy);
    pair9_bits(planes, N, W, snps[0], snps[2], xz);
    pair9_bits(planes, N, W, snps[1], snps[2], yz);
    triplet27_from8(n8, xy, xz, yz, cells);
}

/* ---------------------------------------------------------------- scoring */

/* T[m] = ln(m!) = lgamma(m+1) for m in [0, mmax]  (search-triplets.cu:72-76,
 * extended to the full index range the K2 formula reaches). */
void orc_lnfact_table(int mmax, double *T)
{
    for (int m = 0; m <= mmax; m++) T[m] = lgamma((double)m + 1.0);
}

/* L[0] = 0, L[m] = log2(m / N) for m in [1, N]  (search-triplets.cu:63-69). */
void orc_log2_table(int N, double *L)
{
    L[0] = 0.0;
    for (int m = 1; m <= N; m++) L[m] = log2((double)m / (double)N);
}

/* K2 in FP64 (scoring.hpp:5-7 and 78-85): s += (T[ctl] + T[cas]) - T[ctl+cas+1]
 * over the cells in index order, score = |s|.  Lower is better. */
double orc_score_k2(const uint32_t *cas, const uint32_t *ctl, int ncells, const double *T)
{
    double s = 0.0;
    for (int c = 0; c < ncells; c++) {
        double t = T[ctl[c]] + T[cas[c]];
        t = t - T[ctl[c] + cas[c] + 1];
        s = s + t;
    }
    return fabs(s);
}

/* MI in FP64 (scoring.hpp:9-16 and 57-74): H(G) + H(Y) - H(G,Y) with
 * p = count / N and L[m] = log2(m/N).  Higher is better. */
double orc_score_mi(const uint32_t *cas, const uint32_t *ctl, int ncells, int Nc, int Nt, const double *L)
{
    double invN = 1.0 / (double)(Nc + Nt);
    double hY = -((double)Nc * invN) * L[Nc];
    hY = hY - ((double)Nt * invN) * L[Nt];
    double hG = 0.0, hGY = 0.0;
    for (int c = 0; c < ncells; c++) {
        double pc = (double)cas[c] * invN;
        double pt = (double)ctl[c] * invN;
        hGY = hGY - pc * L[cas[c]];
        hGY = hGY - pt * L[ctl[c]];
        hG = hG - (pc + pt) * L[cas[c] + ctl[c]];
    }
    return (hG + hY) - hGY;
}

/* ----------------------------------------------------------------- search */

uint64_t orc_table_hash(uint64_t rank, const uint32_t *cas, const uint32_t *ctl, int ncells);

static inline uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

typedef struct {
    double score;       /* FP64 score of the winner                          */
    uint64_t packed;    /* X | Y<<21 | Z<<42, or A | B<<32                   */
    int64_t sets;       /* number of sets evaluated                          */
    int32_t idx[3];
} orc_result;

static inline int better(int score, double s, uint64_t p, double bs, uint64_t bp)
{
    if (score == ORC_K2) {
        if (s < bs) return 1;
        if (s > bs) return 0;
    } else {
        if (s > bs) return 1;
        if (s < bs) return 0;
    }
    return p < bp;
}

/* 64-bit packed copies of one class: [snp][plane][W64], zero padded. */
static uint64_t *pack64(const uint32_t *planes, int n, int W, int *W64out)
{
    int W64 = (W + 1) / 2;
    uint64_t *q = (uint64_t *)calloc((size_t)n * 2 * W64 + 8, sizeof(uint64_t));
    for (int v = 0; v < n; v++)
        for (int p = 0; p < 2; p++) {
            const uint32_t *src = planes + ((size_t)v * 2 + p) * W;
            uint64_t *dst = q + ((size_t)v * 2 + p) * W64;
            for (int w = 0; w < W; w++) dst[w >> 1] |= (uint64_t)src[w] << ((w & 1) ? 0 : 32);
        }
    *W64out = W64;
    return q;
}

static inline uint32_t popc_and(const uint64_t *a, const uint64_t *b, int W64)
{
    uint64_t c = 0;
    for (int w = 0; w < W64; w++) c += (uint64_t)__builtin_popcountll(a[w] & b[w]);
    return (uint32_t)c;
}

/* Two X rows against one Q at once: for X in {a, b} (plane 0 at x, plane 1
 * at x + W64), out[i*4+jk] = popcount(X_i & Q_jk), i in {0,1}, jk in 0..3,
 * in one pass over the words.  Q holds four rows of qs words each.  Each Q
 * word is loaded once for both rows. */
static void count16_scalar(const uint64_t *xa, const uint64_t *xb, const uint64_t *Q, size_t qs, int W64,
                           uint32_t *oa, uint32_t *ob)
{
    uint64_t c[16] = {0};
    for (int w = 0; w < W64; w++) {
        const uint64_t a0 = xa[w], a1 = xa[W64 + w], b0 = xb[w], b1 = xb[W64 + w];
        for (int jk = 0; jk < 4; jk++) {
            const uint64_t q = Q[jk * qs + w];
            c[jk] += (uint64_t)__builtin_popcountll(a0 & q);
            c[4 + jk] += (uint64_t)__builtin_popcountll(a1 & q);
            c[8 + jk] += (uint64_t)__builtin_popcountll(b0 & q);
            c[12 + jk] += (uint64_t)__builtin_popcountll(b1 & q);
        }
    }
    for (int i = 0; i < 8; i++) {
        oa[i] = (uint32_t)c[i];
        ob[i] = (uint32_t)c[8 + i];
    }
}

#if defined(__x86_64__)
#include <immintrin.h>
__attribute__((target("avx512f,avx512vpopcntdq")))
static void count16_avx512(const uint64_t *xa, const uint64_t *xb, const uint64_t *Q, size_t qs, int W64,
                           uint32_t *oa, uint32_t *ob)
{
    __m512i c[16];
    for (int i = 0; i < 16; i++) c[i] = _mm512_setzero_si512();
    int w = 0;
    for (; w + 8 <= W64; w += 8) {
        const __m512i a0 = _mm512_loadu_si512((const void *)(xa + w));
        const __m512i a1 = _mm512_loadu_si512((const void *)(xa + W64 + w));
        const __m512i b0 = _mm512_loadu_si512((const void *)(xb + w));
        const __m512i b1 = _mm512_loadu_si512((const void *)(xb + W64 + w));
        for (int jk = 0; jk < 4; jk++) {
            const __m512i q = _mm512_loadu_si512((const void *)(Q + jk * qs + w));
            c[jk] = _mm512_add_epi64(c[jk], _mm512_popcnt_epi64(_mm512_and_si512(a0, q)));
            c[4 + jk] = _mm512_add_epi64(c[4 + jk], _mm512_popcnt_epi64(_mm512_and_si512(a1, q)));
            c[8 + jk] = _mm512_add_epi64(c[8 + jk], _mm512_popcnt_epi64(_mm512_and_si512(b0, q)));
            c[12 + jk] = _mm512_add_epi64(c[12 + jk], _mm512_popcnt_epi64(_mm512_and_si512(b1, q)));
        }
    }
    uint64_t t[16];
    for (int i = 0; i < 16; i++) t[i] = (uint64_t)_mm512_reduce_add_epi64(c[i]);
    for (; w < W64; w++) {
        const uint64_t a0 = xa[w], a1 = xa[W64 + w], b0 = xb[w], b1 = xb[W64 + w];
        for (int jk = 0; jk < 4; jk++) {
            const uint64_t q = Q[jk * qs + w];
            t[jk] += (uint64_t)__builtin_popcountll(a0 & q);
            t[4 + jk] += (uint64_t)__builtin_popcountll(a1 & q);
            t[8 + jk] += (uint64_t)__builtin_popcountll(b0 & q);
            t[12 + jk] += (uint64_t)__builtin_popcountll(b1 & q);
        }
    }
    for (int i = 0; i < 8; i++) {
        oa[i] = (uint32_t)t[i];
        ob[i] = (uint32_t)t[8 + i];
    }
}
#endif

typedef void (*count16_fn)(const uint64_t *, const uint64_t *, const uint64_t *, size_t, int, uint32_t *,
                           uint32_t *);

static count16_fn pick_count16(void)
{
#if defined(__x86_64__)
    __builtin_cpu_init();
    if (__builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512vpopcntdq")) return count16_avx512;
#endif
    return count16_scalar;
}

/* 1 when the AVX-512 VPOPCNTDQ counting loop is in use on this host. */
int orc_uses_avx512(void) { return pick_count16() != count16_scalar; }

typedef struct {
    int n, N[2], W64[2];
    uint64_t *q[2];      /* packed planes per class                 */
    uint32_t *single;    /* [class][snp][3]                         */
    uint32_t *pair;      /* [class][b][a][4] AND counts, a < b < pn: contiguous in the
                            smaller index, which the search's innermost loop walks */
    int pn;
} ctx_t;

static void ctx_pairs(ctx_t *cx, int pn)
{
    cx->pn = pn;
    cx->pair = (uint32_t *)malloc(sizeof(uint32_t) * 2 * (size_t)pn * pn * 4);
    for (int c = 0; c < 2; c++) {
        int W64 = cx->W64[c];
        const uint64_t *q = cx->q[c];
        #pragma omp parallel for schedule(dynamic, 4)
        for (int a = 0; a < pn; a++)
            for (int b = a + 1; b < pn; b++) {
                uint32_t *o = cx->pair + (((size_t)c * pn + b) * pn + a) * 4;
                for (int i = 0; i < 2; i++)
                    for (int j = 0; j < 2; j++)
                        o[i * 2 + j] = popc_and(q + ((size_t)a * 2 + i) * W64, q + ((size_t)b * 2 + j) * W64, W64);
            }
    }
}

static void ctx_pair9(const ctx_t *cx, int c, int a, int b, uint32_t t[9])
{
    pair9_from4(cx->pair + (((size_t)c * cx->pn + b) * cx->pn + a) * 4, cx->single + ((size_t)c * cx->n + a) * 3,
                cx->single + ((size_t)c * cx->n + b) * 3, t);
}

static double score_tables(int order, int score, const uint32_t *cas, const uint32_t *ctl, const ctx_t *cx,
                           const double *tab)
{
    int nc = order == 3 ? 27 : 9;
    if (score == ORC_K2) return orc_score_k2(cas, ctl, nc, tab);
    return orc_score_mi(cas, ctl, nc, cx->N[0], cx->N[1], tab);
}

/* Eight sets at a time: the per-set operation sequence is exactly that of
 * orc_score_k2 / orc_score_mi / orc_table_hash (so every result is the same
 * to the bit), with the eight independent dependency chains interleaved. */
#define ILV 8

static void score8_k2(const uint32_t cas[ILV][27], const uint32_t ctl[ILV][27], int ncells, const double *T,
                      double out[ILV])
{
    double s[ILV];
    for (int j = 0; j < ILV; j++) s[j] = 0.0;
    for (int c = 0; c < ncells; c++)
        for (int j = 0; j < ILV; j++) {
            double t = T[ctl[j][c]] + T[cas[j][c]];
            t = t - T[ctl[j][c] + cas[j][c] + 1];
            s[j] = s[j] + t;
        }
    for (int j = 0; j < ILV; j++) out[j] = fabs(s[j]);
}

static void score8_mi(const uint32_t cas[ILV][27], const uint32_t ctl[ILV][27], int ncells, int Nc, int Nt,
                      const double *L, double out[ILV])
{
    double invN = 1.0 / (double)(Nc + Nt);
    double hY = -((double)Nc * invN) * L[Nc];
    hY = hY - ((double)Nt * invN) * L[Nt];
    double hG[ILV], hGY[ILV];
    for (int j = 0; j < ILV; j++) hG[j] = hGY[j] = 0.0;
    for (int c = 0; c < ncells; c++)
        for (int j = 0; j < ILV; j++) {
            double pc = (double)cas[j][c] * invN;
            double pt = (double)ctl[j][c] * invN;
            hGY[j] = hGY[j] - pc * L[cas[j][c]];
            hGY[j] = hGY[j] - pt * L[ctl[j][c]];
            hG[j] = hG[j] - (pc + pt) * L[cas[j][c] + ctl[j][c]];
        }
    for (int j = 0; j < ILV; j++) out[j] = (hG[j] + hY) - hGY[j];
}

static void hash8(const uint64_t rank[ILV], const uint32_t cas[ILV][27], const uint32_t ctl[ILV][27], int ncells,
                  uint64_t out[ILV])
{
    uint64_t h[ILV];
    for (int j = 0; j < ILV; j++) h[j] = (rank[j] + 1) * 0x9E3779B97F4A7C15ull;
    for (int c = 0; c < ncells; c++)
        for (int j = 0; j < ILV; j++) h[j] = mix64(h[j] ^ cas[j][c]);
    for (int c = 0; c < ncells; c++)
        for (int j = 0; j < ILV; j++) h[j] = mix64(h[j] ^ ctl[j][c]);
    for (int j = 0; j < ILV; j++) out[j] = h[j];
}

/* Exhaustive search.  cases/controls: [n][2][Wc] / [n][2][Wt] u32 planes.
 * Only sets whose largest index lies in [last_lo, last_hi) are evaluated, so
 * a contiguous slice of the search can be timed; [0, n) is the full search.
 * nthreads <= 0 uses the OpenMP default.  checksum, when not NULL, receives
 * the order-independent table checksum (orc_table_hash summed mod 2^64) of
 * every evaluated set, the same quantity orc_all_tables returns, so a full
 * search pins every table it scored as well as the winner.  Returns 0 on
 * success. */
int orc_search_ex(const uint32_t *cases, const uint32_t *controls, int n, int Nc, int Nt, int Wc, int Wt, int order,
                  int score, int nthreads, int last_lo, int last_hi, orc_result *out, uint64_t *checksum)
{
    if (order != 2 && order != 3) return 1;
    if (n < order || last_lo < 0 || last_hi > n || last_lo >= last_hi) return 1;
    if (nthreads > 0) omp_set_num_threads(nthreads);
    ctx_t cx;
    memset(&cx, 0, sizeof cx);
    cx.n = n;
    cx.N[0] = Nc;
    cx.N[1] = Nt;
    cx.q[0] = pack64(cases, n, Wc, &cx.W64[0]);
    cx.q[1] = pack64(controls, n, Wt, &cx.W64[1]);
    cx.single = (uint32_t *)malloc(sizeof(uint32_t) * 2 * (size_t)n * 3);
    orc_singles(cases, n, Nc, Wc, cx.single);
    orc_singles(controls, n, Nt, Wt, cx.single + (size_t)n * 3);
    if (order == 3) ctx_pairs(&cx, last_hi);

    int N = Nc + Nt;
    double *tab;
    if (score == ORC_K2) {
        tab = (double *)malloc(sizeof(double) * (size_t)(N + 2));
        orc_lnfact_table(N + 1, tab);
    } else {
        tab = (double *)malloc(sizeof(double) * (size_t)(N + 1));
        orc_log2_table(N, tab);
    }

    double best = score == ORC_K2 ? INFINITY : -INFINITY;
    uint64_t bestp = UINT64_MAX;
    int64_t sets = 0;
    uint64_t hsum = 0;
    const int want_hash = checksum != NULL;

    count16_fn count16 = pick_count16();
    #pragma omp parallel
    {
        double tb = score == ORC_K2 ? INFINITY : -INFINITY;
        uint64_t tp = UINT64_MAX;
        int64_t ts = 0;
        uint64_t th = 0;
        int W64c = cx.W64[0], W64t = cx.W64[1];
        size_t qmax = (size_t)(W64c > W64t ? W64c : W64t);
        uint64_t *Q = (uint64_t *)malloc(sizeof(uint64_t) * 4 * qmax);
        /* X rows per block of the 3-way loop: both planes of both classes of
         * XB SNPs in about 1 MiB, a multiple of ILV between ILV and 64 */
        int XB = (int)((1u << 20) / ((size_t)(W64c + W64t) * 2 * sizeof(uint64_t)));
        XB = XB / ILV * ILV;
        if (XB < ILV) XB = ILV;
        if (XB > 64) XB = 64;
        #pragma omp for schedule(dynamic, 1)
        for (int hi = last_hi - 1; hi >= last_lo; hi--) {
            if (order == 2) {
                int B = hi;
                for (int A = 0; A < B; A++) {
                    uint32_t cas[9], ctl[9];
                    for (int c = 0; c < 2; c++) {
                        int W64 = cx.W64[c];
                        const uint64_t *q = cx.q[c];
                        uint32_t and4[4];
                        for (int i = 0; i < 2; i++)
                            for (int j = 0; j < 2; j++)
                                and4[i * 2 + j] =
                                    popc_and(q + ((size_t)A * 2 + i) * W64, q + ((size_t)B * 2 + j) * W64, W64);
                        pair9_from4(and4, cx.single + ((size_t)c * n + A) * 3, cx.single + ((size_t)c * n + B) * 3,
                                    c == 0 ? cas : ctl);
                    }
                    double s = score_tables(2, score, cas, ctl, &cx, tab);
                    uint64_t p = (uint64_t)A | ((uint64_t)B << 32);
                    if (want_hash) th += orc_table_hash((uint64_t)B * (B - 1) / 2 + (uint64_t)A, cas, ctl, 9);
                    if (better(score, s, p, tb, tp)) { tb = s; tp = p; }
                    ts++;
                }
            } else {
                int Z = hi;
                const uint64_t rz3 = (uint64_t)Z * (Z - 1) * (Z - 2) / 6;   /* colex rank base C(Z,3) */
                /* Blocked over X: a block of XB X rows (both classes, about
                 * 1 MiB) stays in the core's cache while every Y above it is
                 * contracted against it; X rows are then read from memory
                 * once per (block, Z) instead of once per (Y, Z). */
                for (int X0 = 0; X0 < Z - 1; X0 += XB) {
                    const int XE = X0 + XB < Z - 1 ? X0 + XB : Z - 1;
                    for (int Y = X0 + 1; Y < Z; Y++) {
                        const int X1 = XE < Y ? XE : Y;      /* X in [X0, X1), X < Y */
                        uint32_t n8[2][64][8];
                        uint32_t yz[2][9];
                        for (int c = 0; c < 2; c++) ctx_pair9(&cx, c, Y, Z, yz[c]);
                        for (int c = 0; c < 2; c++) {
                            /* Q[j*2+k] = Y_j & Z_k of this class */
                            int W64 = cx.W64[c];
                            const uint64_t *q = cx.q[c];
                            for (int j = 0; j < 2; j++)
                                for (int k = 0; k < 2; k++) {
                                    const uint64_t *yj = q + ((size_t)Y * 2 + j) * W64;
                                    const uint64_t *zk = q + ((size_t)Z * 2 + k) * W64;
                                    uint64_t *d = Q + (size_t)(j * 2 + k) * qmax;
                                    for (int w = 0; w < W64; w++) d[w] = yj[w] & zk[w];
                                }
                            for (int X = X0; X < X1; X += 2) {
                                /* an odd last row is paired with itself */
                                const int Xb = X + 1 < X1 ? X + 1 : X;
                                count16(q + (size_t)X * 2 * W64, q + (size_t)Xb * 2 * W64, Q, qmax, W64,
                                        n8[c][X - X0], n8[c][Xb - X0]);
                            }
                        }
                        for (int Xb = X0; Xb < X1; Xb += ILV) {
                            /* ILV sets at once (the last group repeats its final X) */
                            const int m = X1 - Xb < ILV ? X1 - Xb : ILV;
                            uint32_t cas[ILV][27], ctl[ILV][27];
                            uint64_t rk[ILV], hv[ILV];
                            double sv[ILV];
                            for (int j = 0; j < ILV; j++) {
                                const int X = Xb + (j < m ? j : m - 1);
                                uint32_t xy[9], xz[9];
                                ctx_pair9(&cx, 0, X, Y, xy);
                                ctx_pair9(&cx, 0, X, Z, xz);
                                triplet27_from8(n8[0][X - X0], xy, xz, yz[0], cas[j]);
                                ctx_pair9(&cx, 1, X, Y, xy);
                                ctx_pair9(&cx, 1, X, Z, xz);
                                triplet27_from8(n8[1][X - X0], xy, xz, yz[1], ctl[j]);
                                rk[j] = rz3 + (uint64_t)Y * (Y - 1) / 2 + (uint64_t)X;
                            }
                            if (score == ORC_K2) score8_k2(cas, ctl, 27, tab, sv);
                            else score8_mi(cas, ctl, 27, cx.N[0], cx.N[1], tab, sv);
                            if (want_hash) {
                                hash8(rk, cas, ctl, 27, hv);
                                for (int j = 0; j < m; j++) th += hv[j];
                            }
                            for (int j = 0; j < m; j++) {
                                const int X = Xb + j;
                                uint64_t p = (uint64_t)X | ((uint64_t)Y << 21) | ((uint64_t)Z << 42);
                                if (better(score, sv[j], p, tb, tp)) { tb = sv[j]; tp = p; }
                            }
                            ts += m;
                        }
                    }
                }
            }
        }
        free(Q);
        #pragma omp critical
        {
            if (better(score, tb, tp, best, bestp)) { best = tb; bestp = tp; }
            sets += ts;
            hsum += th;
        }
    }
    if (checksum) *checksum = hsum;

    out->score = best;
    out->packed = bestp;
    out->sets = sets;
    if (order == 3) {
        out->idx[0] = (int32_t)(bestp & 0x1FFFFF);
        out->idx[1] = (int32_t)((bestp >> 21) & 0x1FFFFF);
        out->idx[2] = (int32_t)((bestp >> 42) & 0x1FFFFF);
    } else {
        out->idx[0] = (int32_t)(bestp & 0xFFFFFFFFu);
        out->idx[1] = (int32_t)(bestp >> 32);
        out->idx[2] = -1;
    }
    free(tab);
    free(cx.q[0]);
    free(cx.q[1]);
    free(cx.single);
    free(cx.pair);
    return 0;
}

int orc_search(const uint32_t *cases, const uint32_t *controls, int n, int Nc, int Nt, int Wc, int Wt, int order,
               int score, int nthreads, int last_lo, int last_hi, orc_result *out)
{
    return orc_search_ex(cases, controls, n, Nc, Nt, Wc, Wt, order, score, nthreads, last_lo, last_hi, out, NULL);
}

/* ---------------------------------------------------- full-table checking */


/* Order-independent checksum of every set's table:
 *   h = (rank+1) * 0x9E3779B97F4A7C15; h = mix64(h ^ cell) over the cases
 *   cells then the controls cells in index order; checksum = sum of h mod 2^64,
 * where rank is the colex rank of the set (C(Z,3)+C(Y,2)+X, or C(B,2)+A). */
uint64_t orc_table_hash(uint64_t rank, const uint32_t *cas, const uint32_t *ctl, int ncells)
{
    uint64_t h = (rank + 1) * 0x9E3779B97F4A7C15ull;
    for (int c = 0; c < ncells; c++) h = mix64(h ^ cas[c]);
    for (int c = 0; c < ncells; c++) h = mix64(h ^ ctl[c]);
    return h;
}

/* Every set's table, in colex rank order: out[rank][2][3^k] (cases then
 * controls).  For small n only.  Also returns the checksum above. */
uint64_t orc_all_tables(const uint32_t *cases, const uint32_t *controls, int n, int Nc, int Nt, int Wc, int Wt,
                        int order, uint32_t *out)
{
    int nc = order == 3 ? 27 : 9;
    int64_t total = (int64_t)orc_n_choose_k((unsigned)n, (unsigned)order);
    uint64_t sum = 0;
    #pragma omp parallel for schedule(dynamic, 64) reduction(+ : sum)
    for (int64_t r = 0; r < total; r++) {
        int s[3];
        /* colex unrank */
        int64_t rem = r;
        int hi = n;
        for (int o = order; o >= 1; o--) {
            int v = hi - 1;
            while ((int64_t)orc_n_choose_k((unsigned)v, (unsigned)o) > rem) v--;
            s[o - 1] = v;
            rem -= (int64_t)orc_n_choose_k((unsigned)v, (unsigned)o);
            hi = v;
        }
        uint32_t cas[27], ctl[27];
        orc_table_bits(cases, Nc, Wc, order, s, cas);
        orc_table_bits(controls, Nt, Wt, order, s, ctl);
        if (out) {
            memcpy(out + (size_t)r * 2 * nc, cas, sizeof(uint32_t) * nc);
            memcpy(out + (size_t)r * 2 * nc + nc, ctl, sizeof(uint32_t) * nc);
        }
        sum += orc_table_hash((uint64_t)r, cas, ctl, nc);
    }
    return sum;
}
