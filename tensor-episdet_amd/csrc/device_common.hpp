// device_common.hpp -- device-side building blocks shared by every search
// kernel: contingency-cell derivation, FP32 scoring (fused paths), FP64
// scoring (plain path and candidate re-scoring), order-preserving keys.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace episdet {

enum { SCORE_K2 = 0, SCORE_MI = 1 };

// Scoring context for the FP32 fused paths.
struct ScoreCtx {
    int score;          // SCORE_K2 / SCORE_MI
    int ncells;         // 9 or 27
    float k2_total;     // K2: ncells * c, c = ln N - 1
    float k0;           // K2: ln(N)/2 - 1 + ln(2 pi)/2 (see k2_term)
    float c;            // K2: ln N - 1
    float invN;         // 1/N
    float mi_q[2];      // MI: Nc/N and 1 - Nc/N, the expected share of each class
    float mi_scale;     // MI: 1 / (N ln 2)
    float margin;       // candidate margin, 2 x the FP32 error bound
    // Lower bounds of the badness (bound_exceeds): lb_max - lb_scale * S, and
    // for K2 that plus lb_lnw * sum_c log2 max(n_c, 1) + lb_off, with S the
    // table's chi-square sum and pi = Nc/N (lb_lnw = 0: no second bound);
    // lb_top is the largest value either can take; lb_slack is what a bound must
    // clear above the best score before a set's FP32 score may be skipped
    // (2 margins plus the bound's own FP32 error)
    float lb_pi, lb_max, lb_scale, lb_slack, lb_lnw, lb_off, lb_top;
};

// Single-SNP counts of both classes: {cases c0, cases c1, controls c0,
// controls c1}; genotype 2 is N - c0 - c1 (helper.cu:38-49).
typedef uint4 Singles;

// Load through a pointer to global memory.  Pointers read from argument
// blocks in device memory are generic to the compiler, and a generic (flat)
// load also counts against the LDS wait counter, so an epilogue that mixes
// them with LDS reads waits on both.
__device__ __forceinline__ uint4 ld_global(const uint4 *p, size_t i)
{
    typedef unsigned u32x4_t __attribute__((ext_vector_type(4)));
    const u32x4_t v = ((const __attribute__((address_space(1))) u32x4_t *)p)[i];
    return make_uint4(v.x, v.y, v.z, v.w);
}

// 3x3 table of the pair (a, b) from its four AND-counts n_ij = |a_i & b_j|,
// i, j in {0,1}, and the single counts: the five derived cells of
// helper.cu:644-662.  t[i*3+j].
__device__ __forceinline__ void pair9(uint32_t n00, uint32_t n01, uint32_t n10, uint32_t n11, uint32_t a0,
                                      uint32_t a1, uint32_t b0, uint32_t b1, uint32_t N, uint32_t t[9])
{
    t[0] = n00;
    t[1] = n01;
    t[3] = n10;
    t[4] = n11;
    t[2] = a0 - (n00 + n01);
    t[5] = a1 - (n10 + n11);
    t[6] = b0 - (n00 + n10);
    t[7] = b1 - (n01 + n11);
    t[8] = (N - a0 - a1) - (t[6] + t[7]);
}

#define EPI_C3(x, y, z) ((x) * 9 + (y) * 3 + (z))

// 27-cell table from the eight AND-counts n8[i*4+j*2+k] = |X_i & Y_j & Z_k|
// and the 3x3 pair tables xy, xz, yz.  Each of the 19 other cells is a pair
// marginal minus two cells already known, in the order of helper.cu:493-548.
__device__ __forceinline__ void triplet27(const uint32_t n8[8], const uint32_t xy[9], const uint32_t xz[9],
                                          const uint32_t yz[9], uint32_t t[27])
{
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int k = 0; k < 2; k++) t[EPI_C3(i, j, k)] = n8[i * 4 + j * 2 + k];
    // genotype 2 of Z, X and Y in {0,1}
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) t[EPI_C3(i, j, 2)] = xy[i * 3 + j] - (t[EPI_C3(i, j, 0)] + t[EPI_C3(i, j, 1)]);
    // genotype 2 of Y, X in {0,1}
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) t[EPI_C3(i, 2, k)] = xz[i * 3 + k] - (t[EPI_C3(i, 0, k)] + t[EPI_C3(i, 1, k)]);
    // genotype 2 of X
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
        for (int k = 0; k < 2; k++) t[EPI_C3(2, j, k)] = yz[j * 3 + k] - (t[EPI_C3(0, j, k)] + t[EPI_C3(1, j, k)]);
#pragma unroll
    for (int j = 0; j < 2; j++) t[EPI_C3(2, j, 2)] = xy[2 * 3 + j] - (t[EPI_C3(2, j, 0)] + t[EPI_C3(2, j, 1)]);
#pragma unroll
    for (int k = 0; k < 3; k++) t[EPI_C3(2, 2, k)] = xz[2 * 3 + k] - (t[EPI_C3(2, 0, k)] + t[EPI_C3(2, 1, k)]);
}

// ---------------------------------------------------------------- FP32 --

__device__ __forceinline__ void kahan_add(float &s, float &c, float v)
{
    float y = v - c;
    float t = s + y;
    c = (t - s) - y;
    s = t;
}

constexpr float LN2_F = 0.693147180559945309f;

// ln m! for m < LNF_SMALL in FP32, set once per device by the host; each
// scoring workgroup copies it to LDS (load_lnf).
constexpr int LNF_SMALL = 1024;
__device__ float g_lnfact[LNF_SMALL];

__device__ __forceinline__ void load_lnf(float *s_lnf)
{
    for (int i = threadIdx.x; i < LNF_SMALL; i += blockDim.x) s_lnf[i] = g_lnfact[i];
}

// K2 term T'(m) = ln m! - c*m, c = ln N - 1.
//  * m < LNF_SMALL: ln m! from the LDS copy of g_lnfact, shifted with one fma.
//  * larger m: Stirling's series with the shift folded into the logarithm,
//      T'(m) = (x - 1/2) ln(x/N) + K0,   x = m + 1,  K0 = ln(N)/2 - 1 + ln(2 pi)/2,
//    truncated after the constant (the next term, 1/(12x), is below 8.2e-5
//    here); ln(x/N) comes from v_log_f32.  The shift inside the logarithm
//    keeps every intermediate near |T'| <= N/e instead of N ln N.
// No table gather from global memory, one transcendental at most.  The host
// measures the error of exactly this code against the FP64 table over every
// m a search can reach (k_term_error) and widens the candidate margin by it.
// Explicit fmaf and no contraction, so every instantiation rounds the same.
__device__ __forceinline__ float k2_term(uint32_t m, float invN, float k0, float c, const float *s_lnf)
{
#pragma clang fp contract(off)
    if (m < (uint32_t)LNF_SMALL) return fmaf(-c, (float)m, s_lnf[m]);
    const float x = (float)(m + 1u);
    const float l2 = __builtin_amdgcn_logf(x * invN);
    return fmaf(fmaf(x, LN2_F, -0.5f * LN2_F), l2, k0);
}

// K2 in FP32.  Each cell contributes f = ln ctl! + ln cas! - ln (ctl+cas+1)!
// (scoring.hpp:5-7), always <= 0, so K2 = -sum f.  With T'(m) = ln m! - c*m
// and cas + ctl - (cas+ctl+1) = -1, every cell reads f + c, and
// K2 = ncells*c - sum(f + c); the cell sum is Kahan-compensated.
__device__ __forceinline__ float k2_f32(const uint32_t *cas, const uint32_t *ctl, int ncells, float invN, float k0,
                                        float c2, float total, const float *s_lnf)
{
    float s = 0.f, c = 0.f;
    for (int i = 0; i < ncells; i++) {
        float f = (k2_term(ctl[i], invN, k0, c2, s_lnf) + k2_term(cas[i], invN, k0, c2, s_lnf)) -
                  k2_term(ctl[i] + cas[i] + 1, invN, k0, c2, s_lnf);
        kahan_add(s, c, f);
    }
    return total - s;
}

// MI in FP32 from non-negative terms.  With x a cell's count in one class
// and y = q n_c the count expected without association (q = Nc/N for the
// cases, Nt/N for the controls, n_c = cases + controls of the cell),
//   MI = H(G) + H(Y) - H(G,Y) = sum over cells and classes of D(x, y) / (N ln 2),
//   D(x, y) = x ln(x/y) - x + y >= 0
// (scoring.hpp:9-16; the -x + y parts add up to zero over a table).  The
// entropy form subtracts sums of 2-5 bits to leave an MI near 1e-4 bits on
// data without association, so its FP32 rounding is 10^4 times the score;
// every D is >= 0, so this sum has no cancellation and keeps the relative
// error of its terms.  With d = x - y (one fma, rounded once), e = d / y and
// h(r) = r ln r - r + 1, D = y h(1 + e):
//  * |e| <= 1/4: h(1 + e) = e^2 sum_{k>=2} (-e)^(k-2) / (k (k - 1)), summed
//    to k = 12 (truncation below 6e-9 relative), so D = d e s;
//  * otherwise y (r ln r + 1 - r), r = 1 + e, ln r from v_log_f32;
//  * x = 0: D = y; n_c = 0: D = 0.
// The host measures this term's relative error against FP64 over the cells
// a search can produce (k_term_error) and sets the candidate margin from it.
__device__ __forceinline__ float mi_dterm(uint32_t x, uint32_t n, float q)
{
#pragma clang fp contract(off)
    if (n == 0u) return 0.f;
    const float nf = (float)n;
    const float y = q * nf;
    if (x == 0u) return y;
    const float d = fmaf(-q, nf, (float)x);
    const float e = d * __builtin_amdgcn_rcpf(y);
    if (fabsf(e) <= 0.25f) {
        const float t = -e;
        float s = 1.f / 132.f;
        s = fmaf(s, t, 1.f / 110.f);
        s = fmaf(s, t, 1.f / 90.f);
        s = fmaf(s, t, 1.f / 72.f);
        s = fmaf(s, t, 1.f / 56.f);
        s = fmaf(s, t, 1.f / 42.f);
        s = fmaf(s, t, 1.f / 30.f);
        s = fmaf(s, t, 1.f / 20.f);
        s = fmaf(s, t, 1.f / 12.f);
        s = fmaf(s, t, 1.f / 6.f);
        s = fmaf(s, t, 0.5f);
        return (d * e) * s;
    }
    const float r = 1.f + e;
    const float lr = __builtin_amdgcn_logf(r) * LN2_F;
    return y * fmaf(r, lr, 1.f - r);
}

__device__ __forceinline__ float mi_f32(const uint32_t *cas, const uint32_t *ctl, int ncells, const float (&q)[2],
                                        float scale)
{
    float s = 0.f;
    for (int i = 0; i < ncells; i++) {
        const uint32_t n = cas[i] + ctl[i];
        s += mi_dterm(cas[i], n, q[0]) + mi_dterm(ctl[i], n, q[1]);
    }
    return s * scale;
}

// Lower bound of a table's badness (K2, or -MI), from
//   KL(p || pi) <= K (p - pi)^2 on [0, 1],  K = max(-ln(1-pi)/pi^2, -ln(pi)/(1-pi)^2)
// (the largest KL(p||pi)/(p-pi)^2, reached at p = 0 or 1) and the chi-square
// sum S = sum_c (cas_c - pi n_c)^2 / n_c, pi = Nc/N:
//  * K2: ln((n+1)!/(a! b!)) = ln(n+1) + ln C(n, a) >= n H(a/n) since
//    C(n, k) >= e^{n H(k/n)} / (n + 1) (method of types), and
//    n H(a/n) = n H(pi) + (a - pi n) H'(pi) - n KL(a/n || pi); summed over the
//    cells the linear terms cancel:  K2 >= N H(pi) - K S  (nats).
//  * MI: I(G;Y) = sum_c (n_c/N) KL(p_c || pi) <= K S / N nats, so
//    -MI >= -K S / (N ln 2) in bits.
// Both are lb_max - lb_scale * S.  Cells with n = 0 add nothing.
//  * K2 also carries each cell's Stirling term: for n >= 1,
//      ln((n+1)!/(a! b!)) >= n H(a/n) + ln(n)/2 - c0,  c0 = ln(pi/2)/2,
//    (with Robbins' ln m! = m ln m - m + ln(2 pi m)/2 + r_m, 1/(12m+1) < r_m
//    < 1/(12m), the left side minus n H - ln(n)/2 is, for a, b >= 1,
//    ln(n+1) - ln(2 pi a b)/2 + r_n - r_a - r_b; it exceeds -c0 because
//    ln((n+1)/(2 sqrt(ab))) beats r_a + r_b < 1/(6m), m = min(a, b): it is
//    at least 1/(n+1) (4ab <= n^2) when 6m >= n+1, else above ln(1.5)/2
//    (4ab <= 4mn).  a = 0 or b = 0 gives ln(n+1) >= ln(n)/2.)  The infimum is
//    approached at a = b as n grows.  Summed over the cells it adds
//    sum_c ln(max(n_c, 1))/2 - NC c0, the same for every set of a given
//    cell occupancy, so the bound can skip sets on data whose best set is
//    weak.  The logarithms are taken of products of 5 cells (each below
//    2^25, the product below 2^125; the host sets lb_lnw = 0 for larger N)
//    with v_log_f32.
// bound_exceeds tells whether the bound of a table exceeds thr.  NC is the
// cell count (27 or 9, equal to sc.ncells), fixed at compile time so the
// short loops unroll.
template <int NC>
__device__ __forceinline__ bool bound_exceeds(const ScoreCtx &sc, const uint32_t *cas, const uint32_t *ctl, float thr)
{
    float s = 0.f;
#pragma unroll
    for (int i = 0; i < NC; i++) {
        const uint32_t nn = cas[i] + ctl[i];
        const float n = (float)nn, d = fmaf(-sc.lb_pi, n, (float)cas[i]);
        s += nn ? d * d * __builtin_amdgcn_rcpf(n) : 0.f;
    }
    const float lb = fmaf(-sc.lb_scale, s, sc.lb_max);
    // Without and with the Stirling term both are lower bounds: the logs
    // are only taken for a set the first does not already rule out.
    if (lb > thr || sc.lb_lnw == 0.f) return lb > thr;
    float l = 0.f;
#pragma unroll
    for (int g = 0; g < NC; g += 5) {
        float p = 1.f;
#pragma unroll
        for (int i = g; i < (g + 5 < NC ? g + 5 : NC); i++) {
            const uint32_t nn = cas[i] + ctl[i];
            p *= (float)(nn ? nn : 1u);
        }
        l += __builtin_amdgcn_logf(p);
    }
    return fmaf(sc.lb_lnw, l, lb + sc.lb_off) > thr;
}

// "Badness": the FP32 score mapped so that lower is always better.
__device__ __forceinline__ float badness_f32(const ScoreCtx &sc, const uint32_t *cas, const uint32_t *ctl,
                                            const float *s_lnf)
{
    if (sc.score == SCORE_K2) return k2_f32(cas, ctl, sc.ncells, sc.invN, sc.k0, sc.c, sc.k2_total, s_lnf);
    return -mi_f32(cas, ctl, sc.ncells, sc.mi_q, sc.mi_scale);
}

// ---------------------------------------------------------------- FP64 --
// Fixed operation sequences, no contraction: identical to the oracle's.

// The FP64 scores index their tables (T[0..N+1], L[0..N]) by the cells.
// Cells derived from consistent count tables each lie within their class
// and sum to its size; a table that does not is scored NaN instead of being
// looked up (the host reports a NaN score as an internal error).
__device__ __forceinline__ bool cells_valid(const uint32_t *cas, const uint32_t *ctl, int ncells, int Nc, int Nt)
{
    unsigned long long sc = 0, st = 0;
    bool ok = true;
    for (int i = 0; i < ncells; i++) {
        ok = ok && cas[i] <= (uint32_t)Nc && ctl[i] <= (uint32_t)Nt;
        sc += cas[i];
        st += ctl[i];
    }
    return ok && sc == (unsigned long long)Nc && st == (unsigned long long)Nt;
}

__device__ __forceinline__ double k2_f64(const uint32_t *cas, const uint32_t *ctl, int ncells, int Nc, int Nt,
                                         const double *T)
{
#pragma clang fp contract(off)
    if (!cells_valid(cas, ctl, ncells, Nc, Nt)) return __builtin_nan("");
    double s = 0.0;
    for (int i = 0; i < ncells; i++) {
        double t = T[ctl[i]] + T[cas[i]];
        t = t - T[ctl[i] + cas[i] + 1];
        s = s + t;
    }
    return fabs(s);
}

__device__ __forceinline__ double mi_f64(const uint32_t *cas, const uint32_t *ctl, int ncells, int Nc, int Nt,
                                         const double *L)
{
#pragma clang fp contract(off)
    if (!cells_valid(cas, ctl, ncells, Nc, Nt)) return __builtin_nan("");
    double invN = 1.0 / (double)(Nc + Nt);
    double hY = -((double)Nc * invN) * L[Nc];
    hY = hY - ((double)Nt * invN) * L[Nt];
    double hG = 0.0, hGY = 0.0;
    for (int i = 0; i < ncells; i++) {
        double pc = (double)cas[i] * invN;
        double pt = (double)ctl[i] * invN;
        hGY = hGY - pc * L[cas[i]];
        hGY = hGY - pt * L[ctl[i]];
        hG = hG - (pc + pt) * L[cas[i] + ctl[i]];
    }
    return (hG + hY) - hGY;
}

// ------------------------------------------------------------------ keys --

// Order-preserving maps of floats/doubles to unsigned integers.
__device__ __forceinline__ unsigned f2key(float f)
{
    unsigned u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key2f(unsigned k)
{
    unsigned u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
    return __uint_as_float(u);
}

// The badness above which a set's FP32 score may be skipped: the global
// bound published so far (single best: the best FP32 badness; top-k: an
// upper bound of the k-th best, see publish_topk; it only decreases, so a
// stale read only skips less) plus the slack, or +inf while no bound can
// exceed it (bounds never exceed lb_top).
__device__ __forceinline__ float skip_threshold(const ScoreCtx &sc, const unsigned *best_key)
{
    const float t = key2f(__atomic_load_n(best_key, __ATOMIC_RELAXED)) + sc.lb_slack;
    return t < sc.lb_top ? t : INFINITY;
}

__device__ __forceinline__ unsigned long long d2key(double d)
{
    unsigned long long u = (unsigned long long)__double_as_longlong(d);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ double key2d(unsigned long long k)
{
    unsigned long long u = (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
    return __longlong_as_double((long long)u);
}

__device__ __forceinline__ float wave_min(float v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wave_min_d(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ uint32_t wave_sum_u(uint32_t v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Colex ranks: C(Z,3) + C(Y,2) + X and C(B,2) + A.
__device__ __forceinline__ unsigned long long rank3(int x, int y, int z)
{
    unsigned long long Z = (unsigned long long)z, Y = (unsigned long long)y;
    return Z * (Z - 1) * (Z - 2) / 6 + Y * (Y - 1) / 2 + (unsigned long long)x;
}
__device__ __forceinline__ unsigned long long rank2(int a, int b)
{
    unsigned long long B = (unsigned long long)b;
    return B * (B - 1) / 2 + (unsigned long long)a;
}

__device__ __forceinline__ unsigned long long mix64(unsigned long long z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Hash of one set's table, combined by summation into an order-independent
// checksum: h = (rank+1)*golden; h = mix64(h ^ cell) over cases then controls.
__device__ __forceinline__ unsigned long long table_hash(unsigned long long rank, const uint32_t *cas,
                                                         const uint32_t *ctl, int ncells)
{
    unsigned long long h = (rank + 1) * 0x9E3779B97F4A7C15ull;
    for (int i = 0; i < ncells; i++) h = mix64(h ^ cas[i]);
    for (int i = 0; i < ncells; i++) h = mix64(h ^ ctl[i]);
    return h;
}

}  // namespace episdet
