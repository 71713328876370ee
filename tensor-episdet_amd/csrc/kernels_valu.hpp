// kernels_valu.hpp -- gfx950 kernels that count with VALU AND + v_bcnt_u32_b32
// over LDS-staged bit-plane tiles, plus the plain-path objective kernel and
// the FP64 candidate re-scoring.
//
// Device layout of a class: u32 [nrows][plane][W], W a multiple of KW words,
// bits past the last sample zero, rows >= n zero.  A 3-way work unit is one
// SNP X, a tile of TY Y-SNPs and a tile of TZ Z-SNPs; a 2-way unit is a tile
// pair.  Every unit streams its rows over the whole sample axis in KW-word
// chunks staged through LDS; the counts of a unit stay in registers.
#pragma once
#include "device_common.hpp"

namespace episdet {

constexpr int KW = 32;          // words per staged chunk (1024 samples)
constexpr int LDSW = KW + 4;    // LDS row stride in words: conflict-free b128 reads

enum { MODE_SEARCH = 0, MODE_DUMP = 1, MODE_CHECKSUM = 2, MODE_COUNTS = 3, MODE_PAIRTABLE = 4, MODE_TOPK = 5 };

// Modes that score sets: MODE_SEARCH keeps the best set's neighbourhood,
// MODE_TOPK the k best sets' (a separate instantiation, so the single-best
// kernels carry none of the top-k registers).
constexpr bool scores_sets(int mode) { return mode == MODE_SEARCH || mode == MODE_TOPK; }

// ------------------------------------------------------------ data prep --

// Copies caller planes (rows of Win words) into the device layout (rows of
// Wd words), clearing every bit at or past sample N and every SNP >= n.
// dst holds nsnp_rows SNPs x 2 planes x Wd words; src holds n x 2 x Win.
__global__ void k_repack(const uint32_t *__restrict__ src, int Win, uint32_t *__restrict__ dst, int Wd, int n,
                         int nsnp_rows, int N)
{
    const size_t total = (size_t)nsnp_rows * 2 * Wd;
    const int Wv = (N + 31) / 32;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        size_t row = i / Wd;
        int w = (int)(i % Wd);
        uint32_t v = 0;
        if ((int)(row >> 1) < n && w < Wv && w < Win) {
            v = src[row * Win + w];
            int valid = N - w * 32;
            if (valid < 32) v &= 0xFFFFFFFFu << (32 - valid);
        }
        dst[i] = v;
    }
}

// Single-SNP plane counts: one wave per SNP, both classes
// (helper.cu:13-52).  out[v] = {cases c0, cases c1, controls c0, controls c1}.
// A sample marked in both planes of a SNP encodes no genotype, and every
// derived count of that SNP would be off (genotype 2 = N - c0 - c1 can go
// negative); such SNPs lower *bad_snp (initialised to ~0) to their index.
__global__ __launch_bounds__(64) void k_singles(const uint32_t *__restrict__ p0, int W0,
                                                const uint32_t *__restrict__ p1, int W1, int n, uint4 *out,
                                                unsigned *bad_snp)
{
    const int v = blockIdx.x, lane = threadIdx.x;
    if (v >= n) return;
    uint32_t c[4] = {0, 0, 0, 0}, both = 0;
    const uint32_t *r0 = p0 + (size_t)v * 2 * W0, *r1 = p1 + (size_t)v * 2 * W1;
    for (int w = lane; w < W0; w += 64) {
        const uint32_t a = r0[w], b = r0[W0 + w];
        c[0] += __popc(a);
        c[1] += __popc(b);
        both |= a & b;
    }
    for (int w = lane; w < W1; w += 64) {
        const uint32_t a = r1[w], b = r1[W1 + w];
        c[2] += __popc(a);
        c[3] += __popc(b);
        both |= a & b;
    }
#pragma unroll
    for (int i = 0; i < 4; i++) c[i] = wave_sum_u(c[i]);
    const uint32_t nboth = wave_sum_u(both ? 1u : 0u);
    if (lane == 0) {
        out[v] = make_uint4(c[0], c[1], c[2], c[3]);
        if (nboth) atomicMin(bad_snp, (unsigned)v);
    }
}

// -------------------------------------------------------- shared output --

struct CandOut {
    unsigned *best_key;          // order-preserving key of the best FP32 badness
    unsigned long long *cand;    // packed indices of candidates
    float *cand_s32;             // their FP32 score
    unsigned *cand_count;        // [0]: number appended (may exceed cand_cap: overflow);
                                 // [1]: sticky overflow flag, set by every append that
                                 // finds no slot, since [0] can wrap past 2^32
    unsigned cand_cap;
    unsigned topk;               // > 1: keep every set that can be among the topk best
};

// Top-k publication.  The k-th smallest FP32 badness among a block's sets
// of one work unit is the badness of k distinct sets, so it bounds the
// global k-th best from above.  It joins a global bound (atomicMin on
// *best_key), and every finite set within `margin` of the bound is appended.
// The bound only decreases, so each of the k best sets (by exact score) is
// appended when it is evaluated.  The k-th smallest key is found by an
// MSB-first radix select, one block-wide count per bit; it runs only when at
// least k of the block's values lie below the current bound.  Requires
// k <= M * blockDim.x (every kernel here has M * blockDim.x >= 1024).
template <int M, int NW>
__device__ void publish_topk(const CandOut &co, float margin, const float (&bad)[M],
                             const unsigned long long (&packed)[M])
{
    __shared__ unsigned s_cnt[2][NW];
    __shared__ unsigned s_g;
    unsigned key[M];
#pragma unroll
    for (int i = 0; i < M; i++) key[i] = f2key(bad[i]);     // +inf keys equal the initial bound: never counted
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    __syncthreads();                          // s_g / s_cnt of the previous call are consumed
    if (tid == 0) s_g = __hip_atomic_load(co.best_key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    unsigned g = s_g;
    unsigned mine = 0;
#pragma unroll
    for (int i = 0; i < M; i++) mine += key[i] < g;
    mine = wave_sum_u(mine);
    if (lane == 0) s_cnt[0][w] = mine;
    __syncthreads();
    unsigned c = 0;
#pragma unroll
    for (int i = 0; i < NW; i++) c += s_cnt[0][i];
    if (c >= co.topk) {
        unsigned prefix = 0, kk = co.topk;
        for (int b = 31; b >= 0; b--) {
            const unsigned hi = b == 31 ? 0u : (~0u << (b + 1));
            unsigned z = 0;
#pragma unroll
            for (int i = 0; i < M; i++) z += ((key[i] ^ prefix) & hi) == 0 && ((key[i] >> b) & 1u) == 0;
            z = wave_sum_u(z);
            const int buf = b & 1;            // bit 31 uses buffer 1 (buffer 0 still holds c), then alternate
            if (lane == 0) s_cnt[buf][w] = z;
            __syncthreads();
            unsigned cnt = 0;
#pragma unroll
            for (int i = 0; i < NW; i++) cnt += s_cnt[buf][i];
            if (kk > cnt) {
                kk -= cnt;
                prefix |= 1u << b;
            }
        }
        if (tid == 0) {
            const unsigned old = atomicMin(co.best_key, prefix);
            s_g = old < prefix ? old : prefix;
        }
        __syncthreads();
        g = s_g;
    }
    const float thr = g >= 0xFF800000u ? INFINITY : key2f(g) + margin;
#pragma unroll
    for (int i = 0; i < M; i++) {
        if (bad[i] <= thr && bad[i] < INFINITY) {
            unsigned slot = atomicAdd(co.cand_count, 1u);
            if (slot < co.cand_cap) {
                co.cand[slot] = packed[i];
                co.cand_s32[slot] = bad[i];
            } else {
                co.cand_count[1] = 1u;
            }
        }
    }
}

// Block-wide candidate publication for a work unit.  bad[] holds the
// thread's FP32 badness values (+inf for invalid sets).  The block minimum
// joins the global minimum; every set within `margin` of the smaller of the
// two is appended.  Since the global minimum only decreases, the final
// argmin of the exact scores is always among the appended sets when margin
// is twice the FP32 error bound.
// NW = waves per workgroup; s_red holds at least NW + 1 floats.
template <int M, int NW = 4>
__device__ __forceinline__ void publish_candidates(const CandOut &co, float margin, const float (&bad)[M],
                                                   const unsigned long long (&packed)[M], float *s_red)
{
    float tmin = INFINITY;
#pragma unroll
    for (int i = 0; i < M; i++) tmin = fminf(tmin, bad[i]);
    float wmin = wave_min(tmin);
    const int tid = threadIdx.x;
    if ((tid & 63) == 0) s_red[tid >> 6] = wmin;
    __syncthreads();
    if (tid == 0) {
        float b = s_red[0];
#pragma unroll
        for (int w = 1; w < NW; w++) b = fminf(b, s_red[w]);
        float thr = -INFINITY;
        if (b < INFINITY) {
            unsigned old = atomicMin(co.best_key, f2key(b));
            thr = fminf(key2f(old), b) + margin;
        }
        s_red[NW] = thr;
    }
    __syncthreads();
    const float thr = s_red[NW];
#pragma unroll
    for (int i = 0; i < M; i++) {
        if (bad[i] <= thr) {
            unsigned slot = atomicAdd(co.cand_count, 1u);
            if (slot < co.cand_cap) {
                co.cand[slot] = packed[i];
                co.cand_s32[slot] = bad[i];
            } else {
                co.cand_count[1] = 1u;
            }
        }
    }
}

// ============================================================== 2-way ====

struct PairArgs {
    const uint32_t *p0, *p1;     // cases / controls planes
    unsigned bytes0, bytes1;     // size of each class's planes (below 4 GiB on MFMA paths)
    int W0, W1, n, N0, N1;
    const uint4 *singles;
    ScoreCtx sc;
    int ntiles;                  // tiles along the SNP axis
    unsigned long long nunits;   // ntiles*(ntiles+1)/2, or blocked_units(ntiles)
    int blocked;                 // units in XCD blocks (blocked_unrank), else triangular
    unsigned long long *queue;
    CandOut co;
    uint4 *pairs;                // MODE_PAIRTABLE: [n][n][2] uint4 (a < b only)
    uint32_t *dump;              // MODE_DUMP: [rank][2][9]
    unsigned long long *checksum;
};

// Accumulates the four AND-counts of R x R pairs over one class: rows
// a = a0 + ta + 16r, b = b0 + tb + 16s.
template <int R>
__device__ __forceinline__ void pair_count_class(const uint32_t *__restrict__ base, int W, int a0, int b0,
                                                 uint32_t (*sA)[16 * R][LDSW], uint32_t (*sB)[16 * R][LDSW],
                                                 uint32_t (&acc)[R][R][4])
{
    constexpr int T = 16 * R;
    const int tid = threadIdx.x, ta = tid >> 4, tb = tid & 15;
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
        for (int s = 0; s < R; s++)
#pragma unroll
            for (int c = 0; c < 4; c++) acc[r][s][c] = 0;
    for (int w0 = 0; w0 < W; w0 += KW) {
        __syncthreads();
        // 2*T rows for each side, KW/4 uint4 per row
        for (int e = tid; e < 2 * T * (KW / 4); e += 256) {
            int row = e / (KW / 4), q = e % (KW / 4);
            int v = row >> 1, k = row & 1;
            uint4 va = *(const uint4 *)(base + ((size_t)(a0 + v) * 2 + k) * W + w0 + q * 4);
            uint4 vb = *(const uint4 *)(base + ((size_t)(b0 + v) * 2 + k) * W + w0 + q * 4);
            *(uint4 *)&sA[k][v][q * 4] = va;
            *(uint4 *)&sB[k][v][q * 4] = vb;
        }
        __syncthreads();
#pragma unroll 2
        for (int w = 0; w < KW; w += 4) {
            uint4 av[R][2], bv[R][2];
#pragma unroll
            for (int r = 0; r < R; r++)
#pragma unroll
                for (int k = 0; k < 2; k++) av[r][k] = *(const uint4 *)&sA[k][ta + 16 * r][w];
#pragma unroll
            for (int s = 0; s < R; s++)
#pragma unroll
                for (int k = 0; k < 2; k++) bv[s][k] = *(const uint4 *)&sB[k][tb + 16 * s][w];
#pragma unroll
            for (int r = 0; r < R; r++)
#pragma unroll
                for (int s = 0; s < R; s++)
#pragma unroll
                    for (int i = 0; i < 2; i++)
#pragma unroll
                        for (int j = 0; j < 2; j++) {
                            uint32_t c = acc[r][s][i * 2 + j];
                            c += __popc(av[r][i].x & bv[s][j].x);
                            c += __popc(av[r][i].y & bv[s][j].y);
                            c += __popc(av[r][i].z & bv[s][j].z);
                            c += __popc(av[r][i].w & bv[s][j].w);
                            acc[r][s][i * 2 + j] = c;
                        }
        }
    }
}

__host__ __device__ inline void unrank_tri(unsigned long long u, int &ta, int &tb)
{
    // u = tb*(tb+1)/2 + ta, 0 <= ta <= tb
    double d = sqrt(8.0 * (double)u + 1.0);
    long long b = (long long)((d - 1.0) * 0.5);
    while ((unsigned long long)(b + 1) * (unsigned long long)(b + 2) / 2 <= u) b++;
    while ((unsigned long long)b * (unsigned long long)(b + 1) / 2 > u) b--;
    tb = (int)b;
    ta = (int)(u - (unsigned long long)b * (unsigned long long)(b + 1) / 2);
}

// Units of a triangle of T tiles in XCD blocks.  The triangle is cut into
// 8 x 8 blocks of units (S = ceil(T / 8) blocks a side, NB = S (S + 1) / 2 of
// them, triangular order); unit u = 8 v + x belongs to XCD x, and XCD x
// walks blocks x, x + 8, x + 16, ... 64 units at a time.  The 64 units one
// XCD runs at once (one per resident workgroup) then share 8 A tiles and 8
// B tiles, which its L2 holds.  Units outside the triangle, or in a block
// past NB, are empty.
__host__ __device__ inline unsigned long long blocked_units(int T)
{
    const unsigned long long S = (unsigned long long)((T + 7) / 8), NB = S * (S + 1) / 2;
    return 512ull * ((NB + 7) / 8);
}

__host__ __device__ inline bool blocked_unrank(unsigned long long u, int T, int &ta, int &tb)
{
    const unsigned long long v = u >> 3, B = 8 * (v >> 6) + (u & 7);
    const unsigned long long S = (unsigned long long)((T + 7) / 8);
    if (B >= S * (S + 1) / 2) return false;
    int sa, sb;
    unrank_tri(B, sa, sb);
    const int k = (int)(v & 63);
    ta = 8 * sa + (k >> 3);
    tb = 8 * sb + (k & 7);
    return ta <= tb && tb < T;
}

// XCD (accelerator complex die) of the calling wave: bits [3:0] of
// HW_REG_XCC_ID (hardware register 20).
__device__ __forceinline__ unsigned xcc_id()
{
    return (unsigned)__builtin_amdgcn_s_getreg(20 | (0 << 6) | ((4 - 1) << 11)) & 7u;
}

// The next unit from per-XCD queues: q[1 + x] counts the units u = 8 v + x of
// XCD x.  A workgroup takes its own XCD's next unit, and once those are gone
// the other XCDs', so every queue drains; total or more means none is left.
__device__ __forceinline__ unsigned long long fetch_xcd(unsigned long long *q, unsigned long long total)
{
    const unsigned x = xcc_id();
    for (unsigned k = 0; k < 8; k++) {
        const unsigned y = (x + k) & 7u;
        const unsigned long long u = 8ull * atomicAdd(q + 1 + y, 1ull) + y;
        if (u < total) return u;
    }
    return total;
}

// A batch of independent 2-way searches sharing one work queue (see
// BatchTri); a single search or pair-table build is a batch of one.
struct BatchPair {
    const PairArgs *jobs;
    const unsigned long long *job_start;
    int njobs;
    unsigned long long total;
    unsigned long long *queue;
    // MODE_COUNTS (k_pairs_mfma, the plain path on the matrix cores):
    // `rounds` rounds (A block ab, B blocks bb0 .. bb0 + rounds - 1 of 256)
    // per launch, 8 workgroups each; round r writes its AND-counts to
    // counts + r * slot as [class][256 A][256 B][4] (k_pairs_counts' layout)
    int ab = 0, bb0 = 0, rounds = 0;
    uint32_t *counts = nullptr;
    size_t slot = 0;
    // 1: units come from the per-XCD queues queue[1..8] (fetch_xcd), which
    // hand them out of order
    int xcd = 0;
};

template <int R, int MODE>
__global__ __launch_bounds__(256) void k_pairs_valu(BatchPair bp)
{
    constexpr int T = 16 * R;
    __shared__ uint32_t sA[2][T][LDSW];
    __shared__ uint32_t sB[2][T][LDSW];
    __shared__ unsigned long long s_u;
    __shared__ float s_red[8];
    __shared__ float s_lnf[scores_sets(MODE) ? LNF_SMALL : 1];
    if constexpr (scores_sets(MODE)) load_lnf(s_lnf);
    const int tid = threadIdx.x, ta = tid >> 4, tb = tid & 15;
    int j = 0;
    for (;;) {
        if (tid == 0) s_u = atomicAdd(bp.queue, 1ull);
        __syncthreads();
        const unsigned long long uu = s_u;
        if (uu >= bp.total) break;
        while (j + 1 < bp.njobs && bp.job_start[j + 1] <= uu) j++;
        const PairArgs &p = bp.jobs[j];
        const unsigned long long u = uu - bp.job_start[j];
        int tA, tB;
        unrank_tri(u, tA, tB);
        const int a0 = tA * T, b0 = tB * T;
        uint32_t acc0[R][R][4], acc1[R][R][4];
        pair_count_class<R>(p.p0, p.W0, a0, b0, sA, sB, acc0);
        pair_count_class<R>(p.p1, p.W1, a0, b0, sA, sB, acc1);

        float bad[R * R];
        unsigned long long packed[R * R];
        unsigned long long hsum = 0;
        float skip_above = INFINITY;     // see skip_threshold
        if constexpr (scores_sets(MODE)) skip_above = skip_threshold(p.sc, p.co.best_key);
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int A = a0 + ta + 16 * r;
#pragma unroll
            for (int s = 0; s < R; s++) {
                const int B = b0 + tb + 16 * s;
                const int m = r * R + s;
                bad[m] = INFINITY;
                packed[m] = (unsigned long long)A | ((unsigned long long)B << 32);
                if (!(A < B && B < p.n)) continue;
                if constexpr (MODE == MODE_PAIRTABLE) {
                    uint4 *o = p.pairs + ((size_t)A * p.n + B) * 2;
                    o[0] = make_uint4(acc0[r][s][0], acc0[r][s][1], acc0[r][s][2], acc0[r][s][3]);
                    o[1] = make_uint4(acc1[r][s][0], acc1[r][s][1], acc1[r][s][2], acc1[r][s][3]);
                } else {
                    const Singles sa = ld_global(p.singles, A), sb = ld_global(p.singles, B);
                    uint32_t cas[9], ctl[9];
                    pair9(acc0[r][s][0], acc0[r][s][1], acc0[r][s][2], acc0[r][s][3], sa.x, sa.y, sb.x, sb.y,
                          (uint32_t)p.N0, cas);
                    pair9(acc1[r][s][0], acc1[r][s][1], acc1[r][s][2], acc1[r][s][3], sa.z, sa.w, sb.z, sb.w,
                          (uint32_t)p.N1, ctl);
                    if constexpr (scores_sets(MODE)) {
                        if (skip_above < INFINITY && bound_exceeds<9>(p.sc, cas, ctl, skip_above))
                            continue;                           // bad[m] stays INFINITY
                        bad[m] = badness_f32(p.sc, cas, ctl, s_lnf);
                    } else if constexpr (MODE == MODE_DUMP) {
                        uint32_t *o = p.dump + rank2(A, B) * 18;
                        for (int c = 0; c < 9; c++) {
                            o[c] = cas[c];
                            o[9 + c] = ctl[c];
                        }
                    } else if constexpr (MODE == MODE_CHECKSUM) {
                        hsum += table_hash(rank2(A, B), cas, ctl, 9);
                    }
                }
            }
        }
        if constexpr (MODE == MODE_SEARCH) publish_candidates<R * R>(p.co, p.sc.margin, bad, packed, s_red);
        if constexpr (MODE == MODE_TOPK) publish_topk<R * R, 4>(p.co, p.sc.margin, bad, packed);
        if constexpr (MODE == MODE_CHECKSUM) {
            if (hsum) atomicAdd(p.checksum, hsum);
        }
    }
}

// ============================================================== 3-way ====

struct TriArgs {
    const uint32_t *p0, *p1;
    unsigned bytes0, bytes1;             // size of each class's planes (below 4 GiB on MFMA paths)
    int W0, W1, n, N0, N1;
    const uint4 *singles;
    const uint4 *pairs;                  // [n][n][2] uint4, a < b
    ScoreCtx sc;
    const unsigned long long *gstart;    // [ngroups+1] first unit of each group
    const int2 *groups;                  // (Y tile, Z tile)
    int ngroups;
    unsigned long long nunits;
    unsigned long long *queue;
    CandOut co;
    uint32_t *dump;                      // MODE_DUMP: [rank][2][27]
    unsigned long long *checksum;
    // MODE_COUNTS (plain path): one round = X, Y block yb, Z block zb of 256
    int X, yb, zb;
    uint32_t *counts;                    // [2][256][256][8]
};

// Group of unit u when the caller's previous unit was in group g0 <= answer
// (a workgroup's units only increase): a short forward scan, then a binary
// search over what is left if the unit jumped far ahead.
__device__ __forceinline__ int find_group(const unsigned long long *gs, int ng, unsigned long long u);
__device__ __forceinline__ int next_group(const unsigned long long *gs, int ng, unsigned long long u, int g0)
{
    int g = g0;
#pragma unroll 1
    for (int step = 0; step < 4; step++) {
        if (g + 1 < ng && gs[g + 1] <= u) g++;
        else return g;
    }
    if (g + 1 < ng && gs[g + 1] <= u) {
        int lo = g, hi = ng;
        while (hi - lo > 1) {
            int mid = (lo + hi) >> 1;
            if (gs[mid] <= u) lo = mid;
            else hi = mid;
        }
        g = lo;
    }
    return g;
}

__device__ __forceinline__ int find_group(const unsigned long long *gs, int ng, unsigned long long u)
{
    int lo = 0, hi = ng;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (gs[mid] <= u) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Eight AND-counts of A x B (Y, Z) combinations with a fixed X over one
// class.  P_ij = X_i & Y_j is formed while staging, so the inner loop is one
// v_and_b32 + one v_bcnt_u32_b32 per cell per word.
template <int A, int B>
__device__ __forceinline__ void tri_count_class(const uint32_t *__restrict__ base, int W, int X, int y0, int z0,
                                                uint32_t (*sP)[16 * A][LDSW], uint32_t (*sZ)[16 * B][LDSW],
                                                uint32_t (&acc)[A][B][8])
{
    constexpr int TY = 16 * A, TZ = 16 * B;
    const int tid = threadIdx.x, ty = tid >> 4, tz = tid & 15;
#pragma unroll
    for (int a = 0; a < A; a++)
#pragma unroll
        for (int b = 0; b < B; b++)
#pragma unroll
            for (int c = 0; c < 8; c++) acc[a][b][c] = 0;
    const uint32_t *xr = base + (size_t)X * 2 * W;
    for (int w0 = 0; w0 < W; w0 += KW) {
        __syncthreads();
        for (int e = tid; e < TY * KW; e += 256) {
            const int y = e / KW, w = e % KW;
            const uint32_t *yr = base + (size_t)(y0 + y) * 2 * W + w0 + w;
            const uint32_t x0 = xr[w0 + w], x1 = xr[W + w0 + w];
            const uint32_t v0 = yr[0], v1 = yr[W];
            sP[0][y][w] = x0 & v0;
            sP[1][y][w] = x0 & v1;
            sP[2][y][w] = x1 & v0;
            sP[3][y][w] = x1 & v1;
        }
        for (int e = tid; e < 2 * TZ * (KW / 4); e += 256) {
            const int row = e / (KW / 4), q = e % (KW / 4);
            const int z = row >> 1, k = row & 1;
            *(uint4 *)&sZ[k][z][q * 4] = *(const uint4 *)(base + ((size_t)(z0 + z) * 2 + k) * W + w0 + q * 4);
        }
        __syncthreads();
#pragma unroll 2
        for (int w = 0; w < KW; w += 4) {
            uint4 pv[A][4], zv[B][2];
#pragma unroll
            for (int a = 0; a < A; a++)
#pragma unroll
                for (int ij = 0; ij < 4; ij++) pv[a][ij] = *(const uint4 *)&sP[ij][ty + 16 * a][w];
#pragma unroll
            for (int b = 0; b < B; b++)
#pragma unroll
                for (int k = 0; k < 2; k++) zv[b][k] = *(const uint4 *)&sZ[k][tz + 16 * b][w];
#pragma unroll
            for (int a = 0; a < A; a++)
#pragma unroll
                for (int b = 0; b < B; b++)
#pragma unroll
                    for (int ij = 0; ij < 4; ij++)
#pragma unroll
                        for (int k = 0; k < 2; k++) {
                            uint32_t c = acc[a][b][ij * 2 + k];
                            c += __popc(pv[a][ij].x & zv[b][k].x);
                            c += __popc(pv[a][ij].y & zv[b][k].y);
                            c += __popc(pv[a][ij].z & zv[b][k].z);
                            c += __popc(pv[a][ij].w & zv[b][k].w);
                            acc[a][b][ij * 2 + k] = c;
                        }
        }
    }
}

// 3x3 pair table of (a, b), a < b, class cls, from the pair AND-count table.
__device__ __forceinline__ void pair9_lookup(const uint4 *pairs, int n, const Singles &sa, const Singles &sb, int a,
                                             int b, int cls, uint32_t N, uint32_t t[9])
{
    const uint4 q = ld_global(pairs, ((size_t)a * n + b) * 2 + cls);
    if (cls == 0) pair9(q.x, q.y, q.z, q.w, sa.x, sa.y, sb.x, sb.y, N, t);
    else pair9(q.x, q.y, q.z, q.w, sa.z, sa.w, sb.z, sb.w, N, t);
}

// Both classes' 27-cell tables of (X, Y, Z) from its counts.
__device__ __forceinline__ void triplet_tables(const TriArgs &p, int X, int Y, int Z, const Singles &sx,
                                               const Singles &sy, const uint32_t n8c[8], const uint32_t n8t[8],
                                               uint32_t cas[27], uint32_t ctl[27])
{
    const Singles sz = ld_global(p.singles, Z);
    uint32_t xy[9], xz[9], yz[9];
    pair9_lookup(p.pairs, p.n, sx, sy, X, Y, 0, (uint32_t)p.N0, xy);
    pair9_lookup(p.pairs, p.n, sx, sz, X, Z, 0, (uint32_t)p.N0, xz);
    pair9_lookup(p.pairs, p.n, sy, sz, Y, Z, 0, (uint32_t)p.N0, yz);
    triplet27(n8c, xy, xz, yz, cas);
    pair9_lookup(p.pairs, p.n, sx, sy, X, Y, 1, (uint32_t)p.N1, xy);
    pair9_lookup(p.pairs, p.n, sx, sz, X, Z, 1, (uint32_t)p.N1, xz);
    pair9_lookup(p.pairs, p.n, sy, sz, Y, Z, 1, (uint32_t)p.N1, yz);
    triplet27(n8t, xy, xz, yz, ctl);
}

template <int A, int B, int MODE>
__global__ __launch_bounds__(256) void k_triplets_valu(TriArgs p)
{
    constexpr int TY = 16 * A, TZ = 16 * B;
    __shared__ uint32_t sP[4][TY][LDSW];
    __shared__ uint32_t sZ[2][TZ][LDSW];
    __shared__ unsigned long long s_u;
    __shared__ float s_red[8];
    __shared__ float s_lnf[scores_sets(MODE) ? LNF_SMALL : 1];
    if constexpr (scores_sets(MODE)) load_lnf(s_lnf);
    const int tid = threadIdx.x, ty = tid >> 4, tz = tid & 15;
    int g = 0;
    for (int iter = 0;; iter++) {
        int X, y0, z0;
        if constexpr (MODE == MODE_COUNTS) {
            if (iter > 0) break;
            const int tiles_z = 256 / TZ;
            X = p.X;
            y0 = p.yb * 256 + (int)(blockIdx.x / tiles_z) * TY;
            z0 = p.zb * 256 + (int)(blockIdx.x % tiles_z) * TZ;
        } else {
            if (tid == 0) s_u = atomicAdd(p.queue, 1ull);
            __syncthreads();
            const unsigned long long u = s_u;
            if (u >= p.nunits) break;
            g = next_group(p.gstart, p.ngroups, u, g);
            X = (int)(u - p.gstart[g]);
            const int2 gt = p.groups[g];
            y0 = gt.x * TY;
            z0 = gt.y * TZ;
        }
        uint32_t acc0[A][B][8], acc1[A][B][8];
        tri_count_class<A, B>(p.p0, p.W0, X, y0, z0, sP, sZ, acc0);
        tri_count_class<A, B>(p.p1, p.W1, X, y0, z0, sP, sZ, acc1);

        if constexpr (MODE == MODE_COUNTS) {
#pragma unroll
            for (int a = 0; a < A; a++)
#pragma unroll
                for (int b = 0; b < B; b++) {
                    const int yl = y0 - p.yb * 256 + ty + 16 * a, zl = z0 - p.zb * 256 + tz + 16 * b;
                    uint4 *o0 = (uint4 *)(p.counts + ((size_t)(0 * 256 + yl) * 256 + zl) * 8);
                    uint4 *o1 = (uint4 *)(p.counts + ((size_t)(1 * 256 + yl) * 256 + zl) * 8);
                    o0[0] = make_uint4(acc0[a][b][0], acc0[a][b][1], acc0[a][b][2], acc0[a][b][3]);
                    o0[1] = make_uint4(acc0[a][b][4], acc0[a][b][5], acc0[a][b][6], acc0[a][b][7]);
                    o1[0] = make_uint4(acc1[a][b][0], acc1[a][b][1], acc1[a][b][2], acc1[a][b][3]);
                    o1[1] = make_uint4(acc1[a][b][4], acc1[a][b][5], acc1[a][b][6], acc1[a][b][7]);
                }
        } else {
            const Singles sx = ld_global(p.singles, X);
            float bad[A * B];
            unsigned long long packed[A * B];
            unsigned long long hsum = 0;
            float skip_above = INFINITY;     // see skip_threshold
            if constexpr (scores_sets(MODE)) skip_above = skip_threshold(p.sc, p.co.best_key);
#pragma unroll
            for (int a = 0; a < A; a++) {
                const int Y = y0 + ty + 16 * a;
                const bool vy = X < Y && Y < p.n;
                const Singles sy = vy ? ld_global(p.singles, Y) : Singles{0, 0, 0, 0};
#pragma unroll
                for (int b = 0; b < B; b++) {
                    const int Z = z0 + tz + 16 * b;
                    const int m = a * B + b;
                    bad[m] = INFINITY;
                    packed[m] = (unsigned long long)X | ((unsigned long long)Y << 21) | ((unsigned long long)Z << 42);
                    if (!(vy && Y < Z && Z < p.n)) continue;
                    uint32_t cas[27], ctl[27];
                    triplet_tables(p, X, Y, Z, sx, sy, acc0[a][b], acc1[a][b], cas, ctl);
                    if constexpr (scores_sets(MODE)) {
                        if (skip_above < INFINITY && bound_exceeds<27>(p.sc, cas, ctl, skip_above))
                            continue;                           // bad[m] stays INFINITY
                        bad[m] = badness_f32(p.sc, cas, ctl, s_lnf);
                    } else if constexpr (MODE == MODE_DUMP) {
                        uint32_t *o = p.dump + rank3(X, Y, Z) * 54;
                        for (int c = 0; c < 27; c++) {
                            o[c] = cas[c];
                            o[27 + c] = ctl[c];
                        }
                    } else if constexpr (MODE == MODE_CHECKSUM) {
                        hsum += table_hash(rank3(X, Y, Z), cas, ctl, 27);
                    }
                }
            }
            if constexpr (MODE == MODE_SEARCH) publish_candidates<A * B>(p.co, p.sc.margin, bad, packed, s_red);
            if constexpr (MODE == MODE_TOPK) publish_topk<A * B, 4>(p.co, p.sc.margin, bad, packed);
            if constexpr (MODE == MODE_CHECKSUM) {
                if (hsum) atomicAdd(p.checksum, hsum);
            }
        }
    }
}

// ==================================================== plain objective ====

// 2-way plain round (search-pairs.cu:137-211): the four AND-counts of every
// (A, B) of A block ab x B block bb (256 x 256 each), both classes, into
// counts[2][256][256][4].  16 workgroups, one per 64 x 64 tile.  Rows up to
// the next multiple of 256 exist in the device layout (zero past n).
__global__ __launch_bounds__(256) void k_pairs_counts(const uint32_t *__restrict__ p0, int W0,
                                                      const uint32_t *__restrict__ p1, int W1, int ab, int bb,
                                                      uint32_t *__restrict__ counts)
{
    constexpr int R = 4, T = 16 * R;
    __shared__ uint32_t sA[2][T][LDSW];
    __shared__ uint32_t sB[2][T][LDSW];
    const int tid = threadIdx.x, ta = tid >> 4, tb = tid & 15;
    const int a0 = ab * 256 + (int)(blockIdx.x >> 2) * T, b0 = bb * 256 + (int)(blockIdx.x & 3) * T;
    uint32_t acc0[R][R][4], acc1[R][R][4];
    pair_count_class<R>(p0, W0, a0, b0, sA, sB, acc0);
    pair_count_class<R>(p1, W1, a0, b0, sA, sB, acc1);
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
        for (int q = 0; q < R; q++) {
            const int al = a0 - ab * 256 + ta + 16 * r, bl = b0 - bb * 256 + tb + 16 * q;
            *(uint4 *)(counts + ((size_t)(0 * 256 + al) * 256 + bl) * 4) =
                make_uint4(acc0[r][q][0], acc0[r][q][1], acc0[r][q][2], acc0[r][q][3]);
            *(uint4 *)(counts + ((size_t)(1 * 256 + al) * 256 + bl) * 4) =
                make_uint4(acc1[r][q][0], acc1[r][q][1], acc1[r][q][2], acc1[r][q][3]);
        }
}

struct PlainObjArgs {
    const uint32_t *counts;              // 3-way: [2][256][256][8] (MODE_COUNTS);
                                         // 2-way: [2][256][256][4] (k_pairs_counts)
    const uint4 *singles;
    const uint4 *pairs;                  // 3-way only
    int order, n, N0, N1, score;
    const double *t64;                   // FP64 table (ln m! or log2(m/N))
    int X, yb, zb;                       // 3-way round: X, Y block, Z block; 2-way: A block yb, B block zb
    unsigned long long *blk_out;         // [256][2] this round's block bests {key, packed}
    const unsigned long long *blk_in;    // the previous round's on this stream, or null
    unsigned long long *best;            // this stream's running best {key, packed}
    unsigned *nan_flag;                  // set when a table is inconsistent (cells_valid)
};

// (key, packed) pairs ordered by the FP64 badness key, then the packed index:
// the tie rule of the exact search, lowest packed index among equal scores.
__device__ __forceinline__ bool lex_less(unsigned long long k1, unsigned long long p1, unsigned long long k2,
                                         unsigned long long p2)
{
    return k1 < k2 || (k1 == k2 && p1 < p2);
}
__device__ __forceinline__ void wave_lexmin(unsigned long long &k, unsigned long long &pk)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const unsigned long long ok = __shfl_xor(k, o, 64), op = __shfl_xor(pk, o, 64);
        if (lex_less(ok, op, k, pk)) {
            k = ok;
            pk = op;
        }
    }
}

// Folds a round's 256 block bests into a stream's running best (one wave).
__device__ __forceinline__ void plain_fold(const unsigned long long *blk, unsigned long long *best)
{
    const int t = threadIdx.x;
    unsigned long long k = ~0ull, pk = ~0ull;
    for (int i = t; i < 256; i += 64) {
        const unsigned long long kk = blk[2 * i], pp = blk[2 * i + 1];
        if (lex_less(kk, pp, k, pk)) {
            k = kk;
            pk = pp;
        }
    }
    wave_lexmin(k, pk);
    if (t == 0 && lex_less(k, pk, best[0], best[1])) {
        best[0] = k;
        best[1] = pk;
    }
}

// One round of the plain path: derives every cell of the round's sets from
// its count tables, scores them in FP64 and keeps the round's best set by
// (exact score, lowest packed index) -- the reference's objective kernel and
// best-set reduction (helper.cu:438-584, 594-698, reduction.hpp:64-122)
// with its nondeterministic index race replaced by a total order.  Workgroup
// b holds Y (A) = yb*256 + b, thread t Z (B) = zb*256 + t.  Each workgroup
// writes its best to blk_out; workgroup 0 first folds the previous round's
// blk_in (that kernel ran before this one on the same stream) into the
// stream's best, so rounds need no cross-workgroup synchronisation and no
// candidate list that ties could overflow.
__global__ __launch_bounds__(256) void k_objective_plain(PlainObjArgs p)
{
    __shared__ unsigned long long s_k[4], s_p[4];
    const int yl = blockIdx.x, zl = threadIdx.x;
    if (yl == 0 && p.blk_in && zl < 64) plain_fold(p.blk_in, p.best);
    double bad = INFINITY;
    bool valid;
    unsigned long long packed;
    if (p.order == 3) {
        const int X = p.X, Y = p.yb * 256 + yl, Z = p.zb * 256 + zl;
        valid = X < Y && Y < Z && Z < p.n;
        packed = (unsigned long long)X | ((unsigned long long)Y << 21) | ((unsigned long long)Z << 42);
        if (valid) {
            const uint32_t *c0 = p.counts + ((size_t)(0 * 256 + yl) * 256 + zl) * 8;
            const uint32_t *c1 = p.counts + ((size_t)(1 * 256 + yl) * 256 + zl) * 8;
            uint32_t n8c[8], n8t[8];
            for (int i = 0; i < 8; i++) {
                n8c[i] = c0[i];
                n8t[i] = c1[i];
            }
            const Singles sx = ld_global(p.singles, X), sy = ld_global(p.singles, Y), sz = ld_global(p.singles, Z);
            uint32_t xy[9], xz[9], yz[9], cas[27], ctl[27];
            pair9_lookup(p.pairs, p.n, sx, sy, X, Y, 0, (uint32_t)p.N0, xy);
            pair9_lookup(p.pairs, p.n, sx, sz, X, Z, 0, (uint32_t)p.N0, xz);
            pair9_lookup(p.pairs, p.n, sy, sz, Y, Z, 0, (uint32_t)p.N0, yz);
            triplet27(n8c, xy, xz, yz, cas);
            pair9_lookup(p.pairs, p.n, sx, sy, X, Y, 1, (uint32_t)p.N1, xy);
            pair9_lookup(p.pairs, p.n, sx, sz, X, Z, 1, (uint32_t)p.N1, xz);
            pair9_lookup(p.pairs, p.n, sy, sz, Y, Z, 1, (uint32_t)p.N1, yz);
            triplet27(n8t, xy, xz, yz, ctl);
            bad = p.score == SCORE_K2 ? k2_f64(cas, ctl, 27, p.N0, p.N1, p.t64)
                                      : -mi_f64(cas, ctl, 27, p.N0, p.N1, p.t64);
        }
    } else {
        const int A = p.yb * 256 + yl, B = p.zb * 256 + zl;
        valid = A < B && B < p.n;
        packed = (unsigned long long)A | ((unsigned long long)B << 32);
        if (valid) {
            const uint4 q0 = *(const uint4 *)(p.counts + ((size_t)(0 * 256 + yl) * 256 + zl) * 4);
            const uint4 q1 = *(const uint4 *)(p.counts + ((size_t)(1 * 256 + yl) * 256 + zl) * 4);
            const Singles sa = ld_global(p.singles, A), sb = ld_global(p.singles, B);
            uint32_t cas[9], ctl[9];
            pair9(q0.x, q0.y, q0.z, q0.w, sa.x, sa.y, sb.x, sb.y, (uint32_t)p.N0, cas);
            pair9(q1.x, q1.y, q1.z, q1.w, sa.z, sa.w, sb.z, sb.w, (uint32_t)p.N1, ctl);
            bad = p.score == SCORE_K2 ? k2_f64(cas, ctl, 9, p.N0, p.N1, p.t64) : -mi_f64(cas, ctl, 9, p.N0, p.N1, p.t64);
        }
    }
    unsigned long long key = ~0ull, pk = ~0ull;
    if (valid) {
        if (bad != bad) *p.nan_flag = 1u;         // inconsistent tables: reported, never ranked
        else {
            key = d2key(bad);
            pk = packed;
        }
    }
    wave_lexmin(key, pk);
    const int w = zl >> 6;
    if ((zl & 63) == 0) {
        s_k[w] = key;
        s_p[w] = pk;
    }
    __syncthreads();
    if (zl == 0) {
        unsigned long long k = s_k[0], q = s_p[0];
        for (int i = 1; i < 4; i++)
            if (lex_less(s_k[i], s_p[i], k, q)) {
                k = s_k[i];
                q = s_p[i];
            }
        p.blk_out[2 * yl] = k;
        p.blk_out[2 * yl + 1] = q;
    }
}

// Folds a stream's last round into its best (one wave, launched on that
// stream after its last objective kernel).
__global__ __launch_bounds__(64) void k_plain_fold(const unsigned long long *blk, unsigned long long *best)
{
    plain_fold(blk, best);
}

// Fills n 64-bit words with v (no memset node, so a captured sequence stays
// replay-safe).
__global__ void k_fill_u64(unsigned long long *dst, int n, unsigned long long v)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[i] = v;
}

// ======================================================= FP64 rescoring ==

struct RescoreArgs {
    const uint32_t *p0, *p1;
    int W0, W1, n, N0, N1, order, score;
    const uint4 *singles;
    const uint4 *pairs;                 // 3-way only
    const double *t64;
    const unsigned long long *cand;     // packed sets
    int count;
    double *out_s64;                    // FP64 score (not badness), may be null
    uint32_t *out_cells;                // [count][2][3^k], may be null
};

// One wave per set: counts the 2^k AND-cells of both classes straight from
// the planes, derives every cell, scores in FP64.
__device__ __forceinline__ void rescore_one(const RescoreArgs &p, int i)
{
    const int lane = threadIdx.x;
    const unsigned long long pk = p.cand[i];
    int s[3];
    if (p.order == 3) {
        s[0] = (int)(pk & 0x1FFFFF);
        s[1] = (int)((pk >> 21) & 0x1FFFFF);
        s[2] = (int)((pk >> 42) & 0x1FFFFF);
    } else {
        s[0] = (int)(pk & 0xFFFFFFFFu);
        s[1] = (int)(pk >> 32);
        s[2] = 0;
    }
    uint32_t n8[2][8];
    for (int cls = 0; cls < 2; cls++) {
        const uint32_t *base = cls ? p.p1 : p.p0;
        const int W = cls ? p.W1 : p.W0;
        uint32_t c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        const uint32_t *r0 = base + (size_t)s[0] * 2 * W, *r1 = base + (size_t)s[1] * 2 * W;
        const uint32_t *r2 = base + (size_t)s[2] * 2 * W;
        for (int w = lane; w < W; w += 64) {
            if (p.order == 3) {
#pragma unroll
                for (int x = 0; x < 2; x++)
#pragma unroll
                    for (int y = 0; y < 2; y++)
#pragma unroll
                        for (int z = 0; z < 2; z++)
                            c[x * 4 + y * 2 + z] += __popc(r0[x * W + w] & r1[y * W + w] & r2[z * W + w]);
            } else {
#pragma unroll
                for (int x = 0; x < 2; x++)
#pragma unroll
                    for (int y = 0; y < 2; y++) c[x * 2 + y] += __popc(r0[x * W + w] & r1[y * W + w]);
            }
        }
#pragma unroll
        for (int k = 0; k < 8; k++) n8[cls][k] = wave_sum_u(c[k]);
    }
    if (lane != 0) return;
    uint32_t cas[27], ctl[27];
    int nc;
    if (p.order == 3) {
        nc = 27;
        const Singles sx = ld_global(p.singles, s[0]), sy = ld_global(p.singles, s[1]), sz = ld_global(p.singles, s[2]);
        uint32_t xy[9], xz[9], yz[9];
        pair9_lookup(p.pairs, p.n, sx, sy, s[0], s[1], 0, (uint32_t)p.N0, xy);
        pair9_lookup(p.pairs, p.n, sx, sz, s[0], s[2], 0, (uint32_t)p.N0, xz);
        pair9_lookup(p.pairs, p.n, sy, sz, s[1], s[2], 0, (uint32_t)p.N0, yz);
        triplet27(n8[0], xy, xz, yz, cas);
        pair9_lookup(p.pairs, p.n, sx, sy, s[0], s[1], 1, (uint32_t)p.N1, xy);
        pair9_lookup(p.pairs, p.n, sx, sz, s[0], s[2], 1, (uint32_t)p.N1, xz);
        pair9_lookup(p.pairs, p.n, sy, sz, s[1], s[2], 1, (uint32_t)p.N1, yz);
        triplet27(n8[1], xy, xz, yz, ctl);
    } else {
        nc = 9;
        const Singles sa = ld_global(p.singles, s[0]), sb = ld_global(p.singles, s[1]);
        pair9(n8[0][0], n8[0][1], n8[0][2], n8[0][3], sa.x, sa.y, sb.x, sb.y, (uint32_t)p.N0, cas);
        pair9(n8[1][0], n8[1][1], n8[1][2], n8[1][3], sa.z, sa.w, sb.z, sb.w, (uint32_t)p.N1, ctl);
    }
    if (p.out_cells) {
        uint32_t *o = p.out_cells + (size_t)i * 2 * nc;
        for (int c = 0; c < nc; c++) {
            o[c] = cas[c];
            o[nc + c] = ctl[c];
        }
    }
    if (p.out_s64)
        p.out_s64[i] =
            p.score == SCORE_K2 ? k2_f64(cas, ctl, nc, p.N0, p.N1, p.t64) : mi_f64(cas, ctl, nc, p.N0, p.N1, p.t64);
}

__global__ __launch_bounds__(64) void k_rescore(RescoreArgs p)
{
    if ((int)blockIdx.x < p.count) rescore_one(p, blockIdx.x);
}

// FP64 re-scoring of the candidates one search appended, with their count
// read on the device (cand[0]; nothing when the overflow flag cand[1] is set
// or the count exceeds cap: the host then searches again in windows), so it
// follows the search on the stream with no host round trip.
// Its outcome also goes straight to pinned host memory (device-mapped), so
// the host reads it after one stream synchronisation, with no copy: the
// candidate count and overflow flag, and the first `slots` candidates.
struct HostStage {
    unsigned *ctl;                    // [0] candidate count, [1] overflow flag
    unsigned long long *packed;       // [slots] packed indices
    float *s32;                       // [slots] FP32 scores from the search
    double *s64;                      // [slots] FP64 scores
    unsigned slots;
};
__global__ __launch_bounds__(64) void k_rescore_after(RescoreArgs p, const unsigned *cand, unsigned cap,
                                                      const float *cand_s32, HostStage hs)
{
    const unsigned count = cand[0], over = cand[1];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        hs.ctl[0] = count;
        hs.ctl[1] = over;
    }
    if (over || count > cap) return;
    for (unsigned i = blockIdx.x; i < count; i += gridDim.x) {
        rescore_one(p, (int)i);
        // lane 0 wrote the FP64 score of set i just now
        if (threadIdx.x == 0 && i < hs.slots) {
            hs.packed[i] = p.cand[i];
            hs.s32[i] = cand_s32[i];
            hs.s64[i] = p.out_s64[i];
        }
    }
}

// FP64 re-scoring of the candidates of many searches; the candidate counts
// are read on the device, so the whole batch needs no host round trip.
struct RescoreJob {
    RescoreArgs r;                 // r.count unused
    const unsigned *count;         // candidates appended by the search, and the
                                   // overflow flag (CandOut::cand_count)
    unsigned cap;                  // buffer capacity (larger counts: overflow)
};

__global__ __launch_bounds__(64) void k_rescore_batch(const RescoreJob *jobs, int njobs)
{
    for (unsigned long long gi = blockIdx.x;; gi += gridDim.x) {
        unsigned long long start = 0;
        int j = 0;
        for (; j < njobs; j++) {
            const unsigned c = jobs[j].count[0];
            const unsigned long long cnt = (c > jobs[j].cap || jobs[j].count[1]) ? 0ull : (unsigned long long)c;
            if (gi < start + cnt) break;
            start += cnt;
        }
        if (j == njobs) return;
        rescore_one(jobs[j].r, (int)(gi - start));
    }
}

// Resets the control words of several searches at once (graph-capturable).
__global__ void k_init_many(unsigned long long *const *queue, unsigned *const *best_key, unsigned *const *cand_count,
                            int n, unsigned long long *batch_queues, int nq)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        *queue[i] = 0ull;
        *best_key[i] = 0xFF800000u;
        cand_count[i][0] = 0u;
        cand_count[i][1] = 0u;
    }
    if (i < nq) batch_queues[i] = 0ull;
}

// FP64 value of mi_dterm's D(x, q n) for the counts as mi_dterm sees them
// (x, n rounded to FP32, exact below 2^24) and the same FP32 share q: the
// series for |e| <= 1/4 (to k = 30, below 1e-18 relative), x ln(x/y) - d
// otherwise, so that no cancellation limits the reference.
__device__ double mi_dterm_f64(double x, double n, float q)
{
    if (n == 0.0) return 0.0;
    const double y = (double)q * n;
    if (x == 0.0) return y;
    const double d = x - y, e = d / y;
    if (fabs(e) <= 0.25) {
        double s = 0.0;
        for (int k = 30; k >= 2; k--) s = s * -e + 1.0 / ((double)k * (double)(k - 1));
        return d * e * s;
    }
    return x * log(x / y) - d;
}

// Largest error of the FP32 score terms over the arguments a search can
// reach, against FP64.  K2: |k2_term(m) - (T64[m] - c m)| for every m in
// [0, N+1] (absolute).  MI: the relative error |D32 - D64| / D64 of
// mi_dterm (against its own inputs: counts past 2^24 round to FP32 first,
// which the host's bound covers apart) for both classes' shares q0, q1 and
// every cell size n in [1, N]
// (every 2^(k)-th n past 2^22 samples, k the smallest that leaves 2^22 of
// them), at the counts x where its two forms meet (|e| = 1/4), around the
// expected count, at 0, n and a spread in between; the terms are >= 0, so
// a table's MI errs by at most this relative error plus the summation's.
// *out holds the bits of a non-negative double, which atomicMax on the bits
// orders correctly.
__global__ void k_term_error(int score, int N, float invN, float k0, float c32, double c, const double *t64,
                             float q0, float q1, unsigned long long *out)
{
    __shared__ float s_lnf[LNF_SMALL];
    load_lnf(s_lnf);
    __syncthreads();
    double e = 0.0;
    if (score == SCORE_K2) {
        for (int m = blockIdx.x * blockDim.x + threadIdx.x; m <= N + 1; m += gridDim.x * blockDim.x) {
            const double exact = t64[m] - c * (double)m;
            e = fmax(e, fabs((double)k2_term((uint32_t)m, invN, k0, c32, s_lnf) - exact));
        }
    } else {
        int step = 1;
        while ((N - 1) / step + 1 > (1 << 22)) step *= 2;
        for (long long i = blockIdx.x * blockDim.x + threadIdx.x; 1 + i * step <= N; i += gridDim.x * blockDim.x) {
            const uint32_t n = (uint32_t)(1 + i * step);
            for (int cls = 0; cls < 2; cls++) {
                const float q = cls ? q1 : q0;
                const double y = (double)q * (double)n;
                const double at[8] = {0.0, (double)n, y, 0.75 * y, 1.25 * y, 0.5 * y, 2.0 * y, 0.1 * y};
                for (int a = 0; a < 8; a++) {
                    const long long x0 = (long long)floor(at[a]);
                    for (int j = (a < 2 ? 0 : -3); j <= (a < 2 ? 0 : 3); j++) {
                        const long long x = x0 + j;
                        if (x < 0 || x > (long long)n) continue;
                        const double ref = mi_dterm_f64((double)(float)x, (double)(float)n, q);
                        const double got = (double)mi_dterm((uint32_t)x, n, q);
                        if (ref > 0.0) e = fmax(e, fabs(got - ref) / ref);
                        else if (got != 0.0) e = INFINITY;
                    }
                }
            }
        }
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) e = fmax(e, __shfl_xor(e, o, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(e));
}

// Small fill kernels (no hipMemsetAsync, so the sequence stays capturable).
// queue0: first work unit; best0: starting global bound (key), 0 for +inf.
// queue[0] is the work queue, queue[1..8] the per-XCD queues (fetch_xcd):
// all start at unit queue0.
__global__ void k_init_search(unsigned long long *queue, unsigned *best_key, unsigned *cand_count,
                              unsigned long long *best_key64, unsigned long long *checksum,
                              unsigned long long queue0, unsigned best0)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        *queue = queue0;
        for (unsigned x = 0; x < 8; x++) queue[1 + x] = queue0 > x ? (queue0 - x + 7) / 8 : 0ull;
        *best_key = best0 ? best0 : 0xFF800000u; // key of +inf
        cand_count[0] = 0u;
        cand_count[1] = 0u;                      // overflow flag
        *best_key64 = 0xFFF0000000000000ull;     // key of +inf (double)
        *checksum = 0ull;
    }
}

}  // namespace episdet
