// kernels_mfma_pairs.hpp -- fused 2-way search counting on the matrix cores.
//
// The four AND-counts of every pair (A, B) are one dense contraction over
// samples:
//   C[(A,i), (B,j)] = sum_s A_i(s) * B_j(s)
// a GEMM with rows (A, plane i), columns (B, plane j) and K = samples.  It
// runs on v_mfma_scale_f32_32x32x64_f8f6f4 with the FP4 encoding of the
// 3-way kernel (class-split nibble masks, E8M0 scales that make every
// product exactly 1.0, exact FP32 accumulation below 2^24 samples per
// class); here the A operand is a plane row itself, with no X & Y product.
// The reference builds these counts as a CUTLASS b1 XOR-popcount GEMM plus
// a recovery step (helper.cu:714-750, 636-637); its pair tables follow
// helper.cu:644-662.
//
// Work unit: 128 A SNPs x 64 B SNPs (A tile tA, B tile tB, tA <= tB / 2 so
// that some A < B exists).  512 threads = 8 waves: waves 0-3 contract the
// cases, 4-7 the controls; wave (class, wy) owns A rows 64wy..64wy+63 (32
// SNPs x 2 planes) x all 128 B rows = 2 x 4 MFMA tiles, 128 accumulators, so
// each A operand feeds four MFMAs and each B operand two.  The rows of a
// 512-sample chunk (256 A rows + 128 B rows, 64 bytes each) go straight from
// global memory to LDS (buffer_load_dwordx4 ... lds), double-buffered; the four
// 16-byte slots of row r are XOR-swizzled by (r >> 1) & 3, so the eight
// consecutive rows a group of lanes reads at one logical slot cover all
// banks once.  After the last chunk the accumulators go to LDS in two
// halves (B 0-31, then B 32-63) and all 512 threads build and score eight
// pair tables per half.
#pragma once
#include "kernels_mfma.hpp"

namespace episdet {

constexpr int P2_TA = 128, P2_TB = 64;
constexpr int P2_KW = 16;                                   // words per chunk (512 samples)
constexpr int P2_AROWS = 2 * P2_TA;                         // 256
constexpr int P2_ROWS = P2_AROWS + 2 * P2_TB;               // 384 rows of a class
constexpr int P2_ROWS_PER_LOAD = 64 * 4 / P2_KW;            // 16: one load instruction = 1 KB
constexpr int P2_LOADS = P2_ROWS / P2_ROWS_PER_LOAD;        // 24 per class and chunk
constexpr int P2_BUF_WORDS = P2_ROWS * P2_KW;               // 6144 (24 KB)
constexpr int P2_STAGE_WORDS = 4 * P2_BUF_WORDS;            // 2 classes x 2 buffers (96 KB)
constexpr int P2_DUMP_WORDS = 2 * 4 * 4 * 16 * 64;          // class x wave x tile x reg x lane (128 KB)
constexpr int P2_SMEM_WORDS = P2_STAGE_WORDS > P2_DUMP_WORDS ? P2_STAGE_WORDS : P2_DUMP_WORDS;
static_assert(P2_LOADS % 4 == 0, "loads split over the four waves of a class");
static_assert(P2_ROWS_PER_LOAD == 16 && P2_KW == 16, "swizzle assumes 4 slots per row");

// Units of the B-tile-major enumeration: B tile tB holds floor(tB/2)+1 units
// (tA = 0..floor(tB/2)), so units before B tile 2m are m(m+1) and before
// 2m+1 are (m+1)^2.
__host__ __device__ inline unsigned long long p2_units(int ntb)
{
    const unsigned long long m = (unsigned long long)(ntb / 2);
    return (ntb & 1) ? (m + 1) * (m + 1) : m * (m + 1);
}

__device__ __forceinline__ void p2_unrank(unsigned long long u, int &ta, int &tb)
{
    long long m = (long long)((sqrt(4.0 * (double)u + 1.0) - 1.0) * 0.5);
    while ((unsigned long long)(m + 1) * (unsigned long long)(m + 2) <= u) m++;
    while ((unsigned long long)m * (unsigned long long)(m + 1) > u) m--;
    const unsigned long long e = (unsigned long long)m * (unsigned long long)(m + 1);
    if (u < e + (unsigned long long)(m + 1)) {
        tb = (int)(2 * m);
        ta = (int)(u - e);
    } else {
        tb = (int)(2 * m + 1);
        ta = (int)(u - (unsigned long long)(m + 1) * (unsigned long long)(m + 1));
    }
}

// The same units in XCD blocks (PairArgs::blocked): 512-SNP super-tiles on
// both axes (4 A tiles, 8 B tiles), S = ceil(n / 512) a side, NB =
// S (S + 1) / 2 blocks of 32 units in triangular order.  Unit u = 8 v + x
// belongs to XCD x, which walks blocks x, x + 8, x + 16, ... 32 units at a
// time, one per resident workgroup of this one-per-CU kernel, so the units
// an XCD runs at once share 4 A and 8 B tiles in its L2.  Units with no
// A < B, past the last tile, or in a block past NB are empty.
__host__ __device__ inline unsigned long long p2_blocked_units(int n)
{
    const unsigned long long S = (unsigned long long)((n + 511) / 512), NB = S * (S + 1) / 2;
    return 256ull * ((NB + 7) / 8);
}

__host__ __device__ inline bool p2_blocked_unrank(unsigned long long u, int n, int &ta, int &tb)
{
    const unsigned long long v = u >> 3, B = 8 * (v >> 5) + (u & 7);
    const unsigned long long S = (unsigned long long)((n + 511) / 512);
    if (B >= S * (S + 1) / 2) return false;
    int sa, sb;
    unrank_tri(B, sa, sb);
    const int k = (int)(v & 31);
    ta = 4 * sa + (k >> 3);
    tb = 8 * sb + (k & 7);
    return ta <= tb / 2 && tb < (n + P2_TB - 1) / P2_TB;
}

// Chunk loads through a scalar buffer descriptor (see m2_row_offsets): the
// row offsets of a unit are computed once, each chunk adds its word offset
// as the scalar soffset.  One instruction fills sixteen 64-byte LDS rows:
// lane L writes physical slot L%4 of row L/4, which holds logical slot
// (L%4) ^ ((L/4 >> 1) & 3).
constexpr int P2_LOADS_PER_WAVE = P2_LOADS / 4;

__device__ __forceinline__ void p2_row_offsets(int W, int a0, int b0, int wih, int lane,
                                               unsigned (&roff)[P2_LOADS_PER_WAVE])
{
    const int rl = lane >> 2, q = (lane & 3) ^ ((rl >> 1) & 3);
#pragma unroll
    for (int k = 0; k < P2_LOADS_PER_WAVE; k++) {
        const int r = (wih + 4 * k) * P2_ROWS_PER_LOAD + rl;
        const int row = r < P2_AROWS ? 2 * a0 + r : 2 * b0 + (r - P2_AROWS);
        roff[k] = ((unsigned)row * (unsigned)W + (unsigned)(q * 4)) * 4u;
    }
}

__device__ __forceinline__ void p2_issue_buf(i32x4_t rsrc, const unsigned (&roff)[P2_LOADS_PER_WAVE], int w0,
                                             unsigned buf, int wih)
{
#pragma unroll
    for (int k = 0; k < P2_LOADS_PER_WAVE; k++) {
        const int ii = wih + 4 * k;
        lds_dma16(rsrc, buf + (unsigned)(ii * P2_ROWS_PER_LOAD * P2_KW * 4), roff[k], (unsigned)(w0 * 4));
    }
}

// Word offset of logical 16-byte slot q of LDS row r.
__device__ __forceinline__ int p2_at(int r, int q) { return r * P2_KW + ((q ^ ((r >> 1) & 3)) << 2); }

template <int C>
__device__ __forceinline__ void p2_step(const uint4 (&aw)[2], const uint4 (&bw)[4], v16f_t (&acc)[2][4])
{
    constexpr int s = mf_scale<C>();
    const v8i_t A0 = mf_operand<C>(aw[0]), A1 = mf_operand<C>(aw[1]);
#pragma unroll
    for (int cb = 0; cb < 4; cb++) {
        const v8i_t B = mf_operand<C>(bw[cb]);
        acc[0][cb] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(A0, B, acc[0][cb], 4, 4, 0, s, 0, s);
        acc[1][cb] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(A1, B, acc[1][cb], 4, 4, 0, s, 0, s);
    }
}

template <int MODE>
__global__ __launch_bounds__(512) void k_pairs_mfma(BatchPair bp)
{
    __shared__ __attribute__((aligned(16))) uint32_t smem[P2_SMEM_WORDS];
    __shared__ unsigned long long s_u;
    __shared__ float s_red[16];
    __shared__ float s_lnf[scores_sets(MODE) ? LNF_SMALL : 1];
    if constexpr (scores_sets(MODE)) load_lnf(s_lnf);
    const int tid = threadIdx.x;
    const int cls = tid >> 8, t = tid & 255;
    const int wy = t >> 6, lane = t & 63, h = lane >> 5, i = lane & 31;
    uint32_t *buf0 = smem + (cls * 2 + 0) * P2_BUF_WORDS;
    uint32_t *buf1 = smem + (cls * 2 + 1) * P2_BUF_WORDS;
    // LDS rows read by this lane: A rows of its two row tiles, B rows of the
    // four column tiles (tile row / column i = 2 * snp + plane)
    const int ar0 = 64 * wy + i, ar1 = 64 * wy + 32 + i;
    int br[4];
#pragma unroll
    for (int cb = 0; cb < 4; cb++) br[cb] = P2_AROWS + 32 * cb + i;

    int j = 0;
    for (int iter = 0;; iter++) {
        int a0, b0;
        if constexpr (MODE == MODE_COUNTS) {
            // one 128 x 64 tile of round blockIdx.x / 8, no queue
            if (iter > 0) break;
            const int r = (int)blockIdx.x >> 3, tl = (int)blockIdx.x & 7;
            if (r >= bp.rounds) break;
            a0 = bp.ab * 256 + (tl >> 2) * P2_TA;
            b0 = (bp.bb0 + r) * 256 + (tl & 3) * P2_TB;
        } else {
            if (tid == 0) s_u = bp.xcd ? fetch_xcd(bp.queue, bp.total) : atomicAdd(bp.queue, 1ull);
            __syncthreads();
            const unsigned long long uu = s_u;
            if (uu >= bp.total) break;
            if (uu < bp.job_start[j]) j = 0;              // per-XCD queues go out of order
            while (j + 1 < bp.njobs && bp.job_start[j + 1] <= uu) j++;
            const unsigned long long u = uu - bp.job_start[j];
            int tA, tB;
            if (!bp.jobs[j].blocked) {
                p2_unrank(u, tA, tB);
            } else if (!p2_blocked_unrank(u, bp.jobs[j].n, tA, tB)) {
                __syncthreads();   // every thread has read s_u before the next fetch writes it
                continue;
            }
            a0 = tA * P2_TA;
            b0 = tB * P2_TB;
        }
        const PairArgs &p = bp.jobs[j];
        const uint32_t *base = cls ? p.p1 : p.p0;
        const int W = cls ? p.W1 : p.W0;
        const int nch_mine = W / P2_KW;
        const int nch = (p.W0 > p.W1 ? p.W0 : p.W1) / P2_KW;
        const i32x4_t rsrc = m2_rsrc(base, cls ? p.bytes1 : p.bytes0);
        unsigned roff[P2_LOADS_PER_WAVE];
        p2_row_offsets(W, a0, b0, wy, lane, roff);

        v16f_t acc[2][4];
#pragma unroll
        for (int rb = 0; rb < 2; rb++)
#pragma unroll
            for (int cb = 0; cb < 4; cb++)
#pragma unroll
                for (int r = 0; r < 16; r++) acc[rb][cb][r] = 0.f;

        if (nch_mine > 0) p2_issue_buf(rsrc, roff, 0, lds_addr(buf0), wy);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        for (int c = 0; c < nch; c++) {
            uint32_t *cur = (c & 1) ? buf1 : buf0;
            if (c + 1 < nch_mine) p2_issue_buf(rsrc, roff, (c + 1) * P2_KW, lds_addr((c & 1) ? buf0 : buf1), wy);
            if (c < nch_mine) {
#pragma unroll
                for (int s = 0; s < P2_KW / 8; s++) {
                    const int q = 2 * s + h;
                    uint4 aw[2], bw[4];
                    aw[0] = *(const uint4 *)&cur[p2_at(ar0, q)];
                    aw[1] = *(const uint4 *)&cur[p2_at(ar1, q)];
#pragma unroll
                    for (int cb = 0; cb < 4; cb++) bw[cb] = *(const uint4 *)&cur[p2_at(br[cb], q)];
                    p2_step<0>(aw, bw, acc);
                    p2_step<1>(aw, bw, acc);
                    p2_step<2>(aw, bw, acc);
                    p2_step<3>(aw, bw, acc);
                }
            }
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
        }

        float bad[16];
        unsigned long long packed[16];
        unsigned long long hsum = 0;
        // sets whose badness lower bound is above this cannot become
        // candidates; their FP32 score is skipped (skip_threshold)
        float skip_above = INFINITY;
        if constexpr (scores_sets(MODE)) skip_above = skip_threshold(p.sc, p.co.best_key);
#pragma unroll
        for (int hf = 0; hf < 2; hf++) {
            // tiles with column blocks 2hf, 2hf+1 -> LDS [cls][wy][t4][reg][lane]
#pragma unroll
            for (int rb = 0; rb < 2; rb++)
#pragma unroll
                for (int cr = 0; cr < 2; cr++)
#pragma unroll
                    for (int r = 0; r < 16; r++)
                        smem[((((cls * 4 + wy) * 4 + rb * 2 + cr) * 16) + r) * 64 + lane] =
                            (uint32_t)acc[rb][2 * hf + cr][r];
            __syncthreads();
#pragma unroll
            for (int q8 = 0; q8 < 8; q8++) {
                const int m = hf * 8 + q8;
                const int T = q8 * 512 + tid;                   // 128 A x 32 B of this half
                const int al = T >> 5, bl = hf * 32 + (T & 31);
                const int A = a0 + al, B = b0 + bl;
                bad[m] = INFINITY;
                packed[m] = (unsigned long long)A | ((unsigned long long)B << 32);
                // the plain path's objective kernel checks the pair itself
                if (MODE != MODE_COUNTS && !(A < B && B < p.n)) continue;
                // row 2*al + pi lives in wave al/32, row tile (al/16)%2, tile
                // row rho = 2*(al%16) + pi; column 2*bl' + pj in column tile
                // bl/16, tile column 2*(bl%16) + pj.  Output register
                // r = 4*(rho/8) + rho%4 of lane 32*((rho/4)%2) + column.
                const int twy = al >> 5, trb = (al >> 4) & 1, a16 = al & 15;
                const int tcr = (bl >> 4) - 2 * hf, b16 = bl & 15;
                const int t4 = trb * 2 + tcr;
                uint32_t n4c[4], n4t[4];
#pragma unroll
                for (int pi = 0; pi < 2; pi++)
#pragma unroll
                    for (int pj = 0; pj < 2; pj++) {
                        const int rho = 2 * a16 + pi;
                        const int reg = 4 * (rho >> 3) + (rho & 3);
                        const int ln = 32 * ((rho >> 2) & 1) + 2 * b16 + pj;
                        const int off = ((t4 * 16) + reg) * 64 + ln;
                        n4c[pi * 2 + pj] = smem[(0 * 4 + twy) * 4 * 16 * 64 + off];
                        n4t[pi * 2 + pj] = smem[(1 * 4 + twy) * 4 * 16 * 64 + off];
                    }
                if constexpr (MODE == MODE_PAIRTABLE) {
                    uint4 *o = p.pairs + ((size_t)A * p.n + B) * 2;
                    o[0] = make_uint4(n4c[0], n4c[1], n4c[2], n4c[3]);
                    o[1] = make_uint4(n4t[0], n4t[1], n4t[2], n4t[3]);
                } else if constexpr (MODE == MODE_COUNTS) {
                    const int ar = A - bp.ab * 256, bbr = B / 256 - bp.bb0, br = B % 256;
                    uint32_t *rc = bp.counts + (size_t)bbr * bp.slot;
                    *(uint4 *)(rc + ((size_t)(0 * 256 + ar) * 256 + br) * 4) = make_uint4(n4c[0], n4c[1], n4c[2], n4c[3]);
                    *(uint4 *)(rc + ((size_t)(1 * 256 + ar) * 256 + br) * 4) = make_uint4(n4t[0], n4t[1], n4t[2], n4t[3]);
                } else {
                    const Singles sa = ld_global(p.singles, A), sb = ld_global(p.singles, B);
                    uint32_t cas[9], ctl[9];
                    pair9(n4c[0], n4c[1], n4c[2], n4c[3], sa.x, sa.y, sb.x, sb.y, (uint32_t)p.N0, cas);
                    pair9(n4t[0], n4t[1], n4t[2], n4t[3], sa.z, sa.w, sb.z, sb.w, (uint32_t)p.N1, ctl);
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
            __syncthreads();
        }
        if constexpr (MODE == MODE_SEARCH) publish_candidates<16, 8>(p.co, p.sc.margin, bad, packed, s_red);
        if constexpr (MODE == MODE_TOPK) publish_topk<16, 8>(p.co, p.sc.margin, bad, packed);
        if constexpr (MODE == MODE_CHECKSUM) {
            if (hsum) atomicAdd(p.checksum, hsum);
        }
    }
}

}  // namespace episdet
