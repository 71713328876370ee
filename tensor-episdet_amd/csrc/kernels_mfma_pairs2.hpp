// kernels_mfma_pairs2.hpp -- fused 2-way search on the matrix cores, two
// workgroups per CU, classes in sequence.
//
// Same contraction, FP4 encoding and pair tables as k_pairs_mfma.  That
// kernel's 512-thread workgroup and 128 KB of LDS leave one workgroup per
// CU, so the matrix cores idle while it builds and scores the tables of a
// unit; and since its cases and controls waves run side by side, the waves
// of the smaller class finish early when the classes differ in size.  This
// version:
//  * work unit: 64 A SNPs x 64 B SNPs (A tile <= B tile), in triangular
//    order, or from 16 tiles up in XCD blocks taken from per-XCD queues
//    (blocked_unrank, fetch_xcd), so that the units one XCD runs at once
//    share their rows in its L2;
//  * 256 threads = 4 waves; wave w owns A SNPs 32(w/2)..+31 (64 rows, 2 row
//    tiles) x B SNPs 32(w%2)..+31 (64 rows, 2 column tiles), for both
//    classes: 2 x 2 x 2 MFMA tiles, 128 accumulators;
//  * the chunks of the cases, then those of the controls, form one
//    sequence that all four waves contract together, so every wave does the
//    same work whatever the class sizes;
//  * a 1024-sample chunk is 128 A rows + 128 B rows of 128 bytes (32 KB),
//    loaded straight to LDS through the scalar buffer descriptor of its
//    class, double-buffered: chunk t + 1 is issued after the barrier that
//    starts chunk t;
//  * after the last chunk the accumulators go to LDS in two halves (column
//    tile 0, then 1, of every wave) and the 256 threads build and score
//    eight pair tables per half.
// LDS: 64 KB stage / dump area plus the ln m! table, so two workgroups fit a
// CU and one's scoring runs beside the other's counting.
#pragma once
#include "kernels_mfma_pairs.hpp"
#include "kernels_mfma4.hpp"

namespace episdet {

constexpr int P3_TA = 64, P3_TB = 64;
constexpr int P3_KW = 32;                                   // words per chunk (1024 samples)
constexpr int P3_AROWS = 2 * P3_TA;                         // 128
constexpr int P3_ROWS = P3_AROWS + 2 * P3_TB;               // 256 rows of a chunk
constexpr int P3_ROWS_PER_LOAD = 8;                         // 128-byte rows: one load = 1 KB
constexpr int P3_LOADS = P3_ROWS / P3_ROWS_PER_LOAD;        // 32 per chunk
constexpr int P3_LOADS_PER_WAVE = P3_LOADS / 4;             // 8
constexpr int P3_BUF_WORDS = P3_ROWS * P3_KW;               // 8192 (32 KB)
constexpr int P3_STAGE_WORDS = 2 * P3_BUF_WORDS;            // double-buffered (64 KB)
constexpr int P3_DUMP_WORDS = 2 * 4 * 2 * 16 * 64;          // class x wave x row tile x reg x lane (64 KB)
constexpr int P3_SMEM_WORDS = P3_STAGE_WORDS > P3_DUMP_WORDS ? P3_STAGE_WORDS : P3_DUMP_WORDS;
static_assert(P3_KW == M2_KW, "128-byte rows use the swizzle of m2_at");

// Byte offsets, within a class's planes, of slot q of each LDS row this
// lane loads (load k of wave w fills rows 8(w + 4k) .. +7); one set per
// class, since the classes' word counts differ.
__device__ __forceinline__ void p3_row_offsets(int W, int a0, int b0, int w, int lane,
                                               unsigned (&roff)[P3_LOADS_PER_WAVE])
{
    const int rl = lane >> 3, q = (lane & 7) ^ rl;
#pragma unroll
    for (int k = 0; k < P3_LOADS_PER_WAVE; k++) {
        const int r = (w + 4 * k) * P3_ROWS_PER_LOAD + rl;
        const int row = r < P3_AROWS ? 2 * a0 + r : 2 * b0 + (r - P3_AROWS);
        roff[k] = ((unsigned)row * (unsigned)W + (unsigned)(q * 4)) * 4u;
    }
}

__device__ __forceinline__ void p3_issue_buf(i32x4_t rsrc, const unsigned (&roff)[P3_LOADS_PER_WAVE], int w0,
                                             unsigned buf, int w)
{
#pragma unroll
    for (int k = 0; k < P3_LOADS_PER_WAVE; k++)
        lds_dma16(rsrc, buf + (unsigned)((w + 4 * k) * P3_ROWS_PER_LOAD * P3_KW * 4), roff[k], (unsigned)(w0 * 4));
}

template <int C>
__device__ __forceinline__ void p3_step(const uint4 (&aw)[2], const uint4 (&bw)[2], v16f_t (&acc)[2][2])
{
    constexpr int s = mf_scale<C>();
    const v8i_t A0 = mf_operand<C>(aw[0]), A1 = mf_operand<C>(aw[1]);
#pragma unroll
    for (int cb = 0; cb < 2; cb++) {
        const v8i_t B = mf_operand<C>(bw[cb]);
        acc[0][cb] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(A0, B, acc[0][cb], 4, 4, 0, s, 0, s);
        acc[1][cb] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(A1, B, acc[1][cb], 4, 4, 0, s, 0, s);
    }
}

// Contracts one staged chunk into one class's accumulators.
__device__ __forceinline__ void p3_chunk(const uint32_t *cur, int ar0, int ar1, int br0, int br1, int h,
                                         v16f_t (&acc)[2][2])
{
#pragma unroll
    for (int s = 0; s < P3_KW / 8; s++) {
        const int q = 2 * s + h;
        uint4 aw[2], bw[2];
        aw[0] = *(const uint4 *)&cur[m2_at(ar0, q)];
        aw[1] = *(const uint4 *)&cur[m2_at(ar1, q)];
        bw[0] = *(const uint4 *)&cur[m2_at(br0, q)];
        bw[1] = *(const uint4 *)&cur[m2_at(br1, q)];
        p3_step<0>(aw, bw, acc);
        p3_step<1>(aw, bw, acc);
        p3_step<2>(aw, bw, acc);
        p3_step<3>(aw, bw, acc);
    }
}

template <int MODE>
__global__ __launch_bounds__(256, 2) void k_pairs_mfma2(BatchPair bp)
{
    static_assert(MODE != MODE_PAIRTABLE, "the pair table is built by k_pairs_mfma");
    __shared__ __attribute__((aligned(16))) uint32_t smem[P3_SMEM_WORDS];
    __shared__ unsigned long long s_u;
    __shared__ float s_red[8];
    __shared__ float s_lnf[scores_sets(MODE) ? LNF_SMALL : 1];
    if constexpr (scores_sets(MODE)) load_lnf(s_lnf);
    const int tid = threadIdx.x;
    const int w = tid >> 6, lane = tid & 63, h = lane >> 5, i = lane & 31;
    const int rp = w >> 1, cp = w & 1;
    // LDS rows read by this lane: A rows of its two row tiles, B rows of its
    // two column tiles (tile row / column i = 2 * snp + plane)
    const int ar0 = 64 * rp + i, ar1 = 64 * rp + 32 + i;
    const int br0 = P3_AROWS + 64 * cp + i, br1 = P3_AROWS + 64 * cp + 32 + i;

    int j = 0;
    for (;;) {
        if (tid == 0) s_u = bp.xcd ? fetch_xcd(bp.queue, bp.total) : atomicAdd(bp.queue, 1ull);
        __syncthreads();
        const unsigned long long uu = s_u;
        if (uu >= bp.total) break;
        if (uu < bp.job_start[j]) j = 0;                  // per-XCD queues go out of order
        while (j + 1 < bp.njobs && bp.job_start[j + 1] <= uu) j++;
        const PairArgs &p = bp.jobs[j];
        const unsigned long long u = uu - bp.job_start[j];
        int tA, tB;
        if (!p.blocked) {
            unrank_tri(u, tA, tB);
        } else if (!blocked_unrank(u, p.ntiles, tA, tB)) {
            __syncthreads();       // every thread has read s_u before the next fetch writes it
            continue;
        }
        const int a0 = tA * P3_TA, b0 = tB * P3_TB;
        const int nch0 = p.W0 / P3_KW, nch = nch0 + p.W1 / P3_KW;
        const i32x4_t rs0 = m2_rsrc(p.p0, p.bytes0), rs1 = m2_rsrc(p.p1, p.bytes1);
        unsigned roff0[P3_LOADS_PER_WAVE], roff1[P3_LOADS_PER_WAVE];
        p3_row_offsets(p.W0, a0, b0, w, lane, roff0);
        p3_row_offsets(p.W1, a0, b0, w, lane, roff1);

        v16f_t acc[2][2][2];                                // [class][row tile][column tile]
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int rb = 0; rb < 2; rb++)
#pragma unroll
                for (int cb = 0; cb < 2; cb++)
#pragma unroll
                    for (int r = 0; r < 16; r++) acc[c][rb][cb][r] = 0.f;

        // chunk t of the sequence: cases chunk t, or controls chunk t - nch0
        if (nch > 0) {
            if (nch0 > 0) p3_issue_buf(rs0, roff0, 0, lds_addr(smem), w);
            else p3_issue_buf(rs1, roff1, 0, lds_addr(smem), w);
        }
        // two loops, one per class, so each contracts into one fixed set of
        // accumulators
#pragma unroll
        for (int c = 0; c < 2; c++) {
            for (int t = c ? nch0 : 0; t < (c ? nch : nch0); t++) {
                // this wave's loads of chunk t are its only outstanding ones
                __builtin_amdgcn_s_waitcnt(m4_wait_vm(0));
                __syncthreads();
                const int tn = t + 1;
                if (tn < nch) {
                    uint32_t *nb = smem + (tn & 1) * P3_BUF_WORDS;
                    if (tn < nch0) p3_issue_buf(rs0, roff0, tn * P3_KW, lds_addr(nb), w);
                    else p3_issue_buf(rs1, roff1, (tn - nch0) * P3_KW, lds_addr(nb), w);
                }
                p3_chunk(smem + (t & 1) * P3_BUF_WORDS, ar0, ar1, br0, br1, h, acc[c]);
            }
        }
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();

        float bad[16];
        unsigned long long packed[16];
        unsigned long long hsum = 0;
        // sets whose badness lower bound is above this cannot become
        // candidates; their FP32 score is skipped (skip_threshold)
        float skip_above = INFINITY;
        if constexpr (scores_sets(MODE)) skip_above = skip_threshold(p.sc, p.co.best_key);
#pragma unroll
        for (int hf = 0; hf < 2; hf++) {
            // column tile hf of every wave -> LDS [class][wave][row tile][reg][lane]
#pragma unroll
            for (int c = 0; c < 2; c++)
#pragma unroll
                for (int rb = 0; rb < 2; rb++)
#pragma unroll
                    for (int r = 0; r < 16; r++)
                        smem[((((c * 4 + w) * 2 + rb) * 16) + r) * 64 + lane] = (uint32_t)acc[c][rb][hf][r];
            __syncthreads();
#pragma unroll
            for (int q8 = 0; q8 < 8; q8++) {
                const int m = hf * 8 + q8;
                const int T = q8 * 256 + tid;                   // 64 A x 32 B of this half
                const int al = T >> 5, bsel = T & 31;
                const int tcp = bsel >> 4, b16 = bsel & 15;
                const int bl = 32 * tcp + 16 * hf + b16;
                const int A = a0 + al, B = b0 + bl;
                bad[m] = INFINITY;
                packed[m] = (unsigned long long)A | ((unsigned long long)B << 32);
                if (!(A < B && B < p.n)) continue;
                // row 2*al + pi lives in the waves of row pair al/32, row tile
                // (al/16)%2, tile row rho = 2*(al%16) + pi; column 2*bl + pj
                // in the wave of column pair bl/32, column tile hf, tile
                // column 2*(bl%16) + pj.  Output register r = 4*(rho/8) +
                // rho%4 of lane 32*((rho/4)%2) + column.
                const int tw = (al >> 5) * 2 + tcp, trb = (al >> 4) & 1, a16 = al & 15;
                uint32_t n4c[4], n4t[4];
#pragma unroll
                for (int pi = 0; pi < 2; pi++)
#pragma unroll
                    for (int pj = 0; pj < 2; pj++) {
                        const int rho = 2 * a16 + pi;
                        const int reg = 4 * (rho >> 3) + (rho & 3);
                        const int ln = 32 * ((rho >> 2) & 1) + 2 * b16 + pj;
                        const int off = (((tw * 2 + trb) * 16) + reg) * 64 + ln;
                        n4c[pi * 2 + pj] = smem[0 * 4 * 2 * 16 * 64 + off];
                        n4t[pi * 2 + pj] = smem[1 * 4 * 2 * 16 * 64 + off];
                    }
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
            __syncthreads();
        }
        if constexpr (MODE == MODE_SEARCH) publish_candidates<16, 4>(p.co, p.sc.margin, bad, packed, s_red);
        if constexpr (MODE == MODE_TOPK) publish_topk<16, 4>(p.co, p.sc.margin, bad, packed);
        if constexpr (MODE == MODE_CHECKSUM) {
            if (hsum) atomicAdd(p.checksum, hsum);
        }
    }
}

}  // namespace episdet
