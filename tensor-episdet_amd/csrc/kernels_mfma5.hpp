// kernels_mfma5.hpp -- fused 3-way search, FP4 MFMA, two workgroups per CU,
// cases and controls contracted in sequence.
//
// Version 4 runs its cases and controls waves side by side, two of each per
// workgroup, so when one class holds more samples than the other, the waves
// of the smaller class (and the SIMDs they run on) finish early.  This
// version keeps version 4's unit and epilogue and gives every wave a share
// of both classes:
//  * work unit: (X, X+1) x 32 Y x 32 Z;
//  * 256 threads = 4 waves; wave w contracts X + w/2 against Y row blocks
//    2(w%2) and 2(w%2)+1 (16 Y) and both Z column blocks, for both classes:
//    2 x 2 x 2 MFMA tiles, 128 accumulators;
//  * the 1024-sample chunks of the cases, then those of the controls, form
//    one sequence that all four waves contract together; a chunk is 4 X +
//    64 Y + 64 Z rows of 128 bytes, loaded straight to LDS through its
//    class's scalar buffer descriptor into one of three buffers, two chunks
//    ahead, each wave waiting only for its own loads;
//  * after the last chunk the accumulators go to LDS in two halves (Z 0-15,
//    then 16-31) and the 256 threads score four triplets per half.
// Each A operand feeds two MFMAs instead of version 4's four, so the operand
// work per MFMA is higher; AUTO picks this version when the classes differ
// in size (see choose_kernel).
// LDS: 64 KB for the stage (3 buffers of 17 KB) / dump area plus the ln m!
// table, so two workgroups fit a CU.
#pragma once
#include "kernels_mfma4.hpp"

namespace episdet {

constexpr int M5_KW = 32;                                   // words per chunk (1024 samples)
constexpr int M5_ROWS_PER_LOAD = 8;                         // 128-byte rows: one load = 1 KB
constexpr int M5_LOADS = (M4_ROWS + M5_ROWS_PER_LOAD - 1) / M5_ROWS_PER_LOAD;   // 17
constexpr int M5_LOADS_PER_WAVE = (M5_LOADS + 3) / 4;       // 5 (wave 0) or 4
constexpr int M5_BUF_WORDS = M5_LOADS * M5_ROWS_PER_LOAD * M5_KW;              // 4352 (17 KB)
constexpr int M5_NBUF = 3;
constexpr int M5_STAGE_WORDS = M5_NBUF * M5_BUF_WORDS;
constexpr int M5_DUMP_WORDS = 2 * 4 * 2 * 16 * 64;          // class x wave x row block x reg x lane (64 KB)
constexpr int M5_SMEM_WORDS = M5_STAGE_WORDS > M5_DUMP_WORDS ? M5_STAGE_WORDS : M5_DUMP_WORDS;
static_assert(M5_KW == M2_KW, "128-byte rows use the swizzle of m2_at");

// Byte offsets, within a class's planes, of slot q of each LDS row this
// lane loads (load k of wave w fills rows 8(w + 4k) .. +7); one set per
// class, since the classes' word counts differ.
__device__ __forceinline__ void m5_row_offsets(int W, int X0, int y0, int z0, int w, int lane,
                                               unsigned (&roff)[M5_LOADS_PER_WAVE])
{
    const int rl = lane >> 3, q = (lane & 7) ^ rl;
#pragma unroll
    for (int k = 0; k < M5_LOADS_PER_WAVE; k++) {
        const int r = (w + 4 * k) * M5_ROWS_PER_LOAD + rl;
        int row;
        if (r < M4_YROW) row = 2 * X0 + r;
        else if (r < M4_ZROW) row = 2 * y0 + (r - M4_YROW);
        else if (r < M4_ROWS) row = 2 * z0 + (r - M4_ZROW);
        else row = 2 * X0;                    // padding rows are never read
        roff[k] = ((unsigned)row * (unsigned)W + (unsigned)(q * 4)) * 4u;
    }
}

__device__ __forceinline__ void m5_issue_buf(i32x4_t rsrc, const unsigned (&roff)[M5_LOADS_PER_WAVE], int w0,
                                             unsigned buf, int w)
{
#pragma unroll
    for (int k = 0; k < M5_LOADS_PER_WAVE; k++) {
        const int ii = w + 4 * k;
        if (ii < M5_LOADS) lds_dma16(rsrc, buf + (unsigned)(ii * M5_ROWS_PER_LOAD * M5_KW * 4), roff[k], (unsigned)(w0 * 4));
    }
}

template <int C>
__device__ __forceinline__ void m5_step(const uint4 &xw, const uint4 (&yw)[2], const uint4 (&zw)[2],
                                        v16f_t (&acc)[2][2], unsigned sh)
{
    constexpr int s = mf_scale<C>();
    const v8i_t B0 = mf_operand<C>(zw[0]), B1 = mf_operand<C>(zw[1]);
#pragma unroll
    for (int rb = 0; rb < 2; rb++) {
        const v8i_t A = m2_aop<C>(xw, yw[rb], sh);
        acc[rb][0] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(A, B0, acc[rb][0], 4, 4, 0, s, 0, s);
        acc[rb][1] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(A, B1, acc[rb][1], 4, 4, 0, s, 0, s);
    }
}

// Contracts one staged chunk into one class's accumulators.
__device__ __forceinline__ void m5_chunk(const uint32_t *cur, int xr, const int (&yr)[2], int zr0, int zr1, int h,
                                         v16f_t (&acc)[2][2], unsigned sh)
{
#pragma unroll
    for (int s = 0; s < M5_KW / 8; s++) {
        const int q = 2 * s + h;
        const uint4 xw = *(const uint4 *)&cur[m2_at(xr, q)];
        uint4 yw[2], zw[2];
        yw[0] = *(const uint4 *)&cur[m2_at(yr[0], q)];
        yw[1] = *(const uint4 *)&cur[m2_at(yr[1], q)];
        zw[0] = *(const uint4 *)&cur[m2_at(zr0, q)];
        zw[1] = *(const uint4 *)&cur[m2_at(zr1, q)];
        m5_step<0>(xw, yw, zw, acc, sh);
        m5_step<1>(xw, yw, zw, acc, sh);
        m5_step<2>(xw, yw, zw, acc, sh);
        m5_step<3>(xw, yw, zw, acc, sh);
    }
}

template <int MODE>
__global__ __launch_bounds__(256, 2) void k_triplets_mfma5(BatchTri bt)
{
    __shared__ __attribute__((aligned(16))) uint32_t smem[M5_SMEM_WORDS];
    __shared__ unsigned long long s_u;
    __shared__ float s_red[8];
    __shared__ float s_lnf[scores_sets(MODE) ? LNF_SMALL : 1];
    if constexpr (scores_sets(MODE)) load_lnf(s_lnf);
    const int tid = threadIdx.x;
    const int w = tid >> 6, lane = tid & 63, h = lane >> 5, i = lane & 31;
    const int xs = w >> 1, yh = w & 1;
    // LDS rows read by this lane: its X's plane, Y rows of its two row
    // blocks, Z rows of the two column blocks
    const int xr = 2 * xs + ((i >> 1) & 1);
    int yr[2];
#pragma unroll
    for (int rl = 0; rl < 2; rl++) yr[rl] = M4_YROW + 2 * ((2 * yh + rl) * 8 + (i >> 2)) + (i & 1);
    const int zr0 = M4_ZROW + i, zr1 = M4_ZROW + 32 + i;
    // loads per chunk issued by this wave
    const int my_loads = (M5_LOADS - w + 3) / 4;

    int g = 0, j = 0;
    for (;;) {
        if (tid == 0) s_u = atomicAdd(bt.queue, 1ull);
        __syncthreads();
        const unsigned long long uu = s_u;
        if (uu >= bt.total) break;
        // job of this unit; a workgroup's units only increase
        int jn = j;
        while (jn + 1 < bt.njobs && bt.job_start[jn + 1] <= uu) jn++;
        if (jn != j) {
            j = jn;
            g = 0;
        }
        const TriArgs &p = bt.jobs[j];
        const unsigned long long u = uu - bt.job_start[j];
        const int nch0 = p.W0 / M5_KW, nch = nch0 + p.W1 / M5_KW;
        g = next_group(p.gstart, p.ngroups, u, g);
        const int X0 = 2 * (int)(u - p.gstart[g]);
        const int2 gt = p.groups[g];
        const int y0 = gt.x * M4_TY, z0 = gt.y * M4_TZ;
        const i32x4_t rs0 = m2_rsrc(p.p0, p.bytes0), rs1 = m2_rsrc(p.p1, p.bytes1);
        unsigned roff0[M5_LOADS_PER_WAVE], roff1[M5_LOADS_PER_WAVE];
        m5_row_offsets(p.W0, X0, y0, z0, w, lane, roff0);
        m5_row_offsets(p.W1, X0, y0, z0, w, lane, roff1);

        v16f_t acc[2][2][2];                                // [class][row block][column block]
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int rb = 0; rb < 2; rb++)
#pragma unroll
                for (int cb = 0; cb < 2; cb++)
#pragma unroll
                    for (int r = 0; r < 16; r++) acc[c][rb][cb][r] = 0.f;

        // chunk t of the sequence: cases chunk t, or controls chunk t - nch0;
        // chunks 0 and 1 in flight, chunk t + 2 goes out once every wave is
        // past chunk t - 1, whose buffer it reuses
#pragma unroll
        for (int k = 0; k < M5_NBUF - 1; k++) {
            if (k < nch) {
                uint32_t *nb = smem + k * M5_BUF_WORDS;
                if (k < nch0) m5_issue_buf(rs0, roff0, k * M5_KW, lds_addr(nb), w);
                else m5_issue_buf(rs1, roff1, (k - nch0) * M5_KW, lds_addr(nb), w);
            }
        }
        // two loops, one per class, so each contracts into one fixed set of
        // accumulators
#pragma unroll
        for (int c = 0; c < 2; c++) {
            for (int t = c ? nch0 : 0; t < (c ? nch : nch0); t++) {
                // this wave's loads of chunk t are done once at most those of
                // chunk t + 1 are outstanding
                if (t + 1 < nch) {
                    if (my_loads == M5_LOADS_PER_WAVE) __builtin_amdgcn_s_waitcnt(m4_wait_vm(M5_LOADS_PER_WAVE));
                    else __builtin_amdgcn_s_waitcnt(m4_wait_vm(M5_LOADS_PER_WAVE - 1));
                } else {
                    __builtin_amdgcn_s_waitcnt(m4_wait_vm(0));
                }
                __syncthreads();
                const int tn = t + M5_NBUF - 1;
                if (tn < nch) {
                    uint32_t *nb = smem + (tn % M5_NBUF) * M5_BUF_WORDS;
                    if (tn < nch0) m5_issue_buf(rs0, roff0, tn * M5_KW, lds_addr(nb), w);
                    else m5_issue_buf(rs1, roff1, (tn - nch0) * M5_KW, lds_addr(nb), w);
                }
                m5_chunk(smem + (t % M5_NBUF) * M5_BUF_WORDS, xr, yr, zr0, zr1, h, acc[c], bt.sh);
            }
        }
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();

        float bad[8];
        unsigned long long packed[8];
        unsigned long long hsum = 0;
        // Sets whose badness lower bound exceeds the best FP32 score found so
        // far by more than the slack cannot become candidates; their FP32
        // score is skipped (skip_threshold).
        float skip_above = INFINITY;
        if constexpr (scores_sets(MODE)) skip_above = skip_threshold(p.sc, p.co.best_key);
#pragma unroll
        for (int hf = 0; hf < 2; hf++) {
            // column block hf -> LDS [class][wave][row block][reg][lane]
#pragma unroll
            for (int c = 0; c < 2; c++)
#pragma unroll
                for (int rb = 0; rb < 2; rb++)
#pragma unroll
                    for (int r = 0; r < 16; r++)
                        smem[((((c * 4 + w) * 2 + rb) * 16) + r) * 64 + lane] = (uint32_t)acc[c][rb][hf][r];
            __syncthreads();
#pragma unroll
            for (int q4 = 0; q4 < 4; q4++) {
                const int m = hf * 4 + q4;
                const int T = q4 * 256 + tid;                   // 2 X x 32 Y x 16 Z of this half
                const int txs = T >> 9, yl = (T >> 4) & 31, zl = hf * 16 + (T & 15);
                const int X = X0 + txs, Y = y0 + yl, Z = z0 + zl;
                bad[m] = INFINITY;
                packed[m] = (unsigned long long)X | ((unsigned long long)Y << 21) | ((unsigned long long)Z << 42);
                if (!(X < Y && Y < Z && Z < p.n)) continue;
                // row block yl/8 = 2*yh + rbl lives in wave 2*txs + yh as its
                // row block rbl; tile row 4*y8 + ij, tile column 2*(zl%16) + k.
                // Output register r = 4*(row/8) + row%4 of lane
                // 32*((row/4)%2) + column.
                const int rbg = yl >> 3, y8 = yl & 7;
                const int tw = 2 * txs + (rbg >> 1), rbl = rbg & 1;
                const int z16 = zl & 15;
                uint32_t n8c[8], n8t[8];
#pragma unroll
                for (int ij = 0; ij < 4; ij++)
#pragma unroll
                    for (int kk = 0; kk < 2; kk++) {
                        const int reg = 4 * (y8 >> 1) + ij;
                        const int ln = 32 * (y8 & 1) + 2 * z16 + kk;
                        const int off = (((tw * 2 + rbl) * 16) + reg) * 64 + ln;
                        n8c[ij * 2 + kk] = smem[0 * 4 * 2 * 16 * 64 + off];
                        n8t[ij * 2 + kk] = smem[1 * 4 * 2 * 16 * 64 + off];
                    }
                const Singles sx = ld_global(p.singles, X), sy = ld_global(p.singles, Y);
                uint32_t cas[27], ctl[27];
                triplet_tables(p, X, Y, Z, sx, sy, n8c, n8t, cas, ctl);
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
            __syncthreads();
        }
        if constexpr (MODE == MODE_SEARCH) publish_candidates<8, 4>(p.co, p.sc.margin, bad, packed, s_red);
        if constexpr (MODE == MODE_TOPK) publish_topk<8, 4>(p.co, p.sc.margin, bad, packed);
        if constexpr (MODE == MODE_CHECKSUM) {
            if (hsum) atomicAdd(p.checksum, hsum);
        }
    }
}

}  // namespace episdet
