// kernels_mfma3.hpp -- fused 3-way search, FP4 MFMA, two workgroups per CU.
//
// Same contraction, FP4 encoding and epilogue as version 2 (kernels_mfma.hpp),
// cut so that two independent workgroups share each CU: while one contracts,
// the other's epilogue (table derivation and FP32 scoring, which leaves the
// matrix cores idle) and barriers run beside it.
//  * work unit: X x 64 Y x 32 Z;
//  * 256 threads = 4 waves: waves 0-1 contract the cases, 2-3 the controls;
//    wave (class, wy) owns 32 Y x 32 Z = 4 x 2 MFMA tiles (128 accumulators),
//    so each A operand feeds two MFMAs and each B operand four;
//  * the X, Y and Z rows of a 512-sample chunk (2 + 128 + 64 rows of 64
//    bytes) go straight to LDS through the scalar buffer descriptor,
//    double-buffered; the four 16-byte slots of row r are XOR-swizzled by
//    (r >> 1) & 3;
//  * after the last chunk the accumulators go to LDS in two halves (Z 0-15,
//    then 16-31) and the 256 threads score four triplets per half.
// LDS is 64 KB for the stage / dump area plus the ln m! table, so two
// workgroups fit the 160 KB of a CU; 256 VGPRs per lane at two waves per
// SIMD.
#pragma once
#include "kernels_mfma.hpp"

namespace episdet {

constexpr int M3_TY = 64, M3_TZ = 32;
constexpr int M3_KW = 16;                                   // words per chunk (512 samples)
constexpr int M3_YROW = 2, M3_ZROW = 2 + 2 * M3_TY;         // first Y row, first Z row
constexpr int M3_ROWS = M3_ZROW + 2 * M3_TZ;                // 194 rows of a class
constexpr int M3_ROWS_PER_LOAD = 64 * 4 / M3_KW;            // 16: one load instruction = 1 KB
constexpr int M3_LOADS = (M3_ROWS + M3_ROWS_PER_LOAD - 1) / M3_ROWS_PER_LOAD;   // 13
constexpr int M3_BUF_WORDS = M3_LOADS * M3_ROWS_PER_LOAD * M3_KW;              // 3328 (13 KB)
constexpr int M3_STAGE_WORDS = 4 * M3_BUF_WORDS;            // 2 classes x 2 buffers
constexpr int M3_DUMP_WORDS = 2 * 2 * 4 * 16 * 64;          // class x wave x tile x reg x lane (64 KB)
constexpr int M3_SMEM_WORDS = M3_STAGE_WORDS > M3_DUMP_WORDS ? M3_STAGE_WORDS : M3_DUMP_WORDS;
constexpr int M3_LOADS_PER_WAVE = (M3_LOADS + 1) / 2;       // two waves per class
static_assert(M3_KW == 16 && M3_ROWS_PER_LOAD == 16, "swizzle assumes four 16-byte slots per row");

// Byte offsets, within the class's planes, of slot q of each LDS row this
// lane loads (load k of wave wih fills rows 16(wih + 2k) .. +15).
__device__ __forceinline__ void m3_row_offsets(int W, int X, int y0, int z0, int wih, int lane,
                                               unsigned (&roff)[M3_LOADS_PER_WAVE])
{
    const int rl = lane >> 2, q = (lane & 3) ^ ((rl >> 1) & 3);
#pragma unroll
    for (int k = 0; k < M3_LOADS_PER_WAVE; k++) {
        const int r = (wih + 2 * k) * M3_ROWS_PER_LOAD + rl;
        int row;
        if (r < M3_YROW) row = 2 * X + r;
        else if (r < M3_ZROW) row = 2 * y0 + (r - M3_YROW);
        else if (r < M3_ROWS) row = 2 * z0 + (r - M3_ZROW);
        else row = 2 * X;                     // padding rows are never read
        roff[k] = ((unsigned)row * (unsigned)W + (unsigned)(q * 4)) * 4u;
    }
}

__device__ __forceinline__ void m3_issue_buf(i32x4_t rsrc, const unsigned (&roff)[M3_LOADS_PER_WAVE], int w0,
                                             unsigned buf, int wih)
{
#pragma unroll
    for (int k = 0; k < M3_LOADS_PER_WAVE; k++) {
        const int ii = wih + 2 * k;
        if (ii < M3_LOADS) lds_dma16(rsrc, buf + (unsigned)(ii * M3_ROWS_PER_LOAD * M3_KW * 4), roff[k], (unsigned)(w0 * 4));
    }
}

// Word offset of logical 16-byte slot q of LDS row r.
__device__ __forceinline__ int m3_at(int r, int q) { return r * M3_KW + ((q ^ ((r >> 1) & 3)) << 2); }

template <int C>
__device__ __forceinline__ void m3_step(const uint4 &xw, const uint4 (&yw)[4], const uint4 (&zw)[2],
                                        v16f_t (&acc)[4][2], unsigned sh)
{
    constexpr int s = mf_scale<C>();
    const v8i_t B0 = mf_operand<C>(zw[0]), B1 = mf_operand<C>(zw[1]);
#pragma unroll
    for (int rb = 0; rb < 4; rb++) {
        const v8i_t A = m2_aop<C>(xw, yw[rb], sh);
        acc[rb][0] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(A, B0, acc[rb][0], 4, 4, 0, s, 0, s);
        acc[rb][1] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(A, B1, acc[rb][1], 4, 4, 0, s, 0, s);
    }
}

template <int MODE>
__global__ __launch_bounds__(256, 2) void k_triplets_mfma3(BatchTri bt)
{
    __shared__ __attribute__((aligned(16))) uint32_t smem[M3_SMEM_WORDS];
    __shared__ unsigned long long s_u;
    __shared__ float s_red[8];
    __shared__ float s_lnf[scores_sets(MODE) ? LNF_SMALL : 1];
    if constexpr (scores_sets(MODE)) load_lnf(s_lnf);
    const int tid = threadIdx.x;
    const int cls = tid >> 7, t = tid & 127;
    const int wy = t >> 6, lane = t & 63, h = lane >> 5, i = lane & 31;
    uint32_t *buf0 = smem + (cls * 2 + 0) * M3_BUF_WORDS;
    uint32_t *buf1 = smem + (cls * 2 + 1) * M3_BUF_WORDS;
    // LDS rows read by this lane: X plane, Y rows of its four row blocks, Z
    // rows of the two column blocks
    const int xr = (i >> 1) & 1;
    int yr[4];
#pragma unroll
    for (int rb = 0; rb < 4; rb++) yr[rb] = M3_YROW + 2 * (wy * 32 + rb * 8 + (i >> 2)) + (i & 1);
    const int zr0 = M3_ZROW + i, zr1 = M3_ZROW + 32 + i;

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
        const uint32_t *base = cls ? p.p1 : p.p0;
        const int W = cls ? p.W1 : p.W0;
        const int nch_mine = W / M3_KW;
        const int nch = (p.W0 > p.W1 ? p.W0 : p.W1) / M3_KW;
        g = next_group(p.gstart, p.ngroups, u, g);
        const int X = (int)(u - p.gstart[g]);
        const int2 gt = p.groups[g];
        const int y0 = gt.x * M3_TY, z0 = gt.y * M3_TZ;
        const i32x4_t rsrc = m2_rsrc(base, cls ? p.bytes1 : p.bytes0);
        unsigned roff[M3_LOADS_PER_WAVE];
        m3_row_offsets(W, X, y0, z0, wy, lane, roff);

        v16f_t acc[4][2];
#pragma unroll
        for (int rb = 0; rb < 4; rb++)
#pragma unroll
            for (int cb = 0; cb < 2; cb++)
#pragma unroll
                for (int r = 0; r < 16; r++) acc[rb][cb][r] = 0.f;

        if (nch_mine > 0) m3_issue_buf(rsrc, roff, 0, lds_addr(buf0), wy);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        for (int c = 0; c < nch; c++) {
            uint32_t *cur = (c & 1) ? buf1 : buf0;
            if (c + 1 < nch_mine) m3_issue_buf(rsrc, roff, (c + 1) * M3_KW, lds_addr((c & 1) ? buf0 : buf1), wy);
            if (c < nch_mine) {
#pragma unroll
                for (int s = 0; s < M3_KW / 8; s++) {
                    const int q = 2 * s + h;
                    const uint4 xw = *(const uint4 *)&cur[m3_at(xr, q)];
                    uint4 yw[4], zw[2];
#pragma unroll
                    for (int rb = 0; rb < 4; rb++) yw[rb] = *(const uint4 *)&cur[m3_at(yr[rb], q)];
                    zw[0] = *(const uint4 *)&cur[m3_at(zr0, q)];
                    zw[1] = *(const uint4 *)&cur[m3_at(zr1, q)];
                    m3_step<0>(xw, yw, zw, acc, bt.sh);
                    m3_step<1>(xw, yw, zw, acc, bt.sh);
                    m3_step<2>(xw, yw, zw, acc, bt.sh);
                    m3_step<3>(xw, yw, zw, acc, bt.sh);
                }
            }
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
        }

        const Singles sx = ld_global(p.singles, X);
        float bad[8];
        unsigned long long packed[8];
        unsigned long long hsum = 0;
        float skip_above = INFINITY;     // see skip_threshold
        if constexpr (scores_sets(MODE)) skip_above = skip_threshold(p.sc, p.co.best_key);
#pragma unroll
        for (int hf = 0; hf < 2; hf++) {
            // column block hf of every row block -> LDS [cls][wy][rb][reg][lane]
#pragma unroll
            for (int rb = 0; rb < 4; rb++)
#pragma unroll
                for (int r = 0; r < 16; r++)
                    smem[((((cls * 2 + wy) * 4 + rb) * 16) + r) * 64 + lane] = (uint32_t)acc[rb][hf][r];
            __syncthreads();
#pragma unroll
            for (int q4 = 0; q4 < 4; q4++) {
                const int m = hf * 4 + q4;
                const int T = q4 * 256 + tid;                   // 64 Y x 16 Z of this half
                const int yl = T >> 4, zl = hf * 16 + (T & 15);
                const int Y = y0 + yl, Z = z0 + zl;
                bad[m] = INFINITY;
                packed[m] = (unsigned long long)X | ((unsigned long long)Y << 21) | ((unsigned long long)Z << 42);
                if (!(X < Y && Y < Z && Z < p.n)) continue;
                // tile row 4*y8 + ij of row block (yl/8)%4 in wave yl/32;
                // tile column 2*(zl%16) + k.  Output register
                // r = 4*(row/8) + row%4 of lane 32*((row/4)%2) + column.
                const int twy = yl >> 5, trb = (yl >> 3) & 3, y8 = yl & 7;
                const int z16 = zl & 15;
                uint32_t n8c[8], n8t[8];
#pragma unroll
                for (int ij = 0; ij < 4; ij++)
#pragma unroll
                    for (int kk = 0; kk < 2; kk++) {
                        const int reg = 4 * (y8 >> 1) + ij;
                        const int ln = 32 * (y8 & 1) + 2 * z16 + kk;
                        const int off = ((trb * 16) + reg) * 64 + ln;
                        n8c[ij * 2 + kk] = smem[(0 * 2 + twy) * 4 * 16 * 64 + off];
                        n8t[ij * 2 + kk] = smem[(1 * 2 + twy) * 4 * 16 * 64 + off];
                    }
                const Singles sy = ld_global(p.singles, Y);
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
