// kernels_mfma4.hpp -- fused 3-way search, FP4 MFMA, two workgroups per CU,
// pairs of X per work unit, multi-buffered chunks.
//
// Same contraction, FP4 encoding and epilogue as versions 2 and 3.  Version
// 3 hides the epilogue by running two workgroups per CU but re-streams its Z
// rows twice as often as version 2.  This version keeps two workgroups per
// CU and wins the rows back by contracting two X at once against the same
// Y and Z rows:
//  * work unit: (X, X+1) x 32 Y x 32 Z;
//  * 256 threads = 4 waves: waves 0-1 the cases, 2-3 the controls; wave
//    (class, xs) owns X + xs against 32 Y x 32 Z = 4 x 2 MFMA tiles (128
//    accumulators);
//  * a chunk of a class is 4 X rows + 64 Y rows + 64 Z rows (132 rows for
//    4096 triplet-samples per class, against 194 in version 3), loaded
//    straight to LDS through the scalar buffer descriptor, NBUF - 1 chunks
//    ahead.  KW = 16 (tile 3): 512-sample chunks, three buffers of 9 KB per
//    class.  KW = 32 (tile 4): 1024-sample chunks, two buffers of 17 KB, so
//    half as many barriers and load rounds per sample;
//  * after the last chunk the accumulators go to LDS in two halves (Z 0-15,
//    then 16-31) and the 256 threads score four triplets per half.
// LDS: at most 68 KB for the stage / dump area plus the ln m! table, so two
// workgroups fit a CU.
#pragma once
#include "kernels_mfma3.hpp"

namespace episdet {

constexpr int M4_TY = 32, M4_TZ = 32;
constexpr int M4_YROW = 4, M4_ZROW = 4 + 2 * M4_TY;         // X rows 0-3, then Y, then Z
constexpr int M4_ROWS = M4_ZROW + 2 * M4_TZ;                // 132 rows of a class
constexpr int M4_DUMP_WORDS = 2 * 2 * 4 * 16 * 64;          // class x wave x tile x reg x lane (64 KB)

// Chunk geometry, by words per row per chunk (KW):
//  * KW 16 (512 samples): 64-byte rows, 16 rows per load, three buffers;
//  * KW 32 (1024 samples): 128-byte rows, 8 rows per load, two buffers,
//    half the chunks (and barriers) per sample.
// Both keep two workgroups per CU (stage and dump area <= 68 KB).
template <int KW>
struct M4Geo {
    static_assert(KW == 16 || KW == 32, "chunk of 16 or 32 words");
    static constexpr int SLOTS = KW / 4;                     // 16-byte slots per row
    static constexpr int ROWS_PER_LOAD = 64 / SLOTS;         // one load = 1 KB
    static constexpr int LOADS = (M4_ROWS + ROWS_PER_LOAD - 1) / ROWS_PER_LOAD;   // 9 / 17
    static constexpr int LOADS_PER_WAVE = (LOADS + 1) / 2;   // two waves per class
    static constexpr int BUF_WORDS = LOADS * ROWS_PER_LOAD * KW;
    static constexpr int NBUF = KW == 16 ? 3 : 2;
    static constexpr int STAGE_WORDS = 2 * NBUF * BUF_WORDS;
    static constexpr int SMEM_WORDS = STAGE_WORDS > M4_DUMP_WORDS ? STAGE_WORDS : M4_DUMP_WORDS;
    // swizzle: logical slot q of row r sits at physical slot q ^ swz(r), so
    // the rows one ds_read_b128 touches spread over every slot
    static __device__ __forceinline__ int swz(int r) { return KW == 16 ? ((r >> 1) & 3) : (r & 7); }
    static __device__ __forceinline__ int at(int r, int q) { return r * KW + ((q ^ swz(r)) << 2); }
};

// s_waitcnt immediates (gfx9 layout: vmcnt [3:0] and [15:14], expcnt [6:4],
// lgkmcnt [11:8]) that wait for vector-memory loads only.
constexpr int m4_wait_vm(int n) { return (n & 0xF) | (0x7 << 4) | (0xF << 8) | ((n >> 4) << 14); }

// Byte offsets, within the class's planes, of slot q of each LDS row this
// lane loads (load k of wave wih fills rows RPL(wih + 2k) .. +RPL-1).
template <int KW>
__device__ __forceinline__ void m4_row_offsets(int W, int X0, int y0, int z0, int wih, int lane,
                                               unsigned (&roff)[M4Geo<KW>::LOADS_PER_WAVE])
{
    using G = M4Geo<KW>;
    const int rl = lane / G::SLOTS;
#pragma unroll
    for (int k = 0; k < G::LOADS_PER_WAVE; k++) {
        const int r = (wih + 2 * k) * G::ROWS_PER_LOAD + rl;
        const int q = (lane % G::SLOTS) ^ G::swz(r);
        int row;
        if (r < M4_YROW) row = 2 * X0 + r;
        else if (r < M4_ZROW) row = 2 * y0 + (r - M4_YROW);
        else if (r < M4_ROWS) row = 2 * z0 + (r - M4_ZROW);
        else row = 2 * X0;                    // padding rows are never read
        roff[k] = ((unsigned)row * (unsigned)W + (unsigned)(q * 4)) * 4u;
    }
}

template <int KW>
__device__ __forceinline__ void m4_issue_buf(i32x4_t rsrc, const unsigned (&roff)[M4Geo<KW>::LOADS_PER_WAVE], int w0,
                                             unsigned buf, int wih)
{
    using G = M4Geo<KW>;
#pragma unroll
    for (int k = 0; k < G::LOADS_PER_WAVE; k++) {
        const int ii = wih + 2 * k;
        if (ii < G::LOADS) lds_dma16(rsrc, buf + (unsigned)(ii * G::ROWS_PER_LOAD * KW * 4), roff[k], (unsigned)(w0 * 4));
    }
}

template <int MODE, int KW>
__global__ __launch_bounds__(256, 2) void k_triplets_mfma4(BatchTri bt)
{
    using G = M4Geo<KW>;
    __shared__ __attribute__((aligned(16))) uint32_t smem[G::SMEM_WORDS];
    __shared__ unsigned long long s_u;
    __shared__ float s_red[8];
    __shared__ float s_lnf[scores_sets(MODE) ? LNF_SMALL : 1];
    if constexpr (scores_sets(MODE)) load_lnf(s_lnf);
    const int tid = threadIdx.x;
    const int cls = tid >> 7, t = tid & 127;
    const int xs = t >> 6, lane = t & 63, h = lane >> 5, i = lane & 31;
    uint32_t *bufs = smem + cls * G::NBUF * G::BUF_WORDS;
    // LDS rows read by this lane: its X's plane, Y rows of the four row
    // blocks, Z rows of the two column blocks
    const int xr = 2 * xs + ((i >> 1) & 1);
    int yr[4];
#pragma unroll
    for (int rb = 0; rb < 4; rb++) yr[rb] = M4_YROW + 2 * (rb * 8 + (i >> 2)) + (i & 1);
    const int zr0 = M4_ZROW + i, zr1 = M4_ZROW + 32 + i;
    // loads per chunk issued by this wave
    const int my_loads = (G::LOADS - xs + 1) / 2;

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
        const int nch_mine = W / KW;
        const int nch = (p.W0 > p.W1 ? p.W0 : p.W1) / KW;
        g = next_group(p.gstart, p.ngroups, u, g);
        const int X0 = 2 * (int)(u - p.gstart[g]);
        const int2 gt = p.groups[g];
        const int y0 = gt.x * M4_TY, z0 = gt.y * M4_TZ;
        const i32x4_t rsrc = m2_rsrc(base, cls ? p.bytes1 : p.bytes0);
        unsigned roff[G::LOADS_PER_WAVE];
        m4_row_offsets<KW>(W, X0, y0, z0, xs, lane, roff);

        v16f_t acc[4][2];
#pragma unroll
        for (int rb = 0; rb < 4; rb++)
#pragma unroll
            for (int cb = 0; cb < 2; cb++)
#pragma unroll
                for (int r = 0; r < 16; r++) acc[rb][cb][r] = 0.f;

        // NBUF - 1 chunks in flight; chunk c + NBUF - 1 goes out once every
        // wave is past chunk c - 1, whose buffer it reuses
#pragma unroll
        for (int k = 0; k + 1 < G::NBUF; k++)
            if (k < nch_mine) m4_issue_buf<KW>(rsrc, roff, k * KW, lds_addr(bufs + k * G::BUF_WORDS), xs);
        for (int c = 0; c < nch; c++) {
            // this wave's loads of chunk c are done once at most those of the
            // chunks after it that are already issued (NBUF - 2 at most) are
            // outstanding
            if (G::NBUF == 3 && c + 1 < nch_mine) {
                if (my_loads == G::LOADS_PER_WAVE) __builtin_amdgcn_s_waitcnt(m4_wait_vm(G::LOADS_PER_WAVE));
                else __builtin_amdgcn_s_waitcnt(m4_wait_vm(G::LOADS / 2));
            } else {
                __builtin_amdgcn_s_waitcnt(m4_wait_vm(0));
            }
            __syncthreads();
            if (c + G::NBUF - 1 < nch_mine)
                m4_issue_buf<KW>(rsrc, roff, (c + G::NBUF - 1) * KW,
                                 lds_addr(bufs + ((c + G::NBUF - 1) % G::NBUF) * G::BUF_WORDS), xs);
            if (c < nch_mine) {
                const uint32_t *cur = bufs + (c % G::NBUF) * G::BUF_WORDS;
                // the words of slice s + 1 are read once the class-3
                // operands of slice s are built, so their LDS latency
                // overlaps that class's MFMAs
                uint4 xw, yw[4], zw[2];
                auto ld = [&](int s) {
                    const int q = 2 * s + h;
                    xw = *(const uint4 *)&cur[G::at(xr, q)];
#pragma unroll
                    for (int rb = 0; rb < 4; rb++) yw[rb] = *(const uint4 *)&cur[G::at(yr[rb], q)];
                    zw[0] = *(const uint4 *)&cur[G::at(zr0, q)];
                    zw[1] = *(const uint4 *)&cur[G::at(zr1, q)];
                };
                ld(0);
#pragma unroll
                for (int s = 0; s < KW / 8; s++) {
                    m3_step<0>(xw, yw, zw, acc, bt.sh);
                    m3_step<1>(xw, yw, zw, acc, bt.sh);
                    m3_step<2>(xw, yw, zw, acc, bt.sh);
                    constexpr int s3 = mf_scale<3>();
                    const v8i_t B0 = mf_operand<3>(zw[0]), B1 = mf_operand<3>(zw[1]);
                    v8i_t A[4];
#pragma unroll
                    for (int rb = 0; rb < 4; rb++) A[rb] = m2_aop<3>(xw, yw[rb], bt.sh);
                    if (s + 1 < KW / 8) ld(s + 1);
#pragma unroll
                    for (int rb = 0; rb < 4; rb++) {
                        acc[rb][0] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(A[rb], B0, acc[rb][0], 4, 4, 0, s3,
                                                                                     0, s3);
                        acc[rb][1] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(A[rb], B1, acc[rb][1], 4, 4, 0, s3,
                                                                                     0, s3);
                    }
                }
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
            // column block hf of every row block -> LDS [cls][xs][rb][reg][lane]
#pragma unroll
            for (int rb = 0; rb < 4; rb++)
#pragma unroll
                for (int r = 0; r < 16; r++)
                    smem[((((cls * 2 + xs) * 4 + rb) * 16) + r) * 64 + lane] = (uint32_t)acc[rb][hf][r];
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
                // tile row 4*y8 + ij of row block yl/8 in wave (class, txs);
                // tile column 2*(zl%16) + k.  Output register
                // r = 4*(row/8) + row%4 of lane 32*((row/4)%2) + column.
                const int trb = yl >> 3, y8 = yl & 7;
                const int z16 = zl & 15;
                uint32_t n8c[8], n8t[8];
#pragma unroll
                for (int ij = 0; ij < 4; ij++)
#pragma unroll
                    for (int kk = 0; kk < 2; kk++) {
                        const int reg = 4 * (y8 >> 1) + ij;
                        const int ln = 32 * (y8 & 1) + 2 * z16 + kk;
                        const int off = ((trb * 16) + reg) * 64 + ln;
                        n8c[ij * 2 + kk] = smem[(0 * 2 + txs) * 4 * 16 * 64 + off];
                        n8t[ij * 2 + kk] = smem[(1 * 2 + txs) * 4 * 16 * 64 + off];
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
