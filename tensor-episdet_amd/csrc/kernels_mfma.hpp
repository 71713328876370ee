// kernels_mfma.hpp -- fused 3-way search counting on the matrix cores.
//
// The eight AND-counts of a triplet are a dense contraction over samples:
//   C[(Y,ij), (Z,k)] = sum_s P_ij(s) * Z_k(s),   P_ij = X_i & Y_j,
// i.e. a GEMM with rows (Y, ij), columns (Z, k) and K = samples, for a fixed
// X.  It runs on v_mfma_scale_f32_32x32x64_f8f6f4 with FP4 (E2M1) operands:
//
//  * bit -> FP4 without unpacking: masking a packed word with 0x11111111 << c
//    leaves, in every nibble, bit c of that nibble's 4 samples.  Classes c = 0,
//    1, 2 read as E2M1 values 0.5, 1, 2; class 3 would land on the sign bit,
//    so it is shifted down one place first (value 2).  A superstep of 8 words
//    per row gives four MFMAs, one per class, each contracting 64 samples.
//  * the per-lane E8M0 scales (2, 1, 1/2 for the classes) make every product
//    exactly 1.0, so every output is an integer count, accumulated exactly in
//    FP32 (requires fewer than 2^24 samples per class).
//  * A (rows) and B (columns) pair by lane half and nibble position, so the
//    sample-to-K mapping needs no shuffles: lane l of A serves row l%32, lane l
//    of B column l%32, lanes 0-31 the first four words, lanes 32-63 the next
//    four.  Output lane l, register r holds C[8(r/4)+4(l/32)+r%4][l%32].
//
// Workgroup: 8 waves on a 32 Y x 64 Z tile (the same work units as the VALU
// kernel's (2,4) tile); it takes a batch of jobs like version 2 below.  Waves 0-3 contract the cases, waves 4-7 the
// controls; wave (wy, wz) of a class owns 16 Y x 32 Z = 2 x 2 MFMA tiles, so
// each wave holds 64 accumulators.  Each half stages its own class: P and Z
// rows of a 2048-sample chunk go to LDS in row-major [row][68] layout
// (row = 4y+ij and 2z+k), so the lanes of a wave read consecutive rows:
// conflict-free ds_read_b128.  The next chunk's global loads are issued
// before the current chunk's MFMAs.  After the last chunk the accumulators
// of both classes go to LDS (over the staging buffers) and all 512 threads
// derive and score four triplets each from there.
#pragma once
#include "kernels_valu.hpp"

namespace episdet {

typedef int v8i_t __attribute__((ext_vector_type(8)));
typedef float v16f_t __attribute__((ext_vector_type(16)));

constexpr int MF_TY = 32, MF_TZ = 64;
constexpr int MF_KW = 64;                     // words per staged chunk (2048 samples)
constexpr int MF_LDSW = MF_KW + 4;            // row stride: 4-bank shift per row
constexpr int MF_ROWS = MF_TY * 4 + MF_TZ * 2;               // 256 rows per class
constexpr int MF_STAGE_WORDS = 2 * MF_ROWS * MF_LDSW;        // both classes
constexpr int MF_DUMP_WORDS = 2 * 4 * 4 * 16 * 64;           // class x wave x tile x reg x lane
constexpr int MF_SMEM_WORDS = MF_STAGE_WORDS > MF_DUMP_WORDS ? MF_STAGE_WORDS : MF_DUMP_WORDS;

struct MfmaStage {
    // one thread's share of a class chunk: X words at its column, 8 (y, w)
    // positions of Y0/Y1 and 8 uint4 of Z rows
    uint32_t x0, x1, y0v[8], y1v[8];
    uint4 z[8];
};

// Global loads of chunk w0 for one class, by the 256 threads t of that half.
// Rows are W words, a multiple of half a chunk: words at or past W (the second
// half of a last chunk) read as zero, which counts nothing.
__device__ __forceinline__ void mf_load(const uint32_t *__restrict__ base, int W, int X, int y0, int z0, int w0,
                                        int t, MfmaStage &st)
{
    const int w = t % MF_KW;                 // every (y, w) position of this thread has the same w
    const bool in = w0 + w < W;
    const uint32_t *xr = base + (size_t)X * 2 * W + w0 + w;
    st.x0 = in ? xr[0] : 0u;
    st.x1 = in ? xr[W] : 0u;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int y = t / MF_KW + 4 * q;     // MF_TY * MF_KW = 2048 positions
        const uint32_t *yr = base + (size_t)(y0 + y) * 2 * W + w0 + w;
        st.y0v[q] = in ? yr[0] : 0u;
        st.y1v[q] = in ? yr[W] : 0u;
    }
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int e = t + 256 * q;           // 2 * MF_TZ * MF_KW / 4 = 2048 uint4
        const int row = e / (MF_KW / 4), c4 = e % (MF_KW / 4);
        st.z[q] = w0 + c4 * 4 < W
                      ? *(const uint4 *)(base + ((size_t)(z0 + (row >> 1)) * 2 + (row & 1)) * W + w0 + c4 * 4)
                      : make_uint4(0u, 0u, 0u, 0u);
    }
}

// P = X & Y and the Z rows into this class's LDS rows: [4y+ij] then [128+2z+k].
__device__ __forceinline__ void mf_store(const MfmaStage &st, int t, uint32_t (*rows)[MF_LDSW])
{
    const int w = t % MF_KW;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int y = t / MF_KW + 4 * q;
        rows[y * 4 + 0][w] = st.x0 & st.y0v[q];
        rows[y * 4 + 1][w] = st.x0 & st.y1v[q];
        rows[y * 4 + 2][w] = st.x1 & st.y0v[q];
        rows[y * 4 + 3][w] = st.x1 & st.y1v[q];
    }
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int e = t + 256 * q;
        const int row = e / (MF_KW / 4), c4 = e % (MF_KW / 4);   // row = 2z + k
        *(uint4 *)&rows[MF_TY * 4 + row][c4 * 4] = st.z[q];
    }
}

// FP4 operand of class c from four packed words.
template <int C>
__device__ __forceinline__ v8i_t mf_operand(const uint4 &v)
{
    v8i_t o = {0, 0, 0, 0, 0, 0, 0, 0};
    if constexpr (C < 3) {
        constexpr uint32_t m = 0x11111111u << C;
        o[0] = (int)(v.x & m);
        o[1] = (int)(v.y & m);
        o[2] = (int)(v.z & m);
        o[3] = (int)(v.w & m);
    } else {
        o[0] = (int)((v.x >> 1) & 0x44444444u);
        o[1] = (int)((v.y >> 1) & 0x44444444u);
        o[2] = (int)((v.z >> 1) & 0x44444444u);
        o[3] = (int)((v.w >> 1) & 0x44444444u);
    }
    return o;
}

// E8M0 scale that turns the class-c value into 1 on both operands.
template <int C>
constexpr int mf_scale()
{
    return C == 0 ? 128 : (C == 1 ? 127 : 126);
}

template <int C>
__device__ __forceinline__ void mf_step(const uint4 (&a)[2], const uint4 (&b)[2], v16f_t (&acc)[2][2])
{
    const v8i_t A0 = mf_operand<C>(a[0]), A1 = mf_operand<C>(a[1]);
    const v8i_t B0 = mf_operand<C>(b[0]), B1 = mf_operand<C>(b[1]);
    constexpr int s = mf_scale<C>();
    acc[0][0] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(A0, B0, acc[0][0], 4, 4, 0, s, 0, s);
    acc[0][1] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(A0, B1, acc[0][1], 4, 4, 0, s, 0, s);
    acc[1][0] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(A1, B0, acc[1][0], 4, 4, 0, s, 0, s);
    acc[1][1] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(A1, B1, acc[1][1], 4, 4, 0, s, 0, s);
}

// A batch of independent 3-way searches sharing one work queue: units
// [job_start[j], job_start[j+1]) belong to job j.  A single search is a batch
// of one.
struct BatchTri {
    const TriArgs *jobs;                    // device array, one per job
    const unsigned long long *job_start;    // [njobs + 1]
    int njobs;
    unsigned long long total;
    unsigned long long *queue;
    unsigned sh = 1;                        // class-3 operand shift (see m2_aop)
    // MODE_COUNTS (k_triplets_mfma2, the plain path on the matrix cores):
    // `rounds` rounds (X0 + r, Y block yb, Z block zb of 256) per launch,
    // 16 workgroups each; round r writes its AND-counts to counts + r * slot
    // as [class][256 Y][256 Z][8] (the layout k_objective_plain reads)
    int X0 = 0, yb = 0, zb = 0, rounds = 0;
    uint32_t *counts = nullptr;
    size_t slot = 0;
};

template <int MODE>
__global__ __launch_bounds__(512) void k_triplets_mfma(BatchTri bt)
{
    __shared__ __attribute__((aligned(16))) uint32_t smem[MF_SMEM_WORDS];
    __shared__ unsigned long long s_u;
    __shared__ float s_red[16];
    __shared__ float s_lnf[scores_sets(MODE) ? LNF_SMALL : 1];
    if constexpr (scores_sets(MODE)) load_lnf(s_lnf);
    const int tid = threadIdx.x;
    const int cls = tid >> 8, t = tid & 255;            // class half and thread within it
    const int wih = t >> 6, lane = t & 63;              // wave within the half
    const int wy = wih >> 1, wz = wih & 1, h = lane >> 5, i = lane & 31;
    uint32_t (*rows)[MF_LDSW] = (uint32_t (*)[MF_LDSW])(smem + cls * MF_ROWS * MF_LDSW);
    const int arow0 = (wy * 16) * 4 + i, arow1 = (wy * 16 + 8) * 4 + i;
    const int brow0 = MF_TY * 4 + (wz * 32) * 2 + i, brow1 = MF_TY * 4 + (wz * 32 + 16) * 2 + i;

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
        // a last chunk may hold half its words (W is a multiple of MF_KW / 2)
        const int nch_mine = (W + MF_KW - 1) / MF_KW;
        const int nch = ((p.W0 > p.W1 ? p.W0 : p.W1) + MF_KW - 1) / MF_KW;
        g = next_group(p.gstart, p.ngroups, u, g);
        const int X = (int)(u - p.gstart[g]);
        const int2 gt = p.groups[g];
        const int y0 = gt.x * MF_TY, z0 = gt.y * MF_TZ;

        v16f_t acc[2][2];
#pragma unroll
        for (int rb = 0; rb < 2; rb++)
#pragma unroll
            for (int cb = 0; cb < 2; cb++)
#pragma unroll
                for (int r = 0; r < 16; r++) acc[rb][cb][r] = 0.f;

        MfmaStage st;
        if (nch_mine > 0) mf_load(base, W, X, y0, z0, 0, t, st);
        for (int c = 0; c < nch; c++) {
            __syncthreads();
            if (c < nch_mine) mf_store(st, t, rows);
            __syncthreads();
            if (c < nch_mine) {
                if (c + 1 < nch_mine) mf_load(base, W, X, y0, z0, (c + 1) * MF_KW, t, st);
                // operands of superstep s+1 are read while superstep s runs
                uint4 a[2], b[2];
                a[0] = *(const uint4 *)&rows[arow0][h * 4];
                a[1] = *(const uint4 *)&rows[arow1][h * 4];
                b[0] = *(const uint4 *)&rows[brow0][h * 4];
                b[1] = *(const uint4 *)&rows[brow1][h * 4];
#pragma unroll
                for (int s = 0; s < MF_KW / 8; s++) {
                    uint4 an[2], bn[2];
                    if (s + 1 < MF_KW / 8) {
                        const int w = (s + 1) * 8 + h * 4;
                        an[0] = *(const uint4 *)&rows[arow0][w];
                        an[1] = *(const uint4 *)&rows[arow1][w];
                        bn[0] = *(const uint4 *)&rows[brow0][w];
                        bn[1] = *(const uint4 *)&rows[brow1][w];
                    }
                    mf_step<0>(a, b, acc);
                    mf_step<1>(a, b, acc);
                    mf_step<2>(a, b, acc);
                    mf_step<3>(a, b, acc);
                    if (s + 1 < MF_KW / 8) {
                        a[0] = an[0];
                        a[1] = an[1];
                        b[0] = bn[0];
                        b[1] = bn[1];
                    }
                }
            }
        }

        // accumulators of both classes -> LDS: [cls][wave][tile][reg][lane]
        __syncthreads();
#pragma unroll
        for (int rb = 0; rb < 2; rb++)
#pragma unroll
            for (int cb = 0; cb < 2; cb++)
#pragma unroll
                for (int r = 0; r < 16; r++)
                    smem[((((cls * 4 + wih) * 4 + rb * 2 + cb) * 16) + r) * 64 + lane] = (uint32_t)acc[rb][cb][r];
        __syncthreads();

        const Singles sx = ld_global(p.singles, X);
        float bad[4];
        unsigned long long packed[4];
        unsigned long long hsum = 0;
        float skip_above = INFINITY;     // see skip_threshold
        if constexpr (scores_sets(MODE)) skip_above = skip_threshold(p.sc, p.co.best_key);
#pragma unroll
        for (int q4 = 0; q4 < 4; q4++) {
            const int T = q4 * 512 + tid;
            const int yl = T >> 6, zl = T & 63;
            const int Y = y0 + yl, Z = z0 + zl;
            bad[q4] = INFINITY;
            packed[q4] = (unsigned long long)X | ((unsigned long long)Y << 21) | ((unsigned long long)Z << 42);
            if (!(X < Y && Y < Z && Z < p.n)) continue;
            // where MFMA tile output put the counts of (yl, zl)
            const int twy = yl >> 4, trb = (yl >> 3) & 1, y8 = yl & 7, tq = y8 >> 1, th = y8 & 1;
            const int twz = zl >> 5, tcb = (zl >> 4) & 1, z16 = zl & 15;
            const int tw = twy * 2 + twz, tt = trb * 2 + tcb;
            uint32_t n8c[8], n8t[8];
#pragma unroll
            for (int ij = 0; ij < 4; ij++)
#pragma unroll
                for (int kk = 0; kk < 2; kk++) {
                    const int off = (tt * 16 + 4 * tq + ij) * 64 + th * 32 + 2 * z16 + kk;
                    n8c[ij * 2 + kk] = smem[((0 * 4 + tw) * 4) * 16 * 64 + off];
                    n8t[ij * 2 + kk] = smem[((1 * 4 + tw) * 4) * 16 * 64 + off];
                }
            const Singles sy = ld_global(p.singles, Y);
            uint32_t cas[27], ctl[27];
            triplet_tables(p, X, Y, Z, sx, sy, n8c, n8t, cas, ctl);
            if constexpr (scores_sets(MODE)) {
                if (skip_above < INFINITY && bound_exceeds<27>(p.sc, cas, ctl, skip_above))
                    continue;                           // bad[q4] stays INFINITY
                bad[q4] = badness_f32(p.sc, cas, ctl, s_lnf);
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
        if constexpr (MODE == MODE_SEARCH) publish_candidates<4, 8>(p.co, p.sc.margin, bad, packed, s_red);
        if constexpr (MODE == MODE_TOPK) publish_topk<4, 8>(p.co, p.sc.margin, bad, packed);
        if constexpr (MODE == MODE_CHECKSUM) {
            if (hsum) atomicAdd(p.checksum, hsum);
        }
    }
}


// ============================================================ version 2 ==
//
// Same contraction and FP4 encoding, restructured so the chunk pipeline costs
// no VALU work and one barrier:
//  * the X, Y and Z rows of a 1024-sample chunk go straight from global
//    memory to LDS (buffer_load_dwordx4 ... lds), into one of two buffers while
//    the other is contracted;
//  * LDS rows are 128 bytes without padding; the eight 16-byte slots of row r
//    are XOR-swizzled by r & 7, so the lanes of a wave, which read up to eight
//    consecutive rows at the same logical slot, hit eight different banks;
//  * P_ij = X_i & Y_j is formed in the operand itself (one v_bitop3 per
//    dword together with the class mask);
//  * a work unit is X x 64 Y x 64 Z; wave (class, wy) owns 16 Y x 64 Z =
//    2 x 4 MFMA tiles (128 accumulators), so each A operand feeds four MFMAs
//    and each B operand two.
// After the last chunk the accumulators go to LDS in two halves (Z 0-31,
// then Z 32-63) and all 512 threads score four triplets per half.

constexpr int M2_TY = 64, M2_TZ = 64;
constexpr int M2_KW = 32;                                   // words per chunk
constexpr int M2_ROWS = 2 + 2 * M2_TY + 2 * M2_TZ;          // X, Y, Z rows of a class
constexpr int M2_ROWS_PAD = (M2_ROWS + 7) / 8 * 8;
constexpr int M2_LOADS = M2_ROWS_PAD / 8;                   // 8 rows per load instruction
constexpr int M2_BUF_WORDS = M2_ROWS_PAD * M2_KW;
constexpr int M2_STAGE_WORDS = 4 * M2_BUF_WORDS;            // 2 classes x 2 buffers
constexpr int M2_DUMP_WORDS = 2 * 4 * 4 * 16 * 64;          // class x wave x tile x reg x lane
constexpr int M2_SMEM_WORDS = M2_STAGE_WORDS > M2_DUMP_WORDS ? M2_STAGE_WORDS : M2_DUMP_WORDS;
constexpr int M2_YROW = 2, M2_ZROW = 2 + 2 * M2_TY;


// Chunk loads, straight from global memory to LDS (buffer_load ... lds)
// through a scalar buffer descriptor: the row offsets of a work unit are
// computed once (m2_row_offsets) and every chunk only adds its word offset
// as the scalar soffset, so a chunk's loads cost no vector address
// arithmetic.  One instruction fills eight 128-byte LDS rows: lane L writes
// physical slot L%8 of row L/8, which holds logical slot (L%8) ^ (L/8).
// Reads at or past num_records (the class's plane bytes) return zero.
typedef int i32x4_t __attribute__((ext_vector_type(4)));

// LDS byte address of a pointer into a __shared__ array: the low 32 bits of
// its generic address (the generic LDS aperture keeps the LDS offset there).
// Taking them directly, rather than casting to an LDS pointer, spares the
// null check of that cast, so the chunk loads compute their destinations in
// 32-bit registers instead of keeping 64-bit pointers live.
__device__ __forceinline__ unsigned lds_addr(const void *p) { return (unsigned)(size_t)p; }

// One 16-byte-per-lane load from the buffer into LDS: lane L's data lands at
// LDS byte address lds + 16 L.  Written as inline assembly on purpose: the
// compiler cannot tell that these writes go to the other buffer of the
// double-buffered stage, and for its own LDS-DMA intrinsic it waits for
// them to finish before the next LDS read, which serialises every chunk's
// loads with its contraction.  The kernels wait for these loads themselves
// (s_waitcnt vmcnt before the barrier that ends each chunk).  m0 is written
// without being named as clobbered (the compiler treats m0 as reserved and
// does not honour it in a clobber list); that is sound because no compiler-generated code in
// these kernels uses m0, which tests/test_device_asm.py checks on the
// assembly the Makefile's flags produce.
__device__ __forceinline__ void lds_dma16(i32x4_t rsrc, unsigned lds, unsigned voffset, unsigned soffset)
{
    const unsigned m0 = (unsigned)__builtin_amdgcn_readfirstlane((int)lds);
    asm volatile("s_mov_b32 m0, %0\n\tbuffer_load_dwordx4 %1, %2, %3 offen lds"
                 :
                 : "s"(m0), "v"(voffset), "s"(rsrc), "s"(soffset)
                 : "memory");
}

constexpr int M2_LOADS_PER_WAVE = (M2_LOADS + 3) / 4;

__device__ __forceinline__ i32x4_t m2_rsrc(const uint32_t *base, unsigned bytes)
{
    // base and bytes are the same for a whole wave (one class per wave);
    // readfirstlane puts the descriptor in scalar registers
    const unsigned long long a = (unsigned long long)base;
    i32x4_t r;
    r.x = __builtin_amdgcn_readfirstlane((int)(unsigned)a);
    r.y = __builtin_amdgcn_readfirstlane((int)(unsigned)(a >> 32));   // stride 0: raw byte-addressed
    r.z = __builtin_amdgcn_readfirstlane((int)bytes);
    r.w = 0x00020000;                                                  // 32-bit data format
    return r;
}

// Byte offset, within the class's planes, of slot q of each LDS row this
// lane loads (load k of wave wih fills rows 8(wih + 4k) .. +7).
__device__ __forceinline__ void m2_row_offsets(int W, int X, int y0, int z0, int wih, int lane,
                                               unsigned (&roff)[M2_LOADS_PER_WAVE])
{
    const int rl = lane >> 3, q = (lane & 7) ^ rl;
#pragma unroll
    for (int k = 0; k < M2_LOADS_PER_WAVE; k++) {
        const int r = (wih + 4 * k) * 8 + rl;
        int row;
        if (r < M2_YROW) row = 2 * X + r;
        else if (r < M2_ZROW) row = 2 * y0 + (r - M2_YROW);
        else if (r < M2_ROWS) row = 2 * z0 + (r - M2_ZROW);
        else row = 2 * X;                     // padding rows are never read
        roff[k] = ((unsigned)row * (unsigned)W + (unsigned)(q * 4)) * 4u;
    }
}

__device__ __forceinline__ void m2_issue_buf(i32x4_t rsrc, const unsigned (&roff)[M2_LOADS_PER_WAVE], int w0,
                                             unsigned buf, int wih)
{
#pragma unroll
    for (int k = 0; k < M2_LOADS_PER_WAVE; k++) {
        const int ii = wih + 4 * k;
        if (ii < M2_LOADS) lds_dma16(rsrc, buf + (unsigned)(ii * 8 * M2_KW * 4), roff[k], (unsigned)(w0 * 4));
    }
}

// Word offset of logical 16-byte slot q of LDS row r.
__device__ __forceinline__ int m2_at(int r, int q) { return r * M2_KW + ((q ^ (r & 7)) << 2); }

// A operand X & Y of class C.  Class 3 sits in the sign bit of each FP4
// nibble and moves to bit 2: (x & y & 0x88888888) >> 1, one v_bitop3_b32 and
// one shift.  The shift count sh (always 1) comes from the kernel arguments:
// with a constant count the compiler rewrites the expression as
// ((x & y) >> 1) & 0x44444444, three instructions.
template <int C>
__device__ __forceinline__ v8i_t m2_aop(const uint4 &x, const uint4 &y, unsigned sh)
{
    v8i_t o = {0, 0, 0, 0, 0, 0, 0, 0};
    if constexpr (C < 3) {
        constexpr uint32_t m = 0x11111111u << C;
        o[0] = (int)(x.x & y.x & m);
        o[1] = (int)(x.y & y.y & m);
        o[2] = (int)(x.z & y.z & m);
        o[3] = (int)(x.w & y.w & m);
    } else {
        o[0] = (int)((x.x & y.x & 0x88888888u) >> sh);
        o[1] = (int)((x.y & y.y & 0x88888888u) >> sh);
        o[2] = (int)((x.z & y.z & 0x88888888u) >> sh);
        o[3] = (int)((x.w & y.w & 0x88888888u) >> sh);
    }
    return o;
}

template <int C>
__device__ __forceinline__ void m2_step(const uint4 &xw, const uint4 (&yw)[2], const uint4 (&zw)[4],
                                        v16f_t (&acc)[2][4], unsigned sh)
{
    constexpr int s = mf_scale<C>();
    const v8i_t A0 = m2_aop<C>(xw, yw[0], sh), A1 = m2_aop<C>(xw, yw[1], sh);
#pragma unroll
    for (int cb = 0; cb < 4; cb++) {
        const v8i_t B = mf_operand<C>(zw[cb]);
        acc[0][cb] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(A0, B, acc[0][cb], 4, 4, 0, s, 0, s);
        acc[1][cb] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(A1, B, acc[1][cb], 4, 4, 0, s, 0, s);
    }
}


template <int MODE>
__global__ __launch_bounds__(512) void k_triplets_mfma2(BatchTri bt)
{
    __shared__ __attribute__((aligned(16))) uint32_t smem[M2_SMEM_WORDS];
    __shared__ unsigned long long s_u;
    __shared__ float s_red[16];
    __shared__ float s_lnf[scores_sets(MODE) ? LNF_SMALL : 1];
    if constexpr (scores_sets(MODE)) load_lnf(s_lnf);
    const int tid = threadIdx.x;
    const int cls = tid >> 8, t = tid & 255;
    const int wy = t >> 6, lane = t & 63, h = lane >> 5, i = lane & 31;
    uint32_t *buf0 = smem + (cls * 2 + 0) * M2_BUF_WORDS;
    uint32_t *buf1 = smem + (cls * 2 + 1) * M2_BUF_WORDS;
    // LDS rows read by this lane: X plane, Y rows of its two row blocks, Z
    // rows of its four column blocks
    const int xr = (i >> 1) & 1;
    const int yr0 = M2_YROW + 2 * (wy * 16 + (i >> 2)) + (i & 1);
    const int yr1 = M2_YROW + 2 * (wy * 16 + 8 + (i >> 2)) + (i & 1);
    int zr[4];
#pragma unroll
    for (int cb = 0; cb < 4; cb++) zr[cb] = M2_ZROW + 2 * (cb * 16 + (i >> 1)) + (i & 1);

    int g = 0, j = 0;
    for (int iter = 0;; iter++) {
        int X, y0, z0;
        if constexpr (MODE == MODE_COUNTS) {
            // one 64 x 64 tile of round blockIdx.x / 16, no queue
            if (iter > 0) break;
            const int r = (int)blockIdx.x >> 4, tl = (int)blockIdx.x & 15;
            if (r >= bt.rounds) break;
            X = bt.X0 + r;
            y0 = bt.yb * 256 + (tl >> 2) * M2_TY;
            z0 = bt.zb * 256 + (tl & 3) * M2_TZ;
        } else {
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
            const unsigned long long u = uu - bt.job_start[j];
            const TriArgs &pj = bt.jobs[j];
            g = next_group(pj.gstart, pj.ngroups, u, g);
            X = (int)(u - pj.gstart[g]);
            const int2 gt = pj.groups[g];
            y0 = gt.x * M2_TY;
            z0 = gt.y * M2_TZ;
        }
        const TriArgs &p = bt.jobs[j];
        const uint32_t *base = cls ? p.p1 : p.p0;
        const int W = cls ? p.W1 : p.W0;
        const int nch_mine = W / M2_KW;
        const int nch = (p.W0 > p.W1 ? p.W0 : p.W1) / M2_KW;
        const i32x4_t rsrc = m2_rsrc(base, cls ? p.bytes1 : p.bytes0);
        unsigned roff[M2_LOADS_PER_WAVE];
        m2_row_offsets(W, X, y0, z0, wy, lane, roff);

        v16f_t acc[2][4];
#pragma unroll
        for (int rb = 0; rb < 2; rb++)
#pragma unroll
            for (int cb = 0; cb < 4; cb++)
#pragma unroll
                for (int r = 0; r < 16; r++) acc[rb][cb][r] = 0.f;

        if (nch_mine > 0) m2_issue_buf(rsrc, roff, 0, lds_addr(buf0), wy);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        for (int c = 0; c < nch; c++) {
            uint32_t *cur = (c & 1) ? buf1 : buf0;
            if (c + 1 < nch_mine) m2_issue_buf(rsrc, roff, (c + 1) * M2_KW, lds_addr((c & 1) ? buf0 : buf1), wy);
            if (c < nch_mine) {
#pragma unroll
                for (int s = 0; s < M2_KW / 8; s++) {
                    const int q = 2 * s + h;
                    const uint4 xw = *(const uint4 *)&cur[m2_at(xr, q)];
                    uint4 yw[2], zw[4];
                    yw[0] = *(const uint4 *)&cur[m2_at(yr0, q)];
                    yw[1] = *(const uint4 *)&cur[m2_at(yr1, q)];
#pragma unroll
                    for (int cb = 0; cb < 4; cb++) zw[cb] = *(const uint4 *)&cur[m2_at(zr[cb], q)];
                    m2_step<0>(xw, yw, zw, acc, bt.sh);
                    m2_step<1>(xw, yw, zw, acc, bt.sh);
                    m2_step<2>(xw, yw, zw, acc, bt.sh);
                    m2_step<3>(xw, yw, zw, acc, bt.sh);
                }
            }
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
        }

        const Singles sx = MODE == MODE_COUNTS ? Singles{0, 0, 0, 0} : ld_global(p.singles, X);
        float bad[8];
        unsigned long long packed[8];
        unsigned long long hsum = 0;
        float skip_above = INFINITY;     // see skip_threshold
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
            for (int q4 = 0; q4 < 4; q4++) {
                const int m = hf * 4 + q4;
                const int T = q4 * 512 + tid;                   // 64 Y x 32 Z of this half
                const int yl = T >> 5, zl = hf * 32 + (T & 31);
                const int Y = y0 + yl, Z = z0 + zl;
                bad[m] = INFINITY;
                packed[m] = (unsigned long long)X | ((unsigned long long)Y << 21) | ((unsigned long long)Z << 42);
                // the plain path's objective kernel checks the set itself
                if (MODE != MODE_COUNTS && !(X < Y && Y < Z && Z < p.n)) continue;
                const int twy = yl >> 4, trb = (yl >> 3) & 1, y8 = yl & 7, tq = y8 >> 1, th = y8 & 1;
                const int tcr = (T & 31) >> 4, z16 = T & 15;
                const int t4 = trb * 2 + tcr;
                uint32_t n8c[8], n8t[8];
#pragma unroll
                for (int ij = 0; ij < 4; ij++)
#pragma unroll
                    for (int kk = 0; kk < 2; kk++) {
                        const int off = ((t4 * 16) + 4 * tq + ij) * 64 + th * 32 + 2 * z16 + kk;
                        n8c[ij * 2 + kk] = smem[(0 * 4 + twy) * 4 * 16 * 64 + off];
                        n8t[ij * 2 + kk] = smem[(1 * 4 + twy) * 4 * 16 * 64 + off];
                    }
                if constexpr (MODE == MODE_COUNTS) {
                    const int yr = Y - bt.yb * 256, zr = Z - bt.zb * 256;
                    uint32_t *rc = bt.counts + (size_t)(X - bt.X0) * bt.slot;
                    uint4 *o0 = (uint4 *)(rc + ((size_t)(0 * 256 + yr) * 256 + zr) * 8);
                    uint4 *o1 = (uint4 *)(rc + ((size_t)(1 * 256 + yr) * 256 + zr) * 8);
                    o0[0] = make_uint4(n8c[0], n8c[1], n8c[2], n8c[3]);
                    o0[1] = make_uint4(n8c[4], n8c[5], n8c[6], n8c[7]);
                    o1[0] = make_uint4(n8t[0], n8t[1], n8t[2], n8t[3]);
                    o1[1] = make_uint4(n8t[4], n8t[5], n8t[6], n8t[7]);
                    continue;
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
        if constexpr (MODE == MODE_SEARCH) publish_candidates<8, 8>(p.co, p.sc.margin, bad, packed, s_red);
        if constexpr (MODE == MODE_TOPK) publish_topk<8, 8>(p.co, p.sc.margin, bad, packed);
        if constexpr (MODE == MODE_CHECKSUM) {
            if (hsum) atomicAdd(p.checksum, hsum);
        }
    }
}

}  // namespace episdet
