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
// kernel's (2,4) tile).  Waves 0-3 contract the cases, waves 4-7 the
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
__device__ __forceinline__ void mf_load(const uint32_t *__restrict__ base, int W, int X, int y0, int z0, int w0,
                                        int t, MfmaStage &st)
{
    const int w = t % MF_KW;                 // every (y, w) position of this thread has the same w
    const uint32_t *xr = base + (size_t)X * 2 * W + w0 + w;
    st.x0 = xr[0];
    st.x1 = xr[W];
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int y = t / MF_KW + 4 * q;     // MF_TY * MF_KW = 2048 positions
        const uint32_t *yr = base + (size_t)(y0 + y) * 2 * W + w0 + w;
        st.y0v[q] = yr[0];
        st.y1v[q] = yr[W];
    }
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int e = t + 256 * q;           // 2 * MF_TZ * MF_KW / 4 = 2048 uint4
        const int row = e / (MF_KW / 4), c4 = e % (MF_KW / 4);
        st.z[q] = *(const uint4 *)(base + ((size_t)(z0 + (row >> 1)) * 2 + (row & 1)) * W + w0 + c4 * 4);
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

template <int MODE>
__global__ __launch_bounds__(512) void k_triplets_mfma(TriArgs p)
{
    __shared__ __attribute__((aligned(16))) uint32_t smem[MF_SMEM_WORDS];
    __shared__ unsigned long long s_u;
    __shared__ float s_red[16];
    const int tid = threadIdx.x;
    const int cls = tid >> 8, t = tid & 255;            // class half and thread within it
    const int wih = t >> 6, lane = t & 63;              // wave within the half
    const int wy = wih >> 1, wz = wih & 1, h = lane >> 5, i = lane & 31;
    uint32_t (*rows)[MF_LDSW] = (uint32_t (*)[MF_LDSW])(smem + cls * MF_ROWS * MF_LDSW);
    const uint32_t *base = cls ? p.p1 : p.p0;
    const int W = cls ? p.W1 : p.W0;
    const int nch_mine = W / MF_KW;
    const int nch = (p.W0 > p.W1 ? p.W0 : p.W1) / MF_KW;
    const int arow0 = (wy * 16) * 4 + i, arow1 = (wy * 16 + 8) * 4 + i;
    const int brow0 = MF_TY * 4 + (wz * 32) * 2 + i, brow1 = MF_TY * 4 + (wz * 32 + 16) * 2 + i;

    for (;;) {
        if (tid == 0) s_u = atomicAdd(p.queue, 1ull);
        __syncthreads();
        const unsigned long long u = s_u;
        if (u >= p.nunits) break;
        const int g = find_group(p.gstart, p.ngroups, u);
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
#pragma unroll
                for (int s = 0; s < MF_KW / 8; s++) {
                    const int w = s * 8 + h * 4;
                    uint4 a[2], b[2];
                    a[0] = *(const uint4 *)&rows[arow0][w];
                    a[1] = *(const uint4 *)&rows[arow1][w];
                    b[0] = *(const uint4 *)&rows[brow0][w];
                    b[1] = *(const uint4 *)&rows[brow1][w];
                    mf_step<0>(a, b, acc);
                    mf_step<1>(a, b, acc);
                    mf_step<2>(a, b, acc);
                    mf_step<3>(a, b, acc);
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

        const Singles sx = p.singles[X];
        float bad[4];
        unsigned long long packed[4];
        unsigned long long hsum = 0;
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
            const Singles sy = p.singles[Y];
            uint32_t cas[27], ctl[27];
            triplet_tables(p, X, Y, Z, sx, sy, n8c, n8t, cas, ctl);
            if constexpr (MODE == MODE_SEARCH) {
                bad[q4] = badness_f32(p.sc, cas, ctl);
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
        if constexpr (MODE == MODE_CHECKSUM) {
            if (hsum) atomicAdd(p.checksum, hsum);
        }
    }
}

}  // namespace episdet
