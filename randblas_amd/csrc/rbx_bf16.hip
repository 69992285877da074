// rbx_bf16.hip -- BF16-split form of the left dense sketch on the gfx950
// BF16 matrix cores (no RandBLAS counterpart; the FP32 analogue is lskge3,
// skge.hh:154-202).
//
//   B = alpha * op(bf16(submat(S))) * op(A) + beta * B
//
// Three kernels:
//   1. split:  op(A) (FP32, any strides) -> T bfloat16 planes A_1..A_T with
//      A = A_1 + ... + A_T + residual.  Every term is the round-to-nearest-even
//      of the previous remainder, and each remainder is exact in FP32.  The
//      planes are stored fragment-packed: for a 32-column block and a 16-deep
//      K step, the 64 lanes' MFMA B operands (8 bf16 each) are one contiguous
//      1 KB run, so every load instruction of the main kernel is a full
//      coalesced 1 KB read.  Zero padding to Kp (multiple of 32) x Np
//      (multiple of 256).
//   2. main:   one 256-thread workgroup per (64-row block, 1024-column block,
//      K-split).  Each 32-wide K-chunk of the operator is generated from its
//      Philox counters, rounded to bfloat16 and double-buffered in LDS; the
//      four waves each own 256 columns and keep a 64 x 256 FP32 accumulator
//      (2 x 8 blocks of v_mfma_f32_32x32x16_bf16, 256 AGPRs).  Each operator
//      entry is generated once per 1024 columns.  B operands rotate through
//      registers: the load of block j for the next K step is issued right
//      after block j's MFMAs of this step, so loads are in flight for a whole
//      step.  The K-split bounds each FP32 accumulation chain; partial tiles
//      go to a row-major scratch.
//   3. reduce: B = alpha * sum_s partial_s (in FP64, fixed order) + beta * B.
// WG-to-tile mapping is XCD-aware: workgroups that share a K-range are
// dispatched to the same XCD back to back, so the A planes they stream are
// served from that XCD's L2.
#include "rbx_internal.hh"
#include "rbx_gemm_common.hh"


namespace rbx {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int BF_MT = 64, BF_NT = 1024, BF_KC = 32, BF_LDK = 40, BF_NTHR = 256;

__host__ __device__ __forceinline__ uint16_t bf16_rne(float x) {
#ifdef __HIP_DEVICE_COMPILE__
    uint32_t u = __float_as_uint(x);
#else
    uint32_t u;
    memcpy(&u, &x, 4);
#endif
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Operator entry rounded to bfloat16 (the value the FP32 operator would hold,
// then RNE).
template <>
__device__ __forceinline__ uint16_t entry_value<uint16_t>(float v, bool gauss) {
    return bf16_rne(gauss ? v : __fmul_rn(v, 0x1.bb67aep+0f));
}

// ---------------------------------------------------------------- split
// Element (k, j) of plane t lives at
//   t*plane_stride + (((j/32)*(Kp/16) + k/16)*64 + j%32 + 32*((k/8)%2))*8 + k%8.
// One wave per (32-column block cb, 16-deep K step s).  Lane l handles column
// 32*cb + l%32 and k = 16*s + 8*(l/32) .. +7, and writes lane l's 16-byte slot
// of the block's 1 KB run in every plane, so the wave's stores are one
// contiguous 1 KB per plane.
__global__ __launch_bounds__(256) void split_kernel(const float *A, int64_t rs, int64_t cs, int64_t K, int64_t N,
                                                    int64_t Kp, int64_t Np, int terms, uint16_t *planes,
                                                    int64_t plane_stride) {
    const int64_t KS = Kp / 16;
    const int64_t n_waves = (Np / 32) * KS;
    const int lane = threadIdx.x & 63;
    const int64_t wave0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t wstride = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t w = wave0; w < n_waves; w += wstride) {
        const int64_t cb = w / KS, st = w - cb * KS;
        const int64_t j = cb * 32 + (lane & 31);
        const int64_t k0 = st * 16 + 8 * (lane >> 5);
        uint16_t out[3][8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int64_t k = k0 + e;
            float a = (k < K && j < N) ? A[k * rs + j * cs] : 0.0f;
#pragma unroll
            for (int t = 0; t < 3; ++t) {
                uint16_t h = bf16_rne(a);
                out[t][e] = h;
                a = __fsub_rn(a, bf16_to_f32(h));  // exact
            }
        }
        const int64_t dst = ((cb * KS + st) * 64 + lane) * 8;
        for (int t = 0; t < terms; ++t) {
            u32x4 v;
            v.x = (uint32_t)out[t][0] | ((uint32_t)out[t][1] << 16);
            v.y = (uint32_t)out[t][2] | ((uint32_t)out[t][3] << 16);
            v.z = (uint32_t)out[t][4] | ((uint32_t)out[t][5] << 16);
            v.w = (uint32_t)out[t][6] | ((uint32_t)out[t][7] << 16);
            *(u32x4 *)(planes + t * plane_stride + dst) = v;
        }
    }
}

struct Bf16Problem {
    int64_t M, N, K, Kp, Np;
    GenMap g;
    const uint16_t *planes;
    int64_t plane_stride;
    float *part;  // nS x M x N, row-major
    int64_t kw;   // K per split (multiple of BF_KC)
    int nMB, nNB, nS;
};

// 32x32 output block, element r of lane l: row 8*(r/4) + 4*(l/32) + r%4, column l%32.
__device__ __forceinline__ int out_row32(int lane, int r) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }

// Operator chunk [m0, m0+64) x [k0, k0+32) -> LDS (row-major, BF_LDK).  When
// the chunk is inside the operator and its major axis runs along K from a
// counter boundary, each Philox counter fills four consecutive entries of a
// row with one 8-byte store.
__device__ __forceinline__ void gen_chunk(const GenMap &g, int64_t m0, int64_t k0, int64_t M, int64_t K,
                                          uint16_t *lds, int tid) {
    const bool fast = !g.b_along_rows && ((g.b_off + k0) & 3) == 0 && m0 + BF_MT <= M && k0 + BF_KC <= K;
    if (!fast) {
        gen_tile<uint16_t, BF_MT, BF_KC, BF_NTHR>(g, m0, k0, M, K, lds, BF_LDK, 1, tid);
        return;
    }
    const bool gauss = g.gauss != 0;
    const int64_t cell0 = (g.b_off + k0) >> 2;
#pragma unroll
    for (int it = 0; it < (BF_MT * BF_KC / 4) / BF_NTHR; ++it) {
        const int idx = tid + it * BF_NTHR;
        const int row = idx >> 3, cell = idx & 7;
        const Ctr128 c = ctr_add(g.ctr, (uint64_t)((g.a_off + m0 + row) * g.P + cell0 + cell));
        const float4 v = gen4(gauss, c, g.k0, g.k1);
        uint2 pk;
        pk.x = (uint32_t)entry_value<uint16_t>(v.x, gauss) | ((uint32_t)entry_value<uint16_t>(v.y, gauss) << 16);
        pk.y = (uint32_t)entry_value<uint16_t>(v.z, gauss) | ((uint32_t)entry_value<uint16_t>(v.w, gauss) << 16);
        *(uint2 *)(lds + row * BF_LDK + cell * 4) = pk;
    }
}

template <int TERMS>
__global__ __launch_bounds__(BF_NTHR, 1) void bf16_skge_kernel(Bf16Problem pb) {
    __shared__ __attribute__((aligned(16))) uint16_t Ss[2][BF_MT * BF_LDK];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

    // XCD-aware tile order: consecutive tiles (same K-split) land on one XCD
    const int W = pb.nMB * pb.nNB * pb.nS;
    const int L = blockIdx.x;
    const int t = (W % 8 == 0) ? (L % 8) * (W / 8) + L / 8 : L;
    const int mb = t % pb.nMB;
    const int rest = t / pb.nMB;
    const int nb = rest % pb.nNB;
    const int sp = rest / pb.nNB;
    const int64_t m0 = (int64_t)mb * BF_MT;
    const int64_t n0 = (int64_t)nb * BF_NT;
    const int64_t kbeg = (int64_t)sp * pb.kw;
    int64_t kend = kbeg + pb.kw;
    if (kend > pb.Kp) kend = pb.Kp;
    const int nchunks = (int)((kend - kbeg) / BF_KC);

    f32x16 acc[2][8];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

    const int64_t wcol = n0 + wave * 256;  // first column of this wave
    const bool active = wcol < pb.Np;     // Np is a multiple of 256: all 8 blocks or none
    const int64_t KS = pb.Kp / 16;
    // lane's B operand for block j, K step s, term t:
    //   planes + t*plane_stride + (((wcol/32 + j)*KS + s)*64 + lane)*8
    const uint16_t *bbase = pb.planes + ((((wcol >> 5) * KS) * 64 + lane) << 3);
    const int64_t bstep_blk = KS * 512;  // elements between column blocks
    const int64_t s_first = kbeg / 16, s_last = kend / 16 - 1;

    bf16x8 bq[8][TERMS];
    if (active) {
#pragma unroll
        for (int j = 0; j < 8; ++j)
#pragma unroll
            for (int tt = 0; tt < TERMS; ++tt)
                bq[j][tt] = __builtin_bit_cast(
                    bf16x8, *(const u32x4 *)(bbase + j * bstep_blk + tt * pb.plane_stride + s_first * 512));
    }

    if (nchunks > 0) gen_chunk(pb.g, m0, kbeg, pb.M, pb.K, Ss[0], tid);
    __syncthreads();

    for (int c = 0; c < nchunks; ++c) {
        const int64_t kc = kbeg + (int64_t)c * BF_KC;
        if (c + 1 < nchunks) gen_chunk(pb.g, m0, kc + BF_KC, pb.M, pb.K, Ss[(c + 1) & 1], tid);
        const uint16_t *S = Ss[c & 1];
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            const int koff = ks * 16 + 8 * (lane >> 5);
            const bf16x8 a0 = __builtin_bit_cast(bf16x8, *(const u32x4 *)(S + (lane & 31) * BF_LDK + koff));
            const bf16x8 a1 = __builtin_bit_cast(bf16x8, *(const u32x4 *)(S + (32 + (lane & 31)) * BF_LDK + koff));
            int64_t s_next = kc / 16 + ks + 1;
            if (s_next > s_last) s_next = s_last;  // redundant in-range reload at the end
            if (active) {
#pragma unroll
                for (int j = 0; j < 8; ++j) {
#pragma unroll
                    for (int tt = 0; tt < TERMS; ++tt) {
                        acc[0][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, bq[j][tt], acc[0][j], 0, 0, 0);
                        acc[1][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, bq[j][tt], acc[1][j], 0, 0, 0);
                    }
#pragma unroll
                    for (int tt = 0; tt < TERMS; ++tt)
                        bq[j][tt] = __builtin_bit_cast(
                            bf16x8, *(const u32x4 *)(bbase + j * bstep_blk + tt * pb.plane_stride + s_next * 512));
                }
            }
        }
        __syncthreads();
    }

    // partial tile -> scratch (row-major M x N per split; lanes walk columns)
    if (!active) return;
    float *P = pb.part + (int64_t)sp * pb.M * pb.N;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int64_t gi = m0 + i * 32 + out_row32(lane, r);
                const int64_t gj = wcol + j * 32 + (lane & 31);
                if (gi < pb.M && gj < pb.N) P[gi * pb.N + gj] = acc[i][j][r];
            }
}

// B[i,j] = alpha * sum_s part[s][i*N + j] + beta * B[i*brs + j*bcs], sums in
// FP64 in split order.  32 x 32 tiles: partials are read along j; a
// column-major B is written along i after a transpose through LDS.
__global__ __launch_bounds__(256) void bf16_reduce_kernel(const float *part, int nS, int64_t M, int64_t N,
                                                          double alpha, double beta, float *B, int64_t brs,
                                                          int64_t bcs) {
    __shared__ double tile[32][33];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    const int64_t i0 = (int64_t)blockIdx.y * 32, j0 = (int64_t)blockIdx.x * 32;
    const int64_t total = M * N;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int64_t i = i0 + ty + 8 * q, j = j0 + tx;
        double v = 0.0;
        if (i < M && j < N)
            for (int s = 0; s < nS; ++s) v += (double)part[(int64_t)s * total + i * N + j];
        tile[ty + 8 * q][tx] = v;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        // column-major: lanes walk i; row-major: lanes walk j
        const bool cm = brs == 1;
        const int64_t i = cm ? i0 + tx : i0 + ty + 8 * q;
        const int64_t j = cm ? j0 + ty + 8 * q : j0 + tx;
        if (i < M && j < N) {
            const double v = cm ? tile[tx][ty + 8 * q] : tile[ty + 8 * q][tx];
            float *b = B + i * brs + j * bcs;
            double out = alpha * v;
            if (beta != 0.0) out += beta * (double)(*b);
            *b = (float)out;
        }
    }
}

static void bf16_split_sketch(char layout, char opS, char opA, int64_t d, int64_t n, int64_t m, float alpha,
                              const rbx_dense_skop *S, int64_t ro_s, int64_t co_s, const float *A, int64_t lda,
                              float beta, float *B, int64_t ldb, int n_terms, hipStream_t stream) {
    RBX_REQUIRE(S != nullptr);
    check_layout(layout);
    check_op(opS);
    check_op(opA);
    RBX_REQUIRE(n_terms >= 1 && n_terms <= 3);
    RBX_REQUIRE(d >= 0 && n >= 0 && m >= 0);
    RBX_REQUIRE(ro_s >= 0 && co_s >= 0);
    if (S->buff != nullptr)
        RBX_FAIL(RBX_UNSUPPORTED, "bf16split: only generated operators (S.buff == NULL) are supported");
    DenseGeom geo = dense_geom(S->dist);
    safe_mul(geo.dim_minor, geo.P);
    int64_t rows_S = opS == 'N' ? d : m, cols_S = opS == 'N' ? m : d;
    RBX_REQUIRE(S->dist.n_rows >= rows_S + ro_s);
    RBX_REQUIRE(S->dist.n_cols >= cols_S + co_s);
    int64_t rows_A = opA == 'N' ? m : n, cols_A = opA == 'N' ? n : m;
    if (layout == 'C') {
        RBX_REQUIRE(lda >= rows_A);
        RBX_REQUIRE(ldb >= d);
    } else {
        RBX_REQUIRE(lda >= cols_A);
        RBX_REQUIRE(ldb >= n);
    }
    const int64_t M = d, N = n, K = m;
    const int64_t brs = layout == 'C' ? 1 : ldb, bcs = layout == 'C' ? ldb : 1;
    if (M == 0 || N == 0) return;
    if (alpha == 0.0f || K == 0) {
        scale_output<float>(B, M, N, brs, bcs, beta, stream);
        return;
    }
    // op(A)[k, j] strides
    int64_t ars, acs;
    if (layout == 'C') {
        ars = opA == 'N' ? 1 : lda;
        acs = opA == 'N' ? lda : 1;
    } else {
        ars = opA == 'N' ? lda : 1;
        acs = opA == 'N' ? 1 : lda;
    }
    Bf16Problem pb{};
    pb.M = M;
    pb.N = N;
    pb.K = K;
    pb.Kp = (K + BF_KC - 1) / BF_KC * BF_KC;
    pb.Np = (N + 255) / 256 * 256;
    pb.g = gen_map(*S, opS, ro_s, co_s, false);
    pb.nMB = (int)((M + BF_MT - 1) / BF_MT);
    pb.nNB = (int)((N + BF_NT - 1) / BF_NT);
    // K per split: bounded FP32 accumulation chains, and enough workgroups
    int64_t kw = 4096;
    int64_t tiles = (int64_t)pb.nMB * pb.nNB;
    int64_t want_splits = (4 * (int64_t)num_cus() + tiles - 1) / tiles;
    int64_t kw_fill = (pb.Kp + want_splits - 1) / want_splits;
    if (kw_fill < kw) kw = kw_fill;
    kw = (kw + BF_KC - 1) / BF_KC * BF_KC;
    if (kw < BF_KC) kw = BF_KC;
    pb.kw = kw;
    pb.nS = (int)((pb.Kp + kw - 1) / kw);
    RBX_REQUIRE((int64_t)pb.nMB * pb.nNB * pb.nS < (1ll << 31));

    const size_t plane_elems = (size_t)safe_mul(pb.Kp, pb.Np);
    const size_t planes_bytes = ((size_t)n_terms * plane_elems * 2 + 255) / 256 * 256;
    const size_t part_bytes = (size_t)safe_mul(safe_mul(pb.nS, M), N) * 4;
    char *ws = (char *)scratch(SCRATCH_BF16, planes_bytes + part_bytes);
    pb.planes = (const uint16_t *)ws;
    pb.plane_stride = (int64_t)plane_elems;
    pb.part = (float *)(ws + planes_bytes);

    {
        int64_t total = (pb.Np / 32) * (pb.Kp / 16) * 64;  // one wave per (block, K step)
        int64_t blocks = (total + 255) / 256;
        if (blocks > 65536) blocks = 65536;
        ProfScope prof("bf16_split_kernel", stream);
        hipLaunchKernelGGL(split_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, A, ars, acs, K, N, pb.Kp, pb.Np,
                           n_terms, (uint16_t *)pb.planes, pb.plane_stride);
        RBX_HIP(hipGetLastError());
    }
    {
        dim3 grid((unsigned)(pb.nMB * pb.nNB * pb.nS));
        ProfScope prof("bf16_skge_kernel", stream);
        if (n_terms == 1) hipLaunchKernelGGL(bf16_skge_kernel<1>, grid, dim3(BF_NTHR), 0, stream, pb);
        else if (n_terms == 2) hipLaunchKernelGGL(bf16_skge_kernel<2>, grid, dim3(BF_NTHR), 0, stream, pb);
        else hipLaunchKernelGGL(bf16_skge_kernel<3>, grid, dim3(BF_NTHR), 0, stream, pb);
        RBX_HIP(hipGetLastError());
    }
    {
        dim3 grid((unsigned)((N + 31) / 32), (unsigned)((M + 31) / 32));
        RBX_REQUIRE(grid.y <= 65535u);
        ProfScope prof("bf16_reduce_kernel", stream);
        hipLaunchKernelGGL(bf16_reduce_kernel, grid, dim3(256), 0, stream, pb.part, pb.nS, M, N,
                           (double)alpha, (double)beta, B, brs, bcs);
        RBX_HIP(hipGetLastError());
    }
}

}  // namespace rbx

using namespace rbx;

extern "C" rbx_status rbx_slskge3_bf16split(char layout, char opS, char opA, int64_t d, int64_t n, int64_t m,
                                            float alpha, const rbx_dense_skop *S, int64_t ro_s, int64_t co_s,
                                            const float *A, int64_t lda, float beta, float *B, int64_t ldb,
                                            int n_terms, void *stream) {
    RBX_API_BEGIN
    bf16_split_sketch(layout, opS, opA, d, n, m, alpha, S, ro_s, co_s, A, lda, beta, B, ldb, n_terms,
                      (hipStream_t)stream);
    RBX_API_END
}
