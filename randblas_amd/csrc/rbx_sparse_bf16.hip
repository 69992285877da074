// rbx_sparse_bf16.hip -- BF16-split form of rsksp3 for CSR data (no RandBLAS
// counterpart; the FP32 analogue is rsksp3, sparse_data/sksp.hh:277-326 ->
// csr_spmm_impl.hh:51-119):
//
//   B = alpha * A_T * op(bf16(submat(S))) + beta * B
//
// A is a CSR matrix with FP32 values and A_T the T-term bfloat16 split of its
// values (a = a_1 + ... + a_T + residual, each a_t the bfloat16 rounding of
// the previous remainder; T = 3 leaves no residual for FP32 values but in
// rare cases one bit); bf16(S) is every FP32 operator entry rounded to
// bfloat16, the same operator as the dense BF16-split sketch.  The operator
// is written once per call as a row-major bfloat16 matrix G = op(bf16(S)),
// half the bytes of the FP32 operator the plain path gathers: the product
// gathers one row of G per stored entry of A, so that halves the dominant
// traffic.  Each output row is one wave's chain over the row's entries in
// stored order: FP32 fused multiply-adds, flushed into FP64 every 256
// entries (long rows), then alpha and beta applied in FP64.
#include "rbx_internal.hh"

namespace rbx {

constexpr int SB_ROWS = 32, SB_NW = 8, SB_CW = 256;  // rows per workgroup, waves, columns per chunk
constexpr int64_t SB_FLUSH = 256;                    // entries per FP32 partial sum

// Value of the TERMS-term bfloat16 split of a (each remainder is exact in
// FP32; the terms after the first add up exactly).
template <int TERMS>
__device__ __forceinline__ float split_value(float a) {
    const float h1 = bf16_to_f32(bf16_rne(a));
    if (TERMS == 1) return h1;
    const float r1 = __fsub_rn(a, h1);
    const float h2 = bf16_to_f32(bf16_rne(r1));
    if (TERMS == 2) return __fadd_rn(h1, h2);
    const float r2 = __fsub_rn(r1, h2);
    const float h3 = bf16_to_f32(bf16_rne(r2));
    return __fadd_rn(h1, __fadd_rn(h2, h3));
}

// One workgroup: SB_ROWS rows of C and SB_CW columns (4 per lane, an 8-byte
// gather of G per stored entry).  Column-major outputs go through an LDS
// tile so stores are 32 rows (128 B) long.
template <int TERMS, bool COLMAJ>
__global__ __launch_bounds__(64 * SB_NW) void csr_rows_bf16_kernel(const int64_t *rowptr, const int64_t *colidx,
                                                                   const float *vals, int64_t mrows, int64_t d,
                                                                   float alpha, float beta, const uint16_t *G,
                                                                   int64_t ldg, float *C, int64_t ldc) {
    __shared__ float tile[COLMAJ ? SB_ROWS : 1][COLMAJ ? SB_CW + 1 : 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t r0 = (int64_t)blockIdx.x * SB_ROWS;
    const int64_t col0 = (int64_t)blockIdx.y * SB_CW;
    const int64_t cl = col0 + 4 * lane;
    const bool lane_on = cl < d;
    for (int rr = wave; rr < SB_ROWS; rr += SB_NW) {
        const int64_t i = r0 + rr;
        if (i >= mrows) break;
        const int64_t p0 = rowptr[i], p1 = rowptr[i + 1];
        float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        double acc64[4] = {0.0, 0.0, 0.0, 0.0};
        for (int64_t pb = p0; pb < p1; pb += 64) {
            const int cnt = (int)(p1 - pb < 64 ? p1 - pb : 64);
            const int64_t myc = lane < cnt ? colidx[pb + lane] : 0;
            const float myv = lane < cnt ? split_value<TERMS>(vals[pb + lane]) : 0.0f;
            // groups of 4 gathers in flight, the last group padded (its
            // entries past cnt gather row 0 of G, a valid address, and are
            // not applied; cnt is uniform, so is that test): no one-at-a-time
            // tail paying a full memory latency per entry
            for (int q = 0; q < cnt; q += 4) {
                uint2 g[4];
                float a[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int64_t c = __shfl(myc, q + u);  // lanes past cnt hold column 0
                    a[u] = __shfl(myv, q + u);
                    g[u] = lane_on ? *(const uint2 *)(G + c * ldg + cl) : make_uint2(0u, 0u);
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (q + u < cnt) {
                        acc[0] = __fmaf_rn(a[u], __uint_as_float(g[u].x << 16), acc[0]);
                        acc[1] = __fmaf_rn(a[u], __uint_as_float(g[u].x & 0xffff0000u), acc[1]);
                        acc[2] = __fmaf_rn(a[u], __uint_as_float(g[u].y << 16), acc[2]);
                        acc[3] = __fmaf_rn(a[u], __uint_as_float(g[u].y & 0xffff0000u), acc[3]);
                    }
            }
            // long rows: flush the FP32 partial into FP64 every SB_FLUSH entries
            if ((pb + 64 - p0) % SB_FLUSH == 0) {
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    acc64[e] += (double)acc[e];
                    acc[e] = 0.0f;
                }
            }
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int64_t j = cl + e;
            double out = (double)alpha * (acc64[e] + (double)acc[e]);
            if (beta != 0.0f && j < d) out += (double)beta * (double)(COLMAJ ? C[i + j * ldc] : C[i * ldc + j]);
            if (COLMAJ) tile[rr][4 * lane + e] = (float)out;
            else if (j < d) C[i * ldc + j] = (float)out;
        }
    }
    if (COLMAJ) {
        __syncthreads();
        const int rr = tid & 31;
        const int64_t i = r0 + rr;
        for (int cc = tid >> 5; cc < SB_CW; cc += 2 * SB_NW) {
            const int64_t j = col0 + cc;
            if (i < mrows && j < d) C[i + j * ldc] = tile[rr][cc];
        }
    }
}

// dst[p * ldg + q] = bf16(src[p * rs + q * cs]) (a materialised FP32 operator)
__global__ __launch_bounds__(256) void gather_bf16_kernel(const float *src, int64_t rs, int64_t cs, int64_t n_p,
                                                          int64_t n_q, uint16_t *dst, int64_t ldg) {
    const int64_t total = n_p * n_q;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = idx / n_q, q = idx - p * n_q;
        dst[p * ldg + q] = bf16_rne(src[p * rs + q * cs]);
    }
}

template <int TERMS>
static void launch_rows_bf16(bool colmaj, dim3 grid, const rbx_csr_matrix *A, int64_t m, int64_t d, float alpha,
                             float beta, const uint16_t *G, int64_t ldg, float *B, int64_t ldb, hipStream_t st) {
    if (colmaj)
        hipLaunchKernelGGL((csr_rows_bf16_kernel<TERMS, true>), grid, dim3(64 * SB_NW), 0, st, A->rowptr, A->colidxs,
                           (const float *)A->vals, m, d, alpha, beta, G, ldg, B, ldb);
    else
        hipLaunchKernelGGL((csr_rows_bf16_kernel<TERMS, false>), grid, dim3(64 * SB_NW), 0, st, A->rowptr,
                           A->colidxs, (const float *)A->vals, m, d, alpha, beta, G, ldg, B, ldb);
    RBX_HIP(hipGetLastError());
}

static void rsksp3_bf16split_impl(char layout, char opA, char opS, int64_t m, int64_t d, int64_t n, float alpha,
                                  const rbx_csr_matrix *A, int64_t ro_a, int64_t co_a, const rbx_dense_skop *S,
                                  int64_t ro_s, int64_t co_s, float beta, float *B, int64_t ldb, int n_terms,
                                  hipStream_t st) {
    RBX_REQUIRE(A != nullptr && S != nullptr);
    check_layout(layout);
    check_op(opA);
    check_op(opS);
    RBX_REQUIRE(n_terms >= 1 && n_terms <= 3);
    RBX_REQUIRE(m >= 0 && d >= 0 && n >= 0);
    RBX_REQUIRE(A->nnz >= 0);
    if (opA != 'N')
        RBX_FAIL(RBX_UNSUPPORTED, "rsksp3_bf16split: only op(A) = NoTrans with CSR data is supported");
    RBX_REQUIRE(ro_a == 0 && co_a == 0);  // CSR data takes the whole matrix (spmm_dispatch.hh:95-103)
    RBX_REQUIRE(A->n_rows == m && A->n_cols == n);
    DenseGeom geo = dense_geom(S->dist);
    safe_mul(geo.dim_minor, geo.P);
    const int64_t rows_S = opS == 'N' ? n : d, cols_S = opS == 'N' ? d : n;
    RBX_REQUIRE(ro_s >= 0 && co_s >= 0);
    RBX_REQUIRE(S->dist.n_rows >= rows_S + ro_s);
    RBX_REQUIRE(S->dist.n_cols >= cols_S + co_s);
    if (layout == 'C') RBX_REQUIRE(ldb >= m);
    else RBX_REQUIRE(ldb >= d);
    RBX_REQUIRE(B != nullptr);
    if (m == 0 || d == 0) return;
    RBX_REQUIRE(A->nnz == 0 || (A->vals != nullptr && A->colidxs != nullptr));
    RBX_REQUIRE(A->rowptr != nullptr);
    // G = op(bf16(submat(S))), n x d row-major, rows padded to 4 entries
    const int64_t ldg = (d + 3) / 4 * 4;
    uint16_t *G = (uint16_t *)scratch(SCRATCH_DENSE_OP, (size_t)safe_mul(n > 0 ? n : 1, ldg) * 2, st);
    if (n > 0) {
        if (S->buff == nullptr) {
            const GenMap gm = gen_map(*S, opS, ro_s, co_s, false);
            fill_generated<uint16_t>(gm, n, d, G, ldg, 1, st);
        } else {
            if (S->layout != 'C' && S->layout != 'R') RBX_FAIL(RBX_INVALID_ARG, "DenseSkOp.layout must be 'C' or 'R'");
            const int64_t frs = S->layout == 'C' ? 1 : S->dist.n_cols, fcs = S->layout == 'C' ? S->dist.n_rows : 1;
            const float *base = (const float *)S->buff + ro_s * frs + co_s * fcs;
            const int64_t rs = opS == 'N' ? frs : fcs, cs = opS == 'N' ? fcs : frs;
            int64_t blocks = (n * d + 255) / 256;
            if (blocks > 65536) blocks = 65536;
            hipLaunchKernelGGL(gather_bf16_kernel, dim3((unsigned)blocks), dim3(256), 0, st, base, rs, cs, n, d, G,
                               ldg);
            RBX_HIP(hipGetLastError());
        }
    }
    const int64_t nrb = (m + SB_ROWS - 1) / SB_ROWS, nch = (d + SB_CW - 1) / SB_CW;
    RBX_REQUIRE(nrb < (1ll << 31) && nch <= 65535);
    const dim3 grid((unsigned)nrb, (unsigned)nch);
    ProfScope prof("csr_rows_bf16_kernel", st);
    const bool cm = layout == 'C';
    if (n_terms == 1) launch_rows_bf16<1>(cm, grid, A, m, d, alpha, beta, G, ldg, B, ldb, st);
    else if (n_terms == 2) launch_rows_bf16<2>(cm, grid, A, m, d, alpha, beta, G, ldg, B, ldb, st);
    else launch_rows_bf16<3>(cm, grid, A, m, d, alpha, beta, G, ldg, B, ldb, st);
}

}  // namespace rbx

using namespace rbx;

extern "C" rbx_status rbx_srsksp3_bf16split(char layout, char opA, char opS, int64_t m, int64_t d, int64_t n,
                                            float alpha, const rbx_csr_matrix *A, int64_t ro_a, int64_t co_a,
                                            const rbx_dense_skop *S, int64_t ro_s, int64_t co_s, float beta, float *B,
                                            int64_t ldb, int n_terms, void *stream) {
    RBX_API_BEGIN
    rsksp3_bf16split_impl(layout, opA, opS, m, d, n, alpha, A, ro_a, co_a, S, ro_s, co_s, beta, B, ldb, n_terms,
                          (hipStream_t)stream);
    RBX_API_END
}
