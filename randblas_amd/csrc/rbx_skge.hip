// rbx_skge.hip -- dense sketches lskge3 / rskge3 (skge.hh:154-355) and their
// adjoints, in FP64 and FP32, on the gfx950 matrix cores.
//
// Every call is reduced to one column-major product
//     C[M x N] = alpha * X[M x K] * Y[K x N] + beta * C
// in which exactly one of X, Y is the sketching operator G and the other is
// the dense data D, both addressed through (row stride, column stride).  A
// row-major call is the transposed column-major problem.  When the operator
// is not materialised (buff == NULL) each K-slice of G is generated from its
// Philox counters straight into LDS; G never reaches HBM.
//
// Tile: 64 x 256 outputs per 256-thread workgroup when the operator is X
// (256 x 64 when it is Y), 4 waves in 2 x 2, K-slice 16,
// v_mfma_{f64,f32}_16x16x4.  Each generated operator entry feeds 256 outputs.
#include "rbx_internal.hh"
#include "rbx_gemm_common.hh"

namespace rbx {

template <typename T>
struct Mfma16;
template <>
struct Mfma16<double> {
    typedef double acc_t __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // output element r of lane l sits at row l/16 + 4r, column l%16
    static __device__ __forceinline__ int out_row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct Mfma16<float> {
    typedef float acc_t __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // output element r of lane l sits at row 4*(l/16) + r, column l%16
    static __device__ __forceinline__ int out_row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

constexpr int BK = 16, NTHR = 256;

// Workgroup tile BM x BN (4 waves in 2 x 2, each (BM/2) x (BN/2) in 16 x 16
// blocks).  The operator tile is BM x BK (GEN_X) or BK x BN (!GEN_X); the
// other extent is made large (256) so each generated entry feeds 256 outputs.
template <typename T, bool GEN_X, bool S_MEM, int BM, int BN>
__global__ __launch_bounds__(NTHR) void skge_kernel(GemmProblem pb) {
    constexpr int LDX = BM + 1, LDY = BN + 1;
    constexpr int MI = BM / 32, NJ = BN / 32;  // 16x16 blocks per wave
    __shared__ T Xs[BK * LDX];
    __shared__ T Ys[BK * LDY];
    using M16 = Mfma16<T>;
    typename M16::acc_t acc[MI][NJ];
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j) acc[i][j] = typename M16::acc_t{0, 0, 0, 0};

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int64_t m0 = (int64_t)blockIdx.x * BM;
    const int64_t n0 = (int64_t)blockIdx.y * BN;
    const T *D = (const T *)pb.D;

    for (int64_t k0 = 0; k0 < pb.K; k0 += BK) {
        // X tile: rows i in [m0, m0+BM), cols k in [k0, k0+BK) -> Xs[k][i]
        if (GEN_X) {
            if (S_MEM)
                load_tile<T, BM, BK, NTHR>((const T *)pb.sbuf + pb.soff, pb.srs, pb.scs, m0, k0, pb.M, pb.K, Xs, 1, LDX, tid);
            else
                gen_tile<T, BM, BK, NTHR>(pb.g, m0, k0, pb.M, pb.K, Xs, 1, LDX, tid);
            load_tile<T, BK, BN, NTHR>(D, pb.drs, pb.dcs, k0, n0, pb.K, pb.N, Ys, LDY, 1, tid);
        } else {
            load_tile<T, BM, BK, NTHR>(D, pb.drs, pb.dcs, m0, k0, pb.M, pb.K, Xs, 1, LDX, tid);
            if (S_MEM)
                load_tile<T, BK, BN, NTHR>((const T *)pb.sbuf + pb.soff, pb.srs, pb.scs, k0, n0, pb.K, pb.N, Ys, LDY, 1, tid);
            else
                gen_tile<T, BK, BN, NTHR>(pb.g, k0, n0, pb.K, pb.N, Ys, LDY, 1, tid);
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            const int kr = kk + (lane >> 4);
            T a[MI], b[NJ];
#pragma unroll
            for (int i = 0; i < MI; ++i) a[i] = Xs[kr * LDX + wm * (BM / 2) + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < NJ; ++j) b[j] = Ys[kr * LDY + wn * (BN / 2) + j * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < MI; ++i)
#pragma unroll
                for (int j = 0; j < NJ; ++j) acc[i][j] = M16::mma(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }

    // epilogue: element r of the lane's accumulator is row M16::out_row(lane, r),
    // column lane%16 of each 16x16 block (the f64 and f32 forms differ)
    T *Cp = (T *)pb.C;
    const T alpha = (T)pb.alpha, beta = (T)pb.beta;
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int64_t gi = m0 + wm * (BM / 2) + i * 16 + M16::out_row(lane, r);
                int64_t gj = n0 + wn * (BN / 2) + j * 16 + (lane & 15);
                if (gi < pb.M && gj < pb.N) {
                    T *c = Cp + gi * pb.crs + gj * pb.ccs;
                    T v = alpha * acc[i][j][r];
                    if (pb.beta != 0.0) v = v + beta * (*c);
                    *c = v;
                }
            }
}

template <typename T>
static void launch_skge(const GemmProblem &pb, hipStream_t stream) {
    bool gx = pb.gen_is_x, mem = pb.sbuf != nullptr;
    ProfScope prof(sizeof(T) == 8 ? "skge_kernel_f64" : "skge_kernel_f32", stream);
    if (gx) {
        constexpr int TM = 64, TN = 256;
        dim3 grid((unsigned)((pb.M + TM - 1) / TM), (unsigned)((pb.N + TN - 1) / TN));
        RBX_REQUIRE(grid.y <= 65535u);
        if (!mem) hipLaunchKernelGGL((skge_kernel<T, true, false, TM, TN>), grid, dim3(NTHR), 0, stream, pb);
        else hipLaunchKernelGGL((skge_kernel<T, true, true, TM, TN>), grid, dim3(NTHR), 0, stream, pb);
    } else {
        constexpr int TM = 256, TN = 64;
        dim3 grid((unsigned)((pb.M + TM - 1) / TM), (unsigned)((pb.N + TN - 1) / TN));
        RBX_REQUIRE(grid.y <= 65535u);
        if (!mem) hipLaunchKernelGGL((skge_kernel<T, false, false, TM, TN>), grid, dim3(NTHR), 0, stream, pb);
        else hipLaunchKernelGGL((skge_kernel<T, false, true, TM, TN>), grid, dim3(NTHR), 0, stream, pb);
    }
    RBX_HIP(hipGetLastError());
}

// C = beta * C on an M x N strided matrix (util::safe_scal semantics).
template <typename T>
__global__ __launch_bounds__(256) void scale_kernel(T *C, int64_t M, int64_t N, int64_t crs, int64_t ccs, double beta) {
    int64_t total = M * N;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        int64_t i = (crs == 1) ? idx % M : idx / N;
        int64_t j = (crs == 1) ? idx / M : idx % N;
        T *c = C + i * crs + j * ccs;
        *c = (beta == 0.0) ? (T)0 : (T)beta * (*c);
    }
}

template <typename T>
void scale_output(T *C, int64_t M, int64_t N, int64_t crs, int64_t ccs, double beta, hipStream_t stream) {
    if (M <= 0 || N <= 0 || beta == 1.0) return;
    int64_t blocks = (M * N + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    ProfScope prof("scale_kernel", stream);
    hipLaunchKernelGGL(scale_kernel<T>, dim3((unsigned)blocks), dim3(256), 0, stream, C, M, N, crs, ccs, beta);
    RBX_HIP(hipGetLastError());
}
template void scale_output<float>(float *, int64_t, int64_t, int64_t, int64_t, double, hipStream_t);
template void scale_output<double>(double *, int64_t, int64_t, int64_t, int64_t, double, hipStream_t);

// Validate a left sketch call exactly as lskge3 does (skge.hh:174-193) and
// build the column-major problem.
template <typename T>
GemmProblem plan_left(char layout, char opS, char opA, int64_t d, int64_t n, int64_t m, double alpha,
                      const rbx_dense_skop *S, int64_t ro_s, int64_t co_s, const T *A, int64_t lda,
                      double beta, T *B, int64_t ldb) {
    RBX_REQUIRE(S != nullptr);
    check_layout(layout);
    check_op(opS);
    check_op(opA);
    RBX_REQUIRE(d >= 0 && n >= 0 && m >= 0);
    RBX_REQUIRE(ro_s >= 0 && co_s >= 0);
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
    GemmProblem pb{};
    pb.alpha = alpha;
    pb.beta = beta;
    pb.D = A;
    pb.C = B;
    // op(A)[k,j] strides in the column-major picture
    if (layout == 'C') {
        pb.M = d; pb.N = n; pb.K = m;
        pb.gen_is_x = 1;
        pb.g = gen_map(*S, opS, ro_s, co_s, false);
        pb.drs = opA == 'N' ? 1 : lda;
        pb.dcs = opA == 'N' ? lda : 1;
        pb.crs = 1; pb.ccs = ldb;
    } else {
        pb.M = n; pb.N = d; pb.K = m;
        pb.gen_is_x = 0;
        pb.g = gen_map(*S, opS, ro_s, co_s, true);
        pb.drs = opA == 'N' ? 1 : lda;
        pb.dcs = opA == 'N' ? lda : 1;
        pb.crs = 1; pb.ccs = ldb;
    }
    set_mem_operand(pb, *S, opS, ro_s, co_s, layout == 'R');
    return pb;
}

// rskge3 (skge.hh:326-346)
template <typename T>
GemmProblem plan_right(char layout, char opA, char opS, int64_t m, int64_t d, int64_t n, double alpha,
                       const T *A, int64_t lda, const rbx_dense_skop *S, int64_t ro_s, int64_t co_s,
                       double beta, T *B, int64_t ldb) {
    RBX_REQUIRE(S != nullptr);
    check_layout(layout);
    check_op(opS);
    check_op(opA);
    RBX_REQUIRE(d >= 0 && n >= 0 && m >= 0);
    RBX_REQUIRE(ro_s >= 0 && co_s >= 0);
    DenseGeom geo = dense_geom(S->dist);
    safe_mul(geo.dim_minor, geo.P);
    int64_t rows_S = opS == 'N' ? n : d, cols_S = opS == 'N' ? d : n;
    RBX_REQUIRE(S->dist.n_rows >= rows_S + ro_s);
    RBX_REQUIRE(S->dist.n_cols >= cols_S + co_s);
    int64_t rows_A = opA == 'N' ? m : n, cols_A = opA == 'N' ? n : m;
    if (layout == 'C') {
        RBX_REQUIRE(lda >= rows_A);
        RBX_REQUIRE(ldb >= m);
    } else {
        RBX_REQUIRE(lda >= cols_A);
        RBX_REQUIRE(ldb >= d);
    }
    GemmProblem pb{};
    pb.alpha = alpha;
    pb.beta = beta;
    pb.D = A;
    pb.C = B;
    if (layout == 'C') {
        pb.M = m; pb.N = d; pb.K = n;
        pb.gen_is_x = 0;
        pb.g = gen_map(*S, opS, ro_s, co_s, false);
        pb.drs = opA == 'N' ? 1 : lda;
        pb.dcs = opA == 'N' ? lda : 1;
        pb.crs = 1; pb.ccs = ldb;
    } else {
        pb.M = d; pb.N = m; pb.K = n;
        pb.gen_is_x = 1;
        pb.g = gen_map(*S, opS, ro_s, co_s, true);
        pb.drs = opA == 'N' ? 1 : lda;
        pb.dcs = opA == 'N' ? lda : 1;
        pb.crs = 1; pb.ccs = ldb;
    }
    set_mem_operand(pb, *S, opS, ro_s, co_s, layout == 'R');
    return pb;
}

template GemmProblem plan_left<float>(char, char, char, int64_t, int64_t, int64_t, double, const rbx_dense_skop *,
                                      int64_t, int64_t, const float *, int64_t, double, float *, int64_t);

template <typename T>
static void run_problem(const GemmProblem &pb, hipStream_t stream) {
    if (pb.M == 0 || pb.N == 0) return;
    if (pb.alpha == 0.0 || pb.K == 0) {
        scale_output<T>((T *)pb.C, pb.M, pb.N, pb.crs, pb.ccs, pb.beta, stream);
        return;
    }
    launch_skge<T>(pb, stream);
}

}  // namespace rbx

using namespace rbx;

extern "C" {

rbx_status rbx_slskge3(char layout, char opS, char opA, int64_t d, int64_t n, int64_t m, float alpha,
                       const rbx_dense_skop *S, int64_t ro_s, int64_t co_s, const float *A, int64_t lda,
                       float beta, float *B, int64_t ldb, void *stream) {
    RBX_API_BEGIN
    run_problem<float>(plan_left<float>(layout, opS, opA, d, n, m, alpha, S, ro_s, co_s, A, lda, beta, B, ldb),
                       (hipStream_t)stream);
    RBX_API_END
}

rbx_status rbx_dlskge3(char layout, char opS, char opA, int64_t d, int64_t n, int64_t m, double alpha,
                       const rbx_dense_skop *S, int64_t ro_s, int64_t co_s, const double *A, int64_t lda,
                       double beta, double *B, int64_t ldb, void *stream) {
    RBX_API_BEGIN
    run_problem<double>(plan_left<double>(layout, opS, opA, d, n, m, alpha, S, ro_s, co_s, A, lda, beta, B, ldb),
                        (hipStream_t)stream);
    RBX_API_END
}

rbx_status rbx_srskge3(char layout, char opA, char opS, int64_t m, int64_t d, int64_t n, float alpha,
                       const float *A, int64_t lda, const rbx_dense_skop *S, int64_t ro_s, int64_t co_s,
                       float beta, float *B, int64_t ldb, void *stream) {
    RBX_API_BEGIN
    run_problem<float>(plan_right<float>(layout, opA, opS, m, d, n, alpha, A, lda, S, ro_s, co_s, beta, B, ldb),
                       (hipStream_t)stream);
    RBX_API_END
}

rbx_status rbx_drskge3(char layout, char opA, char opS, int64_t m, int64_t d, int64_t n, double alpha,
                       const double *A, int64_t lda, const rbx_dense_skop *S, int64_t ro_s, int64_t co_s,
                       double beta, double *B, int64_t ldb, void *stream) {
    RBX_API_BEGIN
    run_problem<double>(plan_right<double>(layout, opA, opS, m, d, n, alpha, A, lda, S, ro_s, co_s, beta, B, ldb),
                        (hipStream_t)stream);
    RBX_API_END
}

// Adjoint of B = op(S) * A with respect to A: Abar = op(S)^T * Bbar, i.e. the
// same generated operator read with the opposite op.
rbx_status rbx_slskge3_adjoint(char layout, char opS, int64_t d, int64_t n, int64_t m, float alpha,
                               const rbx_dense_skop *S, int64_t ro_s, int64_t co_s, const float *Bbar,
                               int64_t ldbb, float beta, float *Abar, int64_t ldab, void *stream) {
    RBX_API_BEGIN
    check_op(opS);
    char opT = opS == 'N' ? 'T' : 'N';
    run_problem<float>(plan_left<float>(layout, opT, 'N', m, n, d, alpha, S, ro_s, co_s, Bbar, ldbb, beta, Abar, ldab),
                       (hipStream_t)stream);
    RBX_API_END
}

rbx_status rbx_dlskge3_adjoint(char layout, char opS, int64_t d, int64_t n, int64_t m, double alpha,
                               const rbx_dense_skop *S, int64_t ro_s, int64_t co_s, const double *Bbar,
                               int64_t ldbb, double beta, double *Abar, int64_t ldab, void *stream) {
    RBX_API_BEGIN
    check_op(opS);
    char opT = opS == 'N' ? 'T' : 'N';
    run_problem<double>(plan_left<double>(layout, opT, 'N', m, n, d, alpha, S, ro_s, co_s, Bbar, ldbb, beta, Abar, ldab),
                        (hipStream_t)stream);
    RBX_API_END
}

rbx_status rbx_srskge3_adjoint(char layout, char opS, int64_t m, int64_t d, int64_t n, float alpha,
                               const float *Bbar, int64_t ldbb, const rbx_dense_skop *S, int64_t ro_s,
                               int64_t co_s, float beta, float *Abar, int64_t ldab, void *stream) {
    RBX_API_BEGIN
    check_op(opS);
    char opT = opS == 'N' ? 'T' : 'N';
    run_problem<float>(plan_right<float>(layout, 'N', opT, m, n, d, alpha, Bbar, ldbb, S, ro_s, co_s, beta, Abar, ldab),
                       (hipStream_t)stream);
    RBX_API_END
}

rbx_status rbx_drskge3_adjoint(char layout, char opS, int64_t m, int64_t d, int64_t n, double alpha,
                               const double *Bbar, int64_t ldbb, const rbx_dense_skop *S, int64_t ro_s,
                               int64_t co_s, double beta, double *Abar, int64_t ldab, void *stream) {
    RBX_API_BEGIN
    check_op(opS);
    char opT = opS == 'N' ? 'T' : 'N';
    run_problem<double>(plan_right<double>(layout, 'N', opT, m, n, d, alpha, Bbar, ldbb, S, ro_s, co_s, beta, Abar, ldab),
                        (hipStream_t)stream);
    RBX_API_END
}

}  // extern "C"
