// rbx_solve.hip -- sketch-and-solve least squares on the device, the
// downstream use of a left sketch (the reference's pattern in
// examples/total-least-squares/tls_dense_skop.cc:139-186, with a QR solve of
// the sketched system in place of its SVD):
//   x = argmin_x || S (A x - b) ||_2,   S = submat(S) (d x m), A m x n, b m,
// SA and Sb by the fused sketch (the operator is never stored), then the QR
// of the d x n sketched system and the triangular solve in FP64 with
// rocSOLVER / rocBLAS for both precisions.  The adjoint gives the gradient of
// <xbar, x(A, b)>:
//   u = R^{-T} xbar,  w = Q [u; 0],  z = R^{-1} u,  r = Sb - SA x,
//   bbar = S^T w,     Abar = S^T (r z^T - w x^T).
#include <mutex>
#include <vector>

#include "rbx_internal.hh"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

namespace rbx {

// one rocBLAS handle per device, bound to the caller's stream on each call
static rocblas_handle blas_handle(hipStream_t st) {
    static std::mutex mu;
    static std::vector<rocblas_handle> handles;
    int dev = 0;
    RBX_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    if ((int)handles.size() <= dev) handles.resize(dev + 1, nullptr);
    if (handles[dev] == nullptr && rocblas_create_handle(&handles[dev]) != rocblas_status_success)
        RBX_FAIL(RBX_HIP_ERROR, "sketch_solve: rocblas_create_handle failed");
    if (rocblas_set_stream(handles[dev], st) != rocblas_status_success)
        RBX_FAIL(RBX_HIP_ERROR, "sketch_solve: rocblas_set_stream failed");
    return handles[dev];
}

static void blas_ok(rocblas_status s, const char *what) {
    if (s != rocblas_status_success) RBX_FAIL(RBX_HIP_ERROR, "sketch_solve: %s failed (rocblas status %d)", what, (int)s);
}

static void inner_ok(rbx_status s) {
    if (s != RBX_OK) throw Fail{s};  // the inner call has set the message
}

// dst (r x c, FP64 column-major, ld r) = src (T, element (i, j) at i*rs + j*cs)
template <typename T>
__global__ __launch_bounds__(256) void to_f64_kernel(const T *src, int64_t rs, int64_t cs, int64_t r, int64_t c,
                                                     double *dst) {
    const int64_t total = r * c;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e % r, j = e / r;
        dst[e] = (double)src[i * rs + j * cs];
    }
}

// dst (T, element (i, j) at i*rs + j*cs) = src (r x c FP64 column-major)
template <typename T>
__global__ __launch_bounds__(256) void from_f64_kernel(const double *src, int64_t r, int64_t c, T *dst, int64_t rs,
                                                       int64_t cs) {
    const int64_t total = r * c;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e % r, j = e / r;
        dst[i * rs + j * cs] = (T)src[e];
    }
}

// M (d x n, FP64 column-major) = r z^T - w x^T
__global__ __launch_bounds__(256) void rank2_kernel(const double *r, const double *z, const double *w, const double *x,
                                                    int64_t d, int64_t n, double *M) {
    const int64_t total = d * n;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e % d, j = e / d;
        M[e] = r[i] * z[j] - w[i] * x[j];
    }
}

__global__ void zero_tail_kernel(double *v, int64_t from, int64_t to) {
    for (int64_t i = from + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < to; i += (int64_t)gridDim.x * blockDim.x)
        v[i] = 0.0;
}

static unsigned grid_of(int64_t n) {
    int64_t b = (n + 255) / 256;
    if (b < 1) b = 1;
    return (unsigned)(b > 4096 ? 4096 : b);
}

template <typename T>
static rbx_status lsk(char layout, int64_t d, int64_t n, int64_t m, const rbx_dense_skop *S, int64_t ro_s, int64_t co_s,
                      const T *A, int64_t lda, T *B, int64_t ldb, hipStream_t st) {
    if constexpr (sizeof(T) == 8) return rbx_dlskge3(layout, 'N', 'N', d, n, m, 1.0, S, ro_s, co_s, A, lda, 0.0, B, ldb, st);
    else return rbx_slskge3(layout, 'N', 'N', d, n, m, 1.0f, S, ro_s, co_s, A, lda, 0.0f, B, ldb, st);
}

template <typename T>
static rbx_status lsk_adj(char layout, int64_t d, int64_t n, int64_t m, const rbx_dense_skop *S, int64_t ro_s,
                          int64_t co_s, const T *G, int64_t ldg, T *Abar, int64_t ldab, hipStream_t st) {
    if constexpr (sizeof(T) == 8)
        return rbx_dlskge3_adjoint(layout, 'N', d, n, m, 1.0, S, ro_s, co_s, G, ldg, 0.0, Abar, ldab, st);
    else return rbx_slskge3_adjoint(layout, 'N', d, n, m, 1.0f, S, ro_s, co_s, G, ldg, 0.0f, Abar, ldab, st);
}

// Workspace of one solve, in the solve scratch slot.
template <typename T>
struct SolveWork {
    double *SA, *QR, *Sb, *y, *tau;  // d x n, d x n, d, d, n
    T *tmpT;                         // d x n in T (the sketches before promotion)
};

template <typename T>
static SolveWork<T> solve_work(int64_t d, int64_t n) {
    auto a = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t dn = (size_t)safe_mul(d, n);
    const size_t bytes = 2 * a(dn * 8) + 2 * a((size_t)d * 8) + a((size_t)n * 8) + a(dn * sizeof(T) + 8);
    char *ws = (char *)scratch(SCRATCH_SOLVE, bytes);
    SolveWork<T> w;
    w.SA = (double *)ws;
    w.QR = (double *)(ws + a(dn * 8));
    w.Sb = (double *)(ws + 2 * a(dn * 8));
    w.y = (double *)(ws + 2 * a(dn * 8) + a((size_t)d * 8));
    w.tau = (double *)(ws + 2 * a(dn * 8) + 2 * a((size_t)d * 8));
    w.tmpT = (T *)(ws + 2 * a(dn * 8) + 2 * a((size_t)d * 8) + a((size_t)n * 8));
    return w;
}

// SA, Sb (FP64), QR = geqrf(SA), y = Q^T Sb, x64 = R^{-1} y[0:n] (left in y)
template <typename T>
static void factor_and_solve(char layout, int64_t d, int64_t m, int64_t n, const rbx_dense_skop *S, int64_t ro_s,
                             int64_t co_s, const T *A, int64_t lda, const T *b, const SolveWork<T> &w,
                             hipStream_t st) {
    RBX_REQUIRE(d <= INT32_MAX && n <= INT32_MAX);
    // SA in T, in `layout` with leading dimension d (C) or n (R), promoted
    const int64_t ldsa = layout == 'C' ? d : n;
    inner_ok(lsk<T>(layout, d, n, m, S, ro_s, co_s, A, lda, w.tmpT, ldsa, st));
    hipLaunchKernelGGL((to_f64_kernel<T>), dim3(grid_of(d * n)), dim3(256), 0, st, (const T *)w.tmpT,
                       layout == 'C' ? (int64_t)1 : ldsa, layout == 'C' ? ldsa : (int64_t)1, d, n, w.SA);
    RBX_HIP(hipGetLastError());
    // Sb: a one-column column-major sketch
    inner_ok(lsk<T>('C', d, 1, m, S, ro_s, co_s, b, m, w.tmpT, d, st));
    hipLaunchKernelGGL((to_f64_kernel<T>), dim3(grid_of(d)), dim3(256), 0, st, (const T *)w.tmpT, (int64_t)1, d, d,
                       (int64_t)1, w.Sb);
    RBX_HIP(hipGetLastError());
    RBX_HIP(hipMemcpyAsync(w.QR, w.SA, (size_t)(d * n) * 8, hipMemcpyDeviceToDevice, st));
    RBX_HIP(hipMemcpyAsync(w.y, w.Sb, (size_t)d * 8, hipMemcpyDeviceToDevice, st));
    rocblas_handle h = blas_handle(st);
    blas_ok(rocsolver_dgeqrf(h, (rocblas_int)d, (rocblas_int)n, w.QR, (rocblas_int)d, w.tau), "geqrf");
    blas_ok(rocsolver_dormqr(h, rocblas_side_left, rocblas_operation_transpose, (rocblas_int)d, 1, (rocblas_int)n,
                             w.QR, (rocblas_int)d, w.tau, w.y, (rocblas_int)d),
            "ormqr");
    blas_ok(rocblas_dtrsv(h, rocblas_fill_upper, rocblas_operation_none, rocblas_diagonal_non_unit, (rocblas_int)n,
                          w.QR, (rocblas_int)d, w.y, 1),
            "trsv");
}

template <typename T>
static void check_solve(char layout, int64_t d, int64_t m, int64_t n, const rbx_dense_skop *S, const T *A, int64_t lda,
                        const T *b) {
    check_layout(layout);
    RBX_REQUIRE(S != nullptr);
    RBX_REQUIRE(n >= 1 && d >= n && m >= 1);
    RBX_REQUIRE(A != nullptr && b != nullptr);
    RBX_REQUIRE(lda >= (layout == 'C' ? m : n));
}

template <typename T>
static void sketch_solve_impl(char layout, int64_t d, int64_t m, int64_t n, const rbx_dense_skop *S, int64_t ro_s,
                              int64_t co_s, const T *A, int64_t lda, const T *b, T *x, hipStream_t st) {
    check_solve<T>(layout, d, m, n, S, A, lda, b);
    RBX_REQUIRE(x != nullptr);
    SolveWork<T> w = solve_work<T>(d, n);
    factor_and_solve<T>(layout, d, m, n, S, ro_s, co_s, A, lda, b, w, st);
    hipLaunchKernelGGL((from_f64_kernel<T>), dim3(grid_of(n)), dim3(256), 0, st, (const double *)w.y, n, (int64_t)1, x,
                       (int64_t)1, (int64_t)1);
    RBX_HIP(hipGetLastError());
}

template <typename T>
static void sketch_solve_adjoint_impl(char layout, int64_t d, int64_t m, int64_t n, const rbx_dense_skop *S,
                                      int64_t ro_s, int64_t co_s, const T *A, int64_t lda, const T *b, const T *xbar,
                                      T *Abar, int64_t ldab, T *bbar, hipStream_t st) {
    check_solve<T>(layout, d, m, n, S, A, lda, b);
    RBX_REQUIRE(xbar != nullptr && Abar != nullptr && bbar != nullptr);
    RBX_REQUIRE(ldab >= (layout == 'C' ? m : n));
    SolveWork<T> w = solve_work<T>(d, n);
    factor_and_solve<T>(layout, d, m, n, S, ro_s, co_s, A, lda, b, w, st);
    // second workspace: x64 (n), u (n), z (n), r (d), wv (d), M (d x n) in FP64, M in T
    auto a = [](size_t bb) { return (bb + 255) / 256 * 256; };
    const size_t dn = (size_t)(d * n);
    char *ws = (char *)scratch(SCRATCH_SOLVE2, 3 * a((size_t)n * 8) + 2 * a((size_t)d * 8) + a(dn * 8) + a(dn * sizeof(T)));
    double *x64 = (double *)ws, *u = (double *)(ws + a((size_t)n * 8)), *z = (double *)(ws + 2 * a((size_t)n * 8));
    double *r = (double *)(ws + 3 * a((size_t)n * 8)), *wv = (double *)(ws + 3 * a((size_t)n * 8) + a((size_t)d * 8));
    double *M = (double *)(ws + 3 * a((size_t)n * 8) + 2 * a((size_t)d * 8));
    T *MT = (T *)(ws + 3 * a((size_t)n * 8) + 2 * a((size_t)d * 8) + a(dn * 8));
    rocblas_handle h = blas_handle(st);
    RBX_HIP(hipMemcpyAsync(x64, w.y, (size_t)n * 8, hipMemcpyDeviceToDevice, st));
    // u = R^{-T} xbar
    hipLaunchKernelGGL((to_f64_kernel<T>), dim3(grid_of(n)), dim3(256), 0, st, xbar, (int64_t)1, n, n, (int64_t)1, u);
    RBX_HIP(hipGetLastError());
    blas_ok(rocblas_dtrsv(h, rocblas_fill_upper, rocblas_operation_transpose, rocblas_diagonal_non_unit, (rocblas_int)n,
                          w.QR, (rocblas_int)d, u, 1),
            "trsv");
    // w = Q [u; 0]
    RBX_HIP(hipMemcpyAsync(wv, u, (size_t)n * 8, hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(zero_tail_kernel, dim3(grid_of(d - n)), dim3(256), 0, st, wv, n, d);
    RBX_HIP(hipGetLastError());
    blas_ok(rocsolver_dormqr(h, rocblas_side_left, rocblas_operation_none, (rocblas_int)d, 1, (rocblas_int)n, w.QR,
                             (rocblas_int)d, w.tau, wv, (rocblas_int)d),
            "ormqr");
    // z = R^{-1} u
    RBX_HIP(hipMemcpyAsync(z, u, (size_t)n * 8, hipMemcpyDeviceToDevice, st));
    blas_ok(rocblas_dtrsv(h, rocblas_fill_upper, rocblas_operation_none, rocblas_diagonal_non_unit, (rocblas_int)n,
                          w.QR, (rocblas_int)d, z, 1),
            "trsv");
    // r = Sb - SA x
    RBX_HIP(hipMemcpyAsync(r, w.Sb, (size_t)d * 8, hipMemcpyDeviceToDevice, st));
    const double one = 1.0, mone = -1.0;
    blas_ok(rocblas_dgemv(h, rocblas_operation_none, (rocblas_int)d, (rocblas_int)n, &mone, w.SA, (rocblas_int)d, x64, 1,
                          &one, r, 1),
            "gemv");
    // M = r z^T - w x^T, in `layout` and T for the adjoint sketch
    hipLaunchKernelGGL(rank2_kernel, dim3(grid_of(d * n)), dim3(256), 0, st, (const double *)r, (const double *)z,
                       (const double *)wv, (const double *)x64, d, n, M);
    const int64_t ldm = layout == 'C' ? d : n;
    hipLaunchKernelGGL((from_f64_kernel<T>), dim3(grid_of(d * n)), dim3(256), 0, st, (const double *)M, d, n, MT,
                       layout == 'C' ? (int64_t)1 : ldm, layout == 'C' ? ldm : (int64_t)1);
    RBX_HIP(hipGetLastError());
    inner_ok(lsk_adj<T>(layout, d, n, m, S, ro_s, co_s, MT, ldm, Abar, ldab, st));
    // bbar = S^T w (one column)
    hipLaunchKernelGGL((from_f64_kernel<T>), dim3(grid_of(d)), dim3(256), 0, st, (const double *)wv, d, (int64_t)1, MT,
                       (int64_t)1, d);
    RBX_HIP(hipGetLastError());
    inner_ok(lsk_adj<T>('C', d, 1, m, S, ro_s, co_s, MT, d, bbar, m, st));
}

}  // namespace rbx

using namespace rbx;

extern "C" {

rbx_status rbx_ssketch_solve(char layout, int64_t d, int64_t m, int64_t n, const rbx_dense_skop *S, int64_t ro_s,
                             int64_t co_s, const float *A, int64_t lda, const float *b, float *x, void *stream) {
    RBX_API_BEGIN
    sketch_solve_impl<float>(layout, d, m, n, S, ro_s, co_s, A, lda, b, x, (hipStream_t)stream);
    RBX_API_END
}

rbx_status rbx_dsketch_solve(char layout, int64_t d, int64_t m, int64_t n, const rbx_dense_skop *S, int64_t ro_s,
                             int64_t co_s, const double *A, int64_t lda, const double *b, double *x, void *stream) {
    RBX_API_BEGIN
    sketch_solve_impl<double>(layout, d, m, n, S, ro_s, co_s, A, lda, b, x, (hipStream_t)stream);
    RBX_API_END
}

rbx_status rbx_ssketch_solve_adjoint(char layout, int64_t d, int64_t m, int64_t n, const rbx_dense_skop *S,
                                     int64_t ro_s, int64_t co_s, const float *A, int64_t lda, const float *b,
                                     const float *xbar, float *Abar, int64_t ldab, float *bbar, void *stream) {
    RBX_API_BEGIN
    sketch_solve_adjoint_impl<float>(layout, d, m, n, S, ro_s, co_s, A, lda, b, xbar, Abar, ldab, bbar,
                                     (hipStream_t)stream);
    RBX_API_END
}

rbx_status rbx_dsketch_solve_adjoint(char layout, int64_t d, int64_t m, int64_t n, const rbx_dense_skop *S,
                                     int64_t ro_s, int64_t co_s, const double *A, int64_t lda, const double *b,
                                     const double *xbar, double *Abar, int64_t ldab, double *bbar, void *stream) {
    RBX_API_BEGIN
    sketch_solve_adjoint_impl<double>(layout, d, m, n, S, ro_s, co_s, A, lda, b, xbar, Abar, ldab, bbar,
                                      (hipStream_t)stream);
    RBX_API_END
}

}  // extern "C"
