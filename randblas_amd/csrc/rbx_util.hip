// rbx_util.hip -- util::require_symmetric (RandBLAS/util.hh:127-148) on the
// device.  Every pair i < j is tested as the reference does, in T:
//   viol = |A(i,j) - A(j,i)|,  rel_tol = (|A(i,j)| + |A(j,i)| + 1) * tol,
// and the first failing pair in the reference's loop order (row i ascending,
// then column j) is reported, i.e. the smallest linear index i*n + j.
#include <cmath>

#include "rbx_internal.hh"

namespace rbx {

__global__ void u64_set_kernel(unsigned long long *p, unsigned long long v) { *p = v; }

template <typename T>
__global__ __launch_bounds__(256) void sym_check_kernel(const T *A, int64_t n, int64_t rs, int64_t cs, T tol,
                                                        unsigned long long *first) {
    const int64_t bi = blockIdx.y, bj = blockIdx.x;
    if (bj < bi) return;  // tile strictly below the diagonal: no pair i < j
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const int64_t j = bj * 32 + tx;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int64_t i = bi * 32 + ty + 8 * q;
        if (i < n && j < n && j > i) {
            const T a = A[i * rs + j * cs], b = A[j * rs + i * cs];
            const T viol = fabs(a - b);
            const T rel = (fabs(a) + fabs(b) + (T)1) * tol;
            if (viol > rel) atomicMin(first, (unsigned long long)(i * n + j));
        }
    }
}

template <typename T>
static void require_symmetric_impl(char layout, const T *A, int64_t n, int64_t lda, T tol, hipStream_t stream) {
    check_layout(layout);
    if (tol < 0) return;
    RBX_REQUIRE(n >= 0);
    if (n < 2) return;
    RBX_REQUIRE(A != nullptr);
    RBX_REQUIRE(lda >= n);
    safe_mul(n, n);
    const int64_t rs = layout == 'C' ? 1 : lda, cs = layout == 'C' ? lda : 1;
    const int64_t nb = (n + 31) / 32;
    RBX_REQUIRE(nb <= 65535);
    unsigned long long *first = (unsigned long long *)scratch(SCRATCH_SMALL, sizeof(unsigned long long), stream);
    const unsigned long long none = ~0ull;
    hipLaunchKernelGGL(u64_set_kernel, dim3(1), dim3(1), 0, stream, first, none);
    RBX_HIP(hipGetLastError());
    {
        ProfScope prof("sym_check_kernel", stream);
        hipLaunchKernelGGL((sym_check_kernel<T>), dim3((unsigned)nb, (unsigned)nb), dim3(256), 0, stream, A, n, rs, cs,
                           tol, first);
        RBX_HIP(hipGetLastError());
    }
    unsigned long long idx = none;
    RBX_HIP(hipMemcpyAsync(&idx, first, sizeof(idx), hipMemcpyDeviceToHost, stream));
    RBX_HIP(hipStreamSynchronize(stream));
    if (idx == none) return;
    const int64_t i = (int64_t)(idx / (unsigned long long)n), j = (int64_t)(idx % (unsigned long long)n);
    T a, b;
    RBX_HIP(hipMemcpy(&a, A + i * rs + j * cs, sizeof(T), hipMemcpyDeviceToHost));
    RBX_HIP(hipMemcpy(&b, A + j * rs + i * cs, sizeof(T), hipMemcpyDeviceToHost));
    const T viol = std::fabs(a - b);
    const T rel = (std::fabs(a) + std::fabs(b) + (T)1) * tol;
    char msg[256];
    snprintf(msg, sizeof(msg),
             "Symmetry check failed. |A(%lld,%lld) - A(%lld,%lld)| was %e, which exceeds tolerance of %e.",
             (long long)i, (long long)j, (long long)j, (long long)i, (double)viol, (double)rel);
    RBX_FAIL(RBX_INVALID_ARG, "%s", msg);
}

}  // namespace rbx

using namespace rbx;

extern "C" {

rbx_status rbx_srequire_symmetric(char layout, const float *A, int64_t n, int64_t lda, float tol, void *stream) {
    RBX_API_BEGIN
    require_symmetric_impl<float>(layout, A, n, lda, tol, (hipStream_t)stream);
    RBX_API_END
}

rbx_status rbx_drequire_symmetric(char layout, const double *A, int64_t n, int64_t lda, double tol, void *stream) {
    RBX_API_BEGIN
    require_symmetric_impl<double>(layout, A, n, lda, tol, (hipStream_t)stream);
    RBX_API_END
}

}  // extern "C"
