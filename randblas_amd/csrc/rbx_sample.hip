// rbx_sample.hip -- index sampling and CDF preparation on the device:
//   sample_indices_iid_uniform  (RandBLAS/util.hh:516-560)
//   sample_indices_iid          (RandBLAS/util.hh:493-514)
//   weights_to_cdf              (RandBLAS/util.hh:459-472)
// Each sample reads the same Philox words as the reference's sequential loop:
// without signs, sample i uses word i mod 4 of counter ctr + floor(i / 4);
// with Rademacher signs, sample i uses words 2(i mod 2), 2(i mod 2) + 1 of
// counter ctr + floor(i / 2).  The returned state is the seed advanced by the
// number of counters the loop touched.  One thread per counter.
#include <cmath>
#include <limits>

#include "rbx_internal.hh"

namespace rbx {

// (sint) ((double) n * ((double) uneg11(w) + 1) / 2), util.hh:531-533
__device__ __forceinline__ int64_t uniform_index(uint32_t w, double dN) {
    const double u = __dmul_rn(__dadd_rn((double)uneg11(w), 1.0), 0.5);
    return (int64_t)__dmul_rn(dN, u);
}

template <typename T>
__global__ __launch_bounds__(256) void iid_uniform_kernel(Ctr128 ctr, uint32_t k0, uint32_t k1, int64_t n, int64_t k,
                                                          int64_t *samples, T *rademachers) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int per = rademachers ? 2 : 4;
    const int64_t i0 = c * per;
    if (i0 >= k) return;
    const uint4 w4 = philox4x32_10(ctr_add(ctr, (uint64_t)c), k0, k1);
    const uint32_t w[4] = {w4.x, w4.y, w4.z, w4.w};
    const double dN = (double)n;
    for (int q = 0; q < per && i0 + q < k; ++q) {
        if (rademachers) {
            samples[i0 + q] = uniform_index(w[2 * q], dN);
            rademachers[i0 + q] = uneg11(w[2 * q + 1]) >= 0.0f ? (T)1 : (T)-1;
        } else {
            samples[i0 + q] = uniform_index(w[q], dN);
        }
    }
}

// u = ((T) uneg11(w) + 1) / 2 in T (uneg11_to_u01<T>, util.hh:474-477), then
// the first index whose CDF value is not below u (std::lower_bound).
template <typename T>
__global__ __launch_bounds__(256) void iid_cdf_kernel(Ctr128 ctr, uint32_t k0, uint32_t k1, int64_t n, const T *cdf,
                                                      int64_t k, int64_t *samples) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t i0 = c * 4;
    if (i0 >= k) return;
    const uint4 w4 = philox4x32_10(ctr_add(ctr, (uint64_t)c), k0, k1);
    const uint32_t w[4] = {w4.x, w4.y, w4.z, w4.w};
    for (int q = 0; q < 4 && i0 + q < k; ++q) {
        const T u = ((T)uneg11(w[q]) + (T)1) / (T)2;
        int64_t lo = 0, len = n;
        while (len > 0) {
            const int64_t half = len / 2, mid = lo + half;
            if (cdf[mid] < u) {
                lo = mid + 1;
                len -= half + 1;
            } else {
                len = half;
            }
        }
        samples[i0 + q] = lo;
    }
}

// Sequential, in the reference's order, so the partial sums round exactly as
// its loop does.  status: 0 ok; 1 + i: w[i] < error_if_below (w[0..i) hold the
// partial sums, as when the reference throws there); -1: the total is below
// sqrt(n) * eps (w holds the partial sums, unscaled).
template <typename T>
__global__ void weights_to_cdf_kernel(int64_t n, T *w, T error_if_below, long long *status) {
    T sum = (T)0;
    for (int64_t i = 0; i < n; ++i) {
        T val = w[i];
        if (!(val >= error_if_below)) {
            *status = 1 + i;
            return;
        }
        val = val > (T)0 ? val : (T)0;
        sum = sum + val;
        w[i] = sum;
    }
    const T floor_sum = (T)sqrt((double)n) * std::numeric_limits<T>::epsilon();
    if (!(sum >= floor_sum)) {
        *status = -1;
        return;
    }
    const T inv = (T)1 / sum;
    for (int64_t i = 0; i < n; ++i) w[i] = w[i] * inv;
    *status = 0;
}

__global__ void i64_zero_kernel(long long *p) { *p = 0; }

static void next_state(const rbx_rng_state *state, uint64_t counters, rbx_rng_state *next) {
    if (next == nullptr) return;
    rbx_rng_state s = *state;
    ctr_to_words(ctr_add(ctr_from_words(s.counter), counters), s.counter);
    *next = s;
}

template <typename T>
static void iid_uniform_impl(int64_t n, int64_t k, int64_t *samples, T *rademachers, const rbx_rng_state *state,
                             rbx_rng_state *next, hipStream_t st) {
    RBX_REQUIRE(state != nullptr);
    RBX_REQUIRE(n >= 0 && k >= 0);
    const int per = rademachers ? 2 : 4;
    const int64_t counters = (k + per - 1) / per;
    if (k > 0) {
        RBX_REQUIRE(samples != nullptr);
        const Ctr128 ctr = ctr_from_words(state->counter);
        const int64_t blocks = (counters + 255) / 256;
        RBX_REQUIRE(blocks < (1ll << 31));
        ProfScope prof("iid_uniform_kernel", st);
        hipLaunchKernelGGL((iid_uniform_kernel<T>), dim3((unsigned)blocks), dim3(256), 0, st, ctr, state->key[0],
                           state->key[1], n, k, samples, rademachers);
        RBX_HIP(hipGetLastError());
    }
    next_state(state, (uint64_t)counters, next);
}

template <typename T>
static void iid_cdf_impl(int64_t n, const T *cdf, int64_t k, int64_t *samples, const rbx_rng_state *state,
                         rbx_rng_state *next, hipStream_t st) {
    RBX_REQUIRE(state != nullptr);
    RBX_REQUIRE(n >= 0 && k >= 0);
    const int64_t counters = (k + 3) / 4;
    if (k > 0) {
        RBX_REQUIRE(samples != nullptr && (n == 0 || cdf != nullptr));
        const Ctr128 ctr = ctr_from_words(state->counter);
        const int64_t blocks = (counters + 255) / 256;
        RBX_REQUIRE(blocks < (1ll << 31));
        ProfScope prof("iid_cdf_kernel", st);
        hipLaunchKernelGGL((iid_cdf_kernel<T>), dim3((unsigned)blocks), dim3(256), 0, st, ctr, state->key[0],
                           state->key[1], n, cdf, k, samples);
        RBX_HIP(hipGetLastError());
    }
    next_state(state, (uint64_t)counters, next);
}

template <typename T>
static void weights_to_cdf_impl(int64_t n, T *w, T error_if_below, hipStream_t st) {
    RBX_REQUIRE(n >= 0);
    RBX_REQUIRE(n == 0 || w != nullptr);
    long long *status = (long long *)scratch(SCRATCH_SMALL, sizeof(long long));
    hipLaunchKernelGGL(i64_zero_kernel, dim3(1), dim3(1), 0, st, status);
    hipLaunchKernelGGL((weights_to_cdf_kernel<T>), dim3(1), dim3(1), 0, st, n, w, error_if_below, status);
    RBX_HIP(hipGetLastError());
    long long h = 0;
    RBX_HIP(hipMemcpyAsync(&h, status, sizeof(h), hipMemcpyDeviceToHost, st));
    RBX_HIP(hipStreamSynchronize(st));
    if (h > 0)
        RBX_FAIL(RBX_INVALID_ARG, "weights_to_cdf: w[%lld] is below error_if_below = %e", h - 1,
                 (double)error_if_below);
    if (h < 0) RBX_FAIL(RBX_INVALID_ARG, "weights_to_cdf: the weights sum to less than sqrt(n) * epsilon");
}

}  // namespace rbx

using namespace rbx;

extern "C" {

rbx_status rbx_sample_indices_iid_uniform(int64_t n, int64_t k, int64_t *samples, const rbx_rng_state *state,
                                          rbx_rng_state *next_state_out, void *stream) {
    RBX_API_BEGIN
    iid_uniform_impl<float>(n, k, samples, nullptr, state, next_state_out, (hipStream_t)stream);
    RBX_API_END
}

rbx_status rbx_ssample_indices_iid_uniform_signed(int64_t n, int64_t k, int64_t *samples, float *rademachers,
                                                  const rbx_rng_state *state, rbx_rng_state *next_state_out,
                                                  void *stream) {
    RBX_API_BEGIN
    RBX_REQUIRE(k == 0 || rademachers != nullptr);
    iid_uniform_impl<float>(n, k, samples, rademachers, state, next_state_out, (hipStream_t)stream);
    RBX_API_END
}

rbx_status rbx_dsample_indices_iid_uniform_signed(int64_t n, int64_t k, int64_t *samples, double *rademachers,
                                                  const rbx_rng_state *state, rbx_rng_state *next_state_out,
                                                  void *stream) {
    RBX_API_BEGIN
    RBX_REQUIRE(k == 0 || rademachers != nullptr);
    iid_uniform_impl<double>(n, k, samples, rademachers, state, next_state_out, (hipStream_t)stream);
    RBX_API_END
}

rbx_status rbx_ssample_indices_iid(int64_t n, const float *cdf, int64_t k, int64_t *samples,
                                   const rbx_rng_state *state, rbx_rng_state *next_state_out, void *stream) {
    RBX_API_BEGIN
    iid_cdf_impl<float>(n, cdf, k, samples, state, next_state_out, (hipStream_t)stream);
    RBX_API_END
}

rbx_status rbx_dsample_indices_iid(int64_t n, const double *cdf, int64_t k, int64_t *samples,
                                   const rbx_rng_state *state, rbx_rng_state *next_state_out, void *stream) {
    RBX_API_BEGIN
    iid_cdf_impl<double>(n, cdf, k, samples, state, next_state_out, (hipStream_t)stream);
    RBX_API_END
}

rbx_status rbx_sweights_to_cdf(int64_t n, float *w, float error_if_below, void *stream) {
    RBX_API_BEGIN
    weights_to_cdf_impl<float>(n, w, error_if_below, (hipStream_t)stream);
    RBX_API_END
}

rbx_status rbx_dweights_to_cdf(int64_t n, double *w, double error_if_below, void *stream) {
    RBX_API_BEGIN
    weights_to_cdf_impl<double>(n, w, error_if_below, (hipStream_t)stream);
    RBX_API_END
}

}  // extern "C"
