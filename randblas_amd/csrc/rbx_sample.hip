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
#include <type_traits>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

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
    // whole counters with 16-byte aligned outputs: vector stores (a thread's
    // samples are contiguous, so its 32 or 16 bytes go out in one or two pieces)
    const bool vec = i0 + per <= k && ((uintptr_t)(samples + i0) & 15) == 0 &&
                     (rademachers == nullptr || ((uintptr_t)(rademachers + i0) & (2 * sizeof(T) - 1)) == 0);
    if (vec) {
        typedef long long i64x2 __attribute__((ext_vector_type(2)));
        typedef T tx2 __attribute__((ext_vector_type(2)));
        if (rademachers) {
            const i64x2 sv = {uniform_index(w[0], dN), uniform_index(w[2], dN)};
            const tx2 rv = {uneg11(w[1]) >= 0.0f ? (T)1 : (T)-1, uneg11(w[3]) >= 0.0f ? (T)1 : (T)-1};
            *(i64x2 *)(samples + i0) = sv;
            *(tx2 *)(rademachers + i0) = rv;
        } else {
            const i64x2 s01 = {uniform_index(w[0], dN), uniform_index(w[1], dN)};
            const i64x2 s23 = {uniform_index(w[2], dN), uniform_index(w[3], dN)};
            *(i64x2 *)(samples + i0) = s01;
            *(i64x2 *)(samples + i0 + 2) = s23;
        }
        return;
    }
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
// the first index whose CDF value is not below u, by std::lower_bound's own
// probe sequence (half = len / 2, mid = lo + half).  The first CDF_LEVELS
// levels of that sequence probe at most 2^CDF_LEVELS - 1 positions, the
// same for every u: each workgroup loads them into LDS as a heap-ordered
// tree (node t, children 2t for cdf[mid] >= u and 2t + 1 for cdf[mid] < u),
// so only the remaining levels read global memory; results are identical
// to the plain search for any array.  The grid is a few workgroups per CU,
// each looping over counters, so a tree serves many draws; a thread's four
// searches of a counter run in lockstep, their global probes in flight
// together.
constexpr int CDF_LEVELS = 12, CDF_NODES = (1 << CDF_LEVELS) - 1;
template <typename T>
__global__ __launch_bounds__(256) void iid_cdf_kernel(Ctr128 ctr, uint32_t k0, uint32_t k1, int64_t n, const T *cdf,
                                                      int64_t k, int64_t *samples) {
    __shared__ T tree[CDF_NODES + 1];
    for (int t = threadIdx.x + 1; t <= CDF_NODES; t += blockDim.x) {
        int64_t lo = 0, len = n;
        const int depth = 31 - __builtin_clz((unsigned)t);
        for (int bit = depth - 1; bit >= 0; --bit) {
            const int64_t half = len / 2, mid = lo + half;
            if ((t >> bit) & 1) {
                lo = mid + 1;
                len -= half + 1;
            } else {
                len = half;
            }
        }
        tree[t] = len > 0 ? cdf[lo + len / 2] : (T)0;  // unused when len == 0
    }
    __syncthreads();
    const int64_t ncounters = (k + 3) / 4;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < ncounters;
         c += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i0 = c * 4;
        const uint4 w4 = philox4x32_10(ctr_add(ctr, (uint64_t)c), k0, k1);
        const uint32_t w[4] = {w4.x, w4.y, w4.z, w4.w};
        T u[4];
        int64_t lo[4], len[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            u[q] = ((T)uneg11(w[q]) + (T)1) / (T)2;
            lo[q] = 0;
            len[q] = n;
            int t = 1;
            for (int lev = 0; lev < CDF_LEVELS && len[q] > 0; ++lev) {
                const int64_t half = len[q] / 2;
                if (tree[t] < u[q]) {
                    lo[q] += half + 1;
                    len[q] -= half + 1;
                    t = 2 * t + 1;
                } else {
                    len[q] = half;
                    t = 2 * t;
                }
            }
        }
        bool more = true;
        while (more) {
            more = false;
            T v[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) v[q] = len[q] > 0 ? cdf[lo[q] + len[q] / 2] : (T)0;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (len[q] > 0) {
                    const int64_t half = len[q] / 2;
                    if (v[q] < u[q]) {
                        lo[q] += half + 1;
                        len[q] -= half + 1;
                    } else {
                        len[q] = half;
                    }
                    more |= len[q] > 0;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (i0 + q < k) samples[i0 + q] = lo[q];
    }
}

// The prefix sums in the reference's order, so they round exactly as its
// loop does: one workgroup streams the weights through LDS in blocks of
// CDF_BLK (double-buffered: waves 1-3 load block b + 1 while thread 0 adds
// up block b), thread 0 runs the sequential chain sum = sum + max(w[i], 0)
// over LDS, and the block's partial sums are written back coalesced.
// status: 0 ok (total[0] = 1 / sum); 1 + i: w[i] < error_if_below (w[0..i)
// hold the partial sums, as when the reference throws there); -1: the total
// is below sqrt(n) * eps (w holds the partial sums, unscaled).
constexpr int CDF_BLK = 2048;
template <typename T>
__global__ __launch_bounds__(256) void weights_to_cdf_kernel(int64_t n, T *w, T error_if_below, long long *status,
                                                             T *inv_total) {
    __shared__ __attribute__((aligned(16))) T buf[2][CDF_BLK];
    __shared__ int stop_sh;
    const int tid = threadIdx.x, wave = tid >> 6;
    const int64_t nb = (n + CDF_BLK - 1) / CDF_BLK;
    auto load = [&](int64_t b, T *dst, int t0, int nt) {
        const int64_t i0 = b * CDF_BLK;
        for (int e = t0; e < CDF_BLK; e += nt)
            if (i0 + e < n) dst[e] = w[i0 + e];
    };
    if (nb > 0) load(0, buf[0], tid, 256);
    T sum = (T)0;  // thread 0's running sum
    for (int64_t b = 0; b < nb; ++b) {
        T *cur = buf[b & 1];
        const int64_t i0 = b * CDF_BLK;
        const int len = (int)(n - i0 < CDF_BLK ? n - i0 : CDF_BLK);
        if (tid == 0) stop_sh = len;
        __syncthreads();  // block b loaded; the previous write-back done
        for (int e = tid; e < len; e += 256)
            if (!(cur[e] >= error_if_below)) atomicMin(&stop_sh, e);
        __syncthreads();
        const int stop = stop_sh;  // entries [0, stop) of the block are summed
        // val = max(val, 0) (std::max order), in parallel, off the chain
        for (int e = tid; e < stop; e += 256)
            if (cur[e] < (T)0) cur[e] = (T)0;
        __syncthreads();
        if (wave != 0 && stop == len && b + 1 < nb) load(b + 1, buf[(b + 1) & 1], tid - 64, 192);
        if (tid == 0) {
            // the chain alone: 16-byte LDS reads and writes; the next group
            // of 8 vectors is read before this group's sums are formed, so
            // the chain never waits on LDS latency (the read position is
            // clamped inside the block buffer; values past `stop` are unused)
            using V = std::conditional_t<sizeof(T) == 8, double2, float4>;
            constexpr int PV = 16 / sizeof(T);  // values per vector
            constexpr int GV = 8 * PV;          // values per group
            T s_ = sum;
            int e = 0;
            V v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = *(const V *)(cur + u * PV);
            for (; e + GV <= stop; e += GV) {
                const int ne = e + GV < CDF_BLK - GV ? e + GV : CDF_BLK - GV;
                V nx[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) nx[u] = *(const V *)(cur + ne + u * PV);
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    T *x = reinterpret_cast<T *>(&v[u]);
#pragma unroll
                    for (int q = 0; q < PV; ++q) {
                        s_ = s_ + x[q];
                        x[q] = s_;
                    }
                    *(V *)(cur + e + u * PV) = v[u];
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) v[u] = nx[u];
            }
            for (; e < stop; ++e) {
                s_ = s_ + cur[e];
                cur[e] = s_;
            }
            sum = s_;
        }
        __syncthreads();
        for (int e = tid; e < stop; e += 256) w[i0 + e] = cur[e];
        if (stop < len) {
            if (tid == 0) *status = 1 + i0 + stop;
            return;
        }
    }
    if (tid == 0) {
        const T floor_sum = (T)sqrt((double)n) * std::numeric_limits<T>::epsilon();
        if (!(sum >= floor_sum)) {
            *status = -1;
        } else {
            *inv_total = (T)1 / sum;
            *status = 0;
        }
    }
}

// w[i] = w[i] * (1 / sum) (blas::scal), when the sums succeeded
template <typename T>
__global__ __launch_bounds__(256) void cdf_scale_kernel(int64_t n, T *w, const long long *status,
                                                        const T *inv_total) {
    if (*status != 0) return;
    const T inv = *inv_total;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        w[i] = w[i] * inv;
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
        int64_t blocks = (counters + 255) / 256;
        if (blocks > 4 * (int64_t)num_cus()) blocks = 4 * (int64_t)num_cus();
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
    long long *status = (long long *)scratch(SCRATCH_SMALL, 2 * sizeof(long long), st);
    T *inv_total = (T *)(status + 1);
    hipLaunchKernelGGL(i64_zero_kernel, dim3(1), dim3(1), 0, st, status);
    hipLaunchKernelGGL((weights_to_cdf_kernel<T>), dim3(1), dim3(256), 0, st, n, w, error_if_below, status,
                       inv_total);
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks > 0)
        hipLaunchKernelGGL((cdf_scale_kernel<T>), dim3((unsigned)blocks), dim3(256), 0, st, n, w,
                           (const long long *)status, (const T *)inv_total);
    RBX_HIP(hipGetLastError());
    long long h = 0;
    RBX_HIP(hipMemcpyAsync(&h, status, sizeof(h), hipMemcpyDeviceToHost, st));
    RBX_HIP(hipStreamSynchronize(st));
    if (h > 0)
        RBX_FAIL(RBX_INVALID_ARG, "weights_to_cdf: w[%lld] is below error_if_below = %e", h - 1,
                 (double)error_if_below);
    if (h < 0) RBX_FAIL(RBX_INVALID_ARG, "weights_to_cdf: the weights sum to less than sqrt(n) * epsilon");
}

// ---------------------------------------------------------------- power-law sparse test data
// A random m x n CSR matrix whose column popularity follows a Zipf law (no
// RandBLAS counterpart; RandBLAS's own generators, sparse_data/
// random_matrix.hh:136-358, draw uniform-density matrices from one sequential
// stream).  Built from the library's counter-addressed samplers, so every
// stage is a pure function of (key, counter) and runs in parallel:
//   w[c]  = (c + 1)^(-1 / (exponent - 1))    host libm pow, rounded to T
//   cdf   = weights_to_cdf(w)                 (util.hh:459-472)
//   cols  = sample_indices_iid(n, cdf, k)     from `state`      (util.hh:493-514)
//   rows  = sample_indices_iid_uniform(m, k)  from the state after it (util.hh:516-546)
//   the k (row, column) pairs sorted, repeated pairs kept once -> CSR
//   vals  = the first nnz entries of a DenseDist(nnz, 1) Gaussian operator
//           from the state after the row draws (dense_skops.hh:563-606)
// Row degrees are Binomial(k, 1/m) before merging; the column degrees have a
// power-law tail with the given exponent.

// keys[i] = rows[i] * n + cols[i]
__global__ __launch_bounds__(256) void plc_keys_kernel(const int64_t *rows, const int64_t *cols, int64_t k,
                                                       int64_t n, uint64_t *keys) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < k) keys[i] = (uint64_t)rows[i] * (uint64_t)n + (uint64_t)cols[i];
}

// first[i] = 1 for the first of each run of equal sorted keys
__global__ __launch_bounds__(256) void plc_first_kernel(const uint64_t *keys, int64_t k, int64_t *first) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < k) first[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}

// Unique entry e = pos[i] (sorted key i opening a run): colidxs[e], and
// rowptr[r] = e for every row r in (row of the previous key, row of key i].
// Rows after the last key's row get nnz (the tail kernel).
__global__ __launch_bounds__(256) void plc_compact_kernel(const uint64_t *keys, const int64_t *first,
                                                          const int64_t *pos, int64_t k, int64_t n,
                                                          int64_t *colidxs, int64_t *rowptr) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= k || first[i] == 0) return;
    const int64_t e = pos[i];
    const uint64_t key = keys[i];
    const int64_t r = (int64_t)(key / (uint64_t)n);
    colidxs[e] = (int64_t)(key - (uint64_t)r * (uint64_t)n);
    const int64_t rp = i == 0 ? -1 : (int64_t)(keys[i - 1] / (uint64_t)n);
    for (int64_t q = rp + 1; q <= r; ++q) rowptr[q] = e;
}

__global__ __launch_bounds__(256) void plc_tail_kernel(const uint64_t *keys, int64_t k, int64_t n, int64_t m,
                                                       const int64_t *first, const int64_t *pos, int64_t *rowptr,
                                                       int64_t *nnz_out) {
    const int64_t r_last = k > 0 ? (int64_t)(keys[k - 1] / (uint64_t)n) : -1;
    const int64_t nnz = k > 0 ? pos[k - 1] + first[k - 1] : 0;
    for (int64_t q = r_last + 1 + (int64_t)threadIdx.x; q <= m; q += blockDim.x) rowptr[q] = nnz;
    if (threadIdx.x == 0) *nnz_out = nnz;
}

static int plc_bits(uint64_t x) {
    int b = 1;
    while (b < 64 && (x >> b) != 0) ++b;
    return b;
}

template <typename T>
static void power_law_csr_impl(int64_t m, int64_t n, int64_t k, double exponent, const rbx_rng_state *state,
                               rbx_csr_matrix *A, rbx_rng_state *next, hipStream_t st) {
    RBX_REQUIRE(state != nullptr && A != nullptr);
    RBX_REQUIRE(m > 0 && n > 0 && k >= 0);
    RBX_REQUIRE(exponent > 1.0);
    RBX_REQUIRE(A->rowptr != nullptr);
    RBX_REQUIRE(k == 0 || (A->colidxs != nullptr && A->vals != nullptr));
    const uint64_t mn = (uint64_t)safe_mul(m, n);
    RBX_REQUIRE(k < (1ll << 31) && mn < (1ull << 63));
    // column weights on the host (libm pow), then the device CDF
    std::vector<T> w((size_t)n);
    const double rank_exp = -1.0 / (exponent - 1.0);
    for (int64_t c = 0; c < n; ++c) w[(size_t)c] = (T)std::pow((double)(c + 1), rank_exp);
    const size_t b_cdf = ((size_t)n * sizeof(T) + 255) / 256 * 256, b_k = ((size_t)(k > 0 ? k : 1) * 8 + 255) / 256 * 256;
    size_t sort_bytes = 0, scan_bytes = 0;
    if (k > 0) {
        RBX_HIP(rocprim::radix_sort_keys(nullptr, sort_bytes, (uint64_t *)nullptr, (uint64_t *)nullptr, (size_t)k, 0,
                                         plc_bits(mn - 1), st));
        RBX_HIP(rocprim::exclusive_scan(nullptr, scan_bytes, (int64_t *)nullptr, (int64_t *)nullptr, (int64_t)0,
                                        (size_t)k, rocprim::plus<int64_t>(), st));
    }
    const size_t b_tmp = ((sort_bytes > scan_bytes ? sort_bytes : scan_bytes) + 255) / 256 * 256;
    char *ws = (char *)scratch(SCRATCH_SPARSE_OP, b_cdf + 5 * b_k + b_tmp + 256, st);
    T *cdf = (T *)ws;
    int64_t *rows = (int64_t *)(ws + b_cdf);
    uint64_t *keys = (uint64_t *)(ws + b_cdf + b_k), *keys2 = (uint64_t *)(ws + b_cdf + 2 * b_k);
    int64_t *first = (int64_t *)(ws + b_cdf + 3 * b_k), *pos = (int64_t *)(ws + b_cdf + 4 * b_k);
    void *tmp = ws + b_cdf + 5 * b_k;
    int64_t *d_nnz = (int64_t *)(ws + b_cdf + 5 * b_k + b_tmp);
    RBX_HIP(hipMemcpyAsync(cdf, w.data(), (size_t)n * sizeof(T), hipMemcpyHostToDevice, st));
    weights_to_cdf_impl<T>(n, cdf, (T)-1, st);  // synchronises the stream (w may go)

    rbx_rng_state s1, s2;
    int64_t *cols = A->colidxs;  // the column draws, overwritten by the merged columns below
    iid_cdf_impl<T>(n, cdf, k, cols, state, &s1, st);
    iid_uniform_impl<T>(m, k, rows, nullptr, &s1, &s2, st);
    const int64_t blocks = (k + 255) / 256;
    if (k > 0) {
        hipLaunchKernelGGL(plc_keys_kernel, dim3((unsigned)blocks), dim3(256), 0, st, rows, cols, k, n, keys);
        RBX_HIP(hipGetLastError());
        RBX_HIP(rocprim::radix_sort_keys(tmp, sort_bytes, keys, keys2, (size_t)k, 0, plc_bits(mn - 1), st));
        hipLaunchKernelGGL(plc_first_kernel, dim3((unsigned)blocks), dim3(256), 0, st, keys2, k, first);
        RBX_HIP(hipGetLastError());
        RBX_HIP(rocprim::exclusive_scan(tmp, scan_bytes, first, pos, (int64_t)0, (size_t)k, rocprim::plus<int64_t>(),
                                        st));
        hipLaunchKernelGGL(plc_compact_kernel, dim3((unsigned)blocks), dim3(256), 0, st, keys2, first, pos, k, n,
                           A->colidxs, A->rowptr);
        RBX_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(plc_tail_kernel, dim3(1), dim3(256), 0, st, keys2, k, n, m, first, pos, A->rowptr, d_nnz);
    RBX_HIP(hipGetLastError());
    int64_t nnz = 0;
    RBX_HIP(hipMemcpyAsync(&nnz, d_nnz, sizeof(nnz), hipMemcpyDeviceToHost, st));
    RBX_HIP(hipStreamSynchronize(st));
    // values: DenseDist(nnz, 1) Gaussian from s2 (natural layout ColMajor:
    // entry i is word i % 4 of counter s2 + i / 4); next = s2 + ceil(nnz / 4)
    rbx_rng_state s3 = s2;
    if (nnz > 0) {
        rbx_dense_skop V{};
        V.dist.n_rows = nnz;
        V.dist.n_cols = 1;
        V.dist.family = 'G';
        V.dist.major_axis = 'L';
        V.seed_state = s2;
        const GenMap g = gen_map(V, 'N', 0, 0, false);
        fill_generated<T>(g, nnz, 1, (T *)A->vals, 1, nnz, st);
        next_state(&s2, (uint64_t)((nnz + 3) / 4), &s3);
    }
    A->n_rows = m;
    A->n_cols = n;
    A->nnz = nnz;
    if (next) *next = s3;
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

rbx_status rbx_spower_law_csr(int64_t m, int64_t n, int64_t k, double exponent, const rbx_rng_state *state,
                              rbx_csr_matrix *A, rbx_rng_state *next_state_out, void *stream) {
    RBX_API_BEGIN
    power_law_csr_impl<float>(m, n, k, exponent, state, A, next_state_out, (hipStream_t)stream);
    RBX_API_END
}

rbx_status rbx_dpower_law_csr(int64_t m, int64_t n, int64_t k, double exponent, const rbx_rng_state *state,
                              rbx_csr_matrix *A, rbx_rng_state *next_state_out, void *stream) {
    RBX_API_BEGIN
    power_law_csr_impl<double>(m, n, k, exponent, state, A, next_state_out, (hipStream_t)stream);
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
