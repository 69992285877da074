// rbx_convert.hip -- sparse-data format conversions on the device
// (sparse_data/conversions.hh:49-126, sparse_data/base.hh:140-300):
//   csr_to_coo / csc_to_coo      copy the stored arrays, expand the pointers
//   coo_to_csr / coo_to_csc      stable sort by (row, column) or (column, row),
//                                then compress the major indices
//   coo sort / determine sort    coo_arrays_apply_sort / coo_arrays_determine_sort
// Output arrays are allocated by the caller (device memory).  Sorting is
// stable: entries with equal (row, column) keep their stored order (the
// reference sorts with std::sort, for which that order is unspecified, and
// keeps it when the arrays are already in the target order).
#include <cstring>

#include "rbx_internal.hh"

#include <rocprim/rocprim.hpp>

namespace rbx {

// idx[p] = i for every p in [ptr[i], ptr[i+1]) (compressed_ptr_to_sorted_idxs);
// positions outside [0, nnz) are never written
__global__ __launch_bounds__(256) void expand_ptr_kernel(const int64_t *ptr, int64_t ncomp, int64_t nnz,
                                                         int64_t *idx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncomp) return;
    const int64_t p0 = ptr[i] < 0 ? 0 : ptr[i], p1 = ptr[i + 1] > nnz ? nnz : ptr[i + 1];
    for (int64_t p = p0; p < p1; ++p) idx[p] = i;
}

// ptr from sorted major indices (sorted_idxs_to_compressed_ptr): ptr[q] is
// the first position whose index is >= q.
__global__ __launch_bounds__(256) void compress_idx_kernel(const int64_t *idx, int64_t nnz, int64_t ncomp,
                                                           int64_t *ptr) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p > nnz) return;
    int64_t lo = p == 0 ? 0 : idx[p - 1] + 1;      // first slot this position owns
    const int64_t hi = p == nnz ? ncomp : idx[p];  // last slot it owns
    if (lo < 0) lo = 0;                            // (indices are in range for valid input;
    for (int64_t q = lo; q <= hi && q <= ncomp; ++q) ptr[q] = p;  //  no slot outside [0, ncomp] is written)
}

__global__ __launch_bounds__(256) void sort_keys_kernel(const int64_t *maj, const int64_t *mnr, int64_t nnz,
                                                        int64_t nmnr, uint64_t *keys, int64_t *perm) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x) {
        keys[e] = (uint64_t)maj[e] * (uint64_t)nmnr + (uint64_t)mnr[e];
        perm[e] = e;
    }
}

__global__ __launch_bounds__(256) void major_keys_kernel(const int64_t *maj, int64_t nnz, uint32_t *keys,
                                                         int32_t *perm) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x) {
        keys[e] = (uint32_t)maj[e];
        perm[e] = (int32_t)e;
    }
}

template <typename T, typename PT>
__global__ __launch_bounds__(256) void gather3_kernel(const PT *perm, int64_t nnz, const T *v, const int64_t *a,
                                                      const int64_t *b, T *vo, int64_t *ao, int64_t *bo) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t s = perm[e];
        vo[e] = v[s];
        ao[e] = a[s];
        bo[e] = b[s];
    }
}

// bit 0: not CSC-ordered, bit 1: not CSR-ordered (increasing_by_csc / _csr,
// sparse_data/base.hh:120-139: non-decreasing in the major index, then in
// the minor one)
__global__ __launch_bounds__(256) void order_flags_kernel(const int64_t *rows, const int64_t *cols, int64_t nnz,
                                                          int *flags) {
    int bad = 0;
    for (int64_t e = 1 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i0 = rows[e - 1], j0 = cols[e - 1], i1 = rows[e], j1 = cols[e];
        if (!(j0 < j1 || (j0 == j1 && i0 <= i1))) bad |= 1;
        if (!(i0 < i1 || (i0 == i1 && j0 <= j1))) bad |= 2;
    }
    // one atomic per wave (one per thread serialises millions of them on the
    // flag's cache line)
    const int wbad = (__ballot(bad & 1) ? 1 : 0) | (__ballot(bad & 2) ? 2 : 0);
    if ((threadIdx.x & 63) == 0 && wbad) atomicOr(flags, wbad);
}

__global__ void int_zero_kernel(int *p) { *p = 0; }

static int bits_of(uint64_t x) {
    int b = 1;
    while (b < 64 && (x >> b) != 0) ++b;
    return b;
}

static int64_t blocks_for(int64_t n) {
    int64_t b = (n + 255) / 256;
    if (b < 1) b = 1;
    return b > 16384 ? 16384 : b;
}

// How the entries already lie: 0 ordered by (maj, mnr), 1 by (mnr, maj), 2
// neither (one pass over the indices and a 4-byte read on the host: the
// reference keeps this as COOMatrix::sort and skips its sort by it,
// conversions.hh:87-117).
static int coo_input_order(const int64_t *maj, const int64_t *mnr, int64_t nnz, hipStream_t st) {
    if (nnz < 2) return 0;
    int *flags = (int *)scratch(SCRATCH_SMALL, sizeof(int), st);
    hipLaunchKernelGGL(int_zero_kernel, dim3(1), dim3(1), 0, st, flags);
    int64_t nb = blocks_for(nnz);
    if (nb > 4 * (int64_t)num_cus()) nb = 4 * (int64_t)num_cus();
    hipLaunchKernelGGL(order_flags_kernel, dim3((unsigned)nb), dim3(256), 0, st, maj, mnr, nnz, flags);
    RBX_HIP(hipGetLastError());
    int h = 0;
    RBX_HIP(hipMemcpyAsync(&h, flags, sizeof(h), hipMemcpyDeviceToHost, st));
    RBX_HIP(hipStreamSynchronize(st));
    return !(h & 2) ? 0 : (!(h & 1) ? 1 : 2);
}

// Stable sort of the COO arrays (vals, maj, mnr) by (maj, mnr) into the
// output arrays (which may be the inputs).  Entries already in that order are
// copied; entries ordered by (mnr, maj) need a stable sort on maj alone
// (within a major index they already ascend in the minor one, duplicates in
// stored order), on 32-bit keys; others sort on the full key.
template <typename T>
static void sort_coo(int64_t nnz, int64_t nmaj, int64_t nmnr, const T *v, const int64_t *maj, const int64_t *mnr, T *vo,
                     int64_t *majo, int64_t *mnro, hipStream_t st) {
    if (nnz == 0) return;
    RBX_REQUIRE(nmaj > 0 && nmnr > 0);
    const int ord = coo_input_order(maj, mnr, nnz, st);
    if (ord == 0) {
        if (vo != v) RBX_HIP(hipMemcpyAsync(vo, v, (size_t)nnz * sizeof(T), hipMemcpyDeviceToDevice, st));
        if (majo != maj) RBX_HIP(hipMemcpyAsync(majo, maj, (size_t)nnz * 8, hipMemcpyDeviceToDevice, st));
        if (mnro != mnr) RBX_HIP(hipMemcpyAsync(mnro, mnr, (size_t)nnz * 8, hipMemcpyDeviceToDevice, st));
        return;
    }
    if (ord == 1 && (uint64_t)nmaj <= 0xFFFFFFFFull && nnz < (int64_t)INT32_MAX) {
        const int end_bit = bits_of((uint64_t)(nmaj - 1));
        size_t sort_bytes = 0;
        RBX_HIP(rocprim::radix_sort_pairs(nullptr, sort_bytes, (uint32_t *)nullptr, (uint32_t *)nullptr,
                                          (int32_t *)nullptr, (int32_t *)nullptr, (size_t)nnz, 0, end_bit, st));
        const size_t b4 = ((size_t)nnz * 4 + 255) / 256 * 256, b8 = ((size_t)nnz * 8 + 255) / 256 * 256,
                     bt = ((size_t)nnz * sizeof(T) + 255) / 256 * 256;
        char *ws = (char *)scratch(SCRATCH_SPARSE_CSR, 4 * b4 + 2 * b8 + bt + sort_bytes, st);
        uint32_t *k_in = (uint32_t *)ws, *k_out = (uint32_t *)(ws + b4);
        int32_t *p_in = (int32_t *)(ws + 2 * b4), *p_out = (int32_t *)(ws + 3 * b4);
        int64_t *a_tmp = (int64_t *)(ws + 4 * b4), *b_tmp = (int64_t *)(ws + 4 * b4 + b8);
        T *v_tmp = (T *)(ws + 4 * b4 + 2 * b8);
        void *sort_tmp = ws + 4 * b4 + 2 * b8 + bt;
        const unsigned nb = (unsigned)blocks_for(nnz);
        hipLaunchKernelGGL(major_keys_kernel, dim3(nb), dim3(256), 0, st, maj, nnz, k_in, p_in);
        RBX_HIP(hipGetLastError());
        RBX_HIP(rocprim::radix_sort_pairs(sort_tmp, sort_bytes, k_in, k_out, p_in, p_out, (size_t)nnz, 0, end_bit, st));
        hipLaunchKernelGGL((gather3_kernel<T, int32_t>), dim3(nb), dim3(256), 0, st, p_out, nnz, v, maj, mnr, v_tmp,
                           a_tmp, b_tmp);
        RBX_HIP(hipGetLastError());
        RBX_HIP(hipMemcpyAsync(vo, v_tmp, (size_t)nnz * sizeof(T), hipMemcpyDeviceToDevice, st));
        RBX_HIP(hipMemcpyAsync(majo, a_tmp, (size_t)nnz * 8, hipMemcpyDeviceToDevice, st));
        RBX_HIP(hipMemcpyAsync(mnro, b_tmp, (size_t)nnz * 8, hipMemcpyDeviceToDevice, st));
        return;
    }
    const uint64_t span = (uint64_t)safe_mul(nmaj, nmnr);
    const int end_bit = bits_of(span - 1);
    size_t sort_bytes = 0;
    RBX_HIP(rocprim::radix_sort_pairs(nullptr, sort_bytes, (uint64_t *)nullptr, (uint64_t *)nullptr,
                                      (int64_t *)nullptr, (int64_t *)nullptr, (size_t)nnz, 0, end_bit, st));
    const size_t b8 = ((size_t)nnz * 8 + 255) / 256 * 256, bt = ((size_t)nnz * sizeof(T) + 255) / 256 * 256;
    char *ws = (char *)scratch(SCRATCH_SPARSE_CSR, 6 * b8 + bt + sort_bytes, st);
    uint64_t *k_in = (uint64_t *)ws, *k_out = (uint64_t *)(ws + b8);
    int64_t *p_in = (int64_t *)(ws + 2 * b8), *p_out = (int64_t *)(ws + 3 * b8);
    int64_t *a_tmp = (int64_t *)(ws + 4 * b8), *b_tmp = (int64_t *)(ws + 5 * b8);
    T *v_tmp = (T *)(ws + 6 * b8);
    void *sort_tmp = ws + 6 * b8 + bt;
    const unsigned nb = (unsigned)blocks_for(nnz);
    hipLaunchKernelGGL(sort_keys_kernel, dim3(nb), dim3(256), 0, st, maj, mnr, nnz, nmnr, k_in, p_in);
    RBX_HIP(hipGetLastError());
    RBX_HIP(rocprim::radix_sort_pairs(sort_tmp, sort_bytes, k_in, k_out, p_in, p_out, (size_t)nnz, 0, end_bit, st));
    // gather through scratch so that the output may alias the input
    hipLaunchKernelGGL((gather3_kernel<T, int64_t>), dim3(nb), dim3(256), 0, st, p_out, nnz, v, maj, mnr, v_tmp, a_tmp,
                       b_tmp);
    RBX_HIP(hipGetLastError());
    RBX_HIP(hipMemcpyAsync(vo, v_tmp, (size_t)nnz * sizeof(T), hipMemcpyDeviceToDevice, st));
    RBX_HIP(hipMemcpyAsync(majo, a_tmp, (size_t)nnz * 8, hipMemcpyDeviceToDevice, st));
    RBX_HIP(hipMemcpyAsync(mnro, b_tmp, (size_t)nnz * 8, hipMemcpyDeviceToDevice, st));
}

template <typename T>
static void compressed_to_coo(int64_t ncomp, int64_t nnz, const int64_t *ptr, const int64_t *minor, const T *vals,
                              int64_t *maj_out, int64_t *mnr_out, T *vals_out, hipStream_t st) {
    if (nnz == 0) return;
    RBX_HIP(hipMemcpyAsync(vals_out, vals, (size_t)nnz * sizeof(T), hipMemcpyDeviceToDevice, st));
    RBX_HIP(hipMemcpyAsync(mnr_out, minor, (size_t)nnz * 8, hipMemcpyDeviceToDevice, st));
    if (ncomp > 0) {
        hipLaunchKernelGGL(expand_ptr_kernel, dim3((unsigned)((ncomp + 255) / 256)), dim3(256), 0, st, ptr, ncomp,
                           nnz, maj_out);
        RBX_HIP(hipGetLastError());
    }
}

// COO -> compressed: sort by (maj, mnr) into the output index / value
// arrays, expanding the sorted major indices into scratch, then compress.
template <typename T>
static void coo_to_compressed(int64_t nmaj, int64_t nmnr, int64_t nnz, const int64_t *maj, const int64_t *mnr,
                              const T *vals, int64_t *ptr, int64_t *mnr_out, T *vals_out, hipStream_t st) {
    RBX_REQUIRE(ptr != nullptr);
    int64_t *maj_sorted = (int64_t *)scratch(SCRATCH_SPARSE_OP, (size_t)(nnz > 0 ? nnz : 1) * 8, st);
    if (nnz > 0) sort_coo<T>(nnz, nmaj, nmnr, vals, maj, mnr, vals_out, maj_sorted, mnr_out, st);
    hipLaunchKernelGGL(compress_idx_kernel, dim3((unsigned)((nnz + 1 + 255) / 256)), dim3(256), 0, st, maj_sorted, nnz,
                       nmaj, ptr);
    RBX_HIP(hipGetLastError());
}

static void check_coo(const rbx_coo_matrix *A) {
    RBX_REQUIRE(A != nullptr);
    RBX_REQUIRE(A->n_rows >= 0 && A->n_cols >= 0 && A->nnz >= 0);
    RBX_REQUIRE(A->nnz == 0 || (A->vals && A->rows && A->cols));
}

}  // namespace rbx

using namespace rbx;

extern "C" {

#define RBX_CONVERT_ENTRIES(PFX, T)                                                                                  \
    rbx_status rbx_##PFX##csr_to_coo(const rbx_csr_matrix *csr, rbx_coo_matrix *coo, void *stream) {                 \
        RBX_API_BEGIN                                                                                                \
        RBX_REQUIRE(csr != nullptr && coo != nullptr && csr->rowptr != nullptr);                                     \
        RBX_REQUIRE(csr->n_rows == coo->n_rows && csr->n_cols == coo->n_cols && csr->nnz >= 0);                      \
        coo->nnz = csr->nnz;                                                                                         \
        check_coo(coo);                                                                                              \
        compressed_to_coo<T>(csr->n_rows, csr->nnz, csr->rowptr, csr->colidxs, (const T *)csr->vals, coo->rows,     \
                             coo->cols, (T *)coo->vals, (hipStream_t)stream);                                        \
        RBX_API_END                                                                                                  \
    }                                                                                                                \
    rbx_status rbx_##PFX##csc_to_coo(const rbx_csc_matrix *csc, rbx_coo_matrix *coo, void *stream) {                 \
        RBX_API_BEGIN                                                                                                \
        RBX_REQUIRE(csc != nullptr && coo != nullptr && csc->colptr != nullptr);                                     \
        RBX_REQUIRE(csc->n_rows == coo->n_rows && csc->n_cols == coo->n_cols && csc->nnz >= 0);                      \
        coo->nnz = csc->nnz;                                                                                         \
        check_coo(coo);                                                                                              \
        compressed_to_coo<T>(csc->n_cols, csc->nnz, csc->colptr, csc->rowidxs, (const T *)csc->vals, coo->cols,     \
                             coo->rows, (T *)coo->vals, (hipStream_t)stream);                                        \
        RBX_API_END                                                                                                  \
    }                                                                                                                \
    rbx_status rbx_##PFX##coo_to_csr(const rbx_coo_matrix *coo, rbx_csr_matrix *csr, void *stream) {                 \
        RBX_API_BEGIN                                                                                                \
        check_coo(coo);                                                                                              \
        RBX_REQUIRE(csr != nullptr && csr->n_rows == coo->n_rows && csr->n_cols == coo->n_cols);                     \
        RBX_REQUIRE(coo->nnz == 0 || (csr->vals && csr->colidxs));                                                   \
        csr->nnz = coo->nnz;                                                                                         \
        coo_to_compressed<T>(coo->n_rows, coo->n_cols, coo->nnz, coo->rows, coo->cols, (const T *)coo->vals,        \
                             csr->rowptr, csr->colidxs, (T *)csr->vals, (hipStream_t)stream);                        \
        RBX_API_END                                                                                                  \
    }                                                                                                                \
    rbx_status rbx_##PFX##coo_to_csc(const rbx_coo_matrix *coo, rbx_csc_matrix *csc, void *stream) {                 \
        RBX_API_BEGIN                                                                                                \
        check_coo(coo);                                                                                              \
        RBX_REQUIRE(csc != nullptr && csc->n_rows == coo->n_rows && csc->n_cols == coo->n_cols);                     \
        RBX_REQUIRE(coo->nnz == 0 || (csc->vals && csc->rowidxs));                                                   \
        csc->nnz = coo->nnz;                                                                                         \
        coo_to_compressed<T>(coo->n_cols, coo->n_rows, coo->nnz, coo->cols, coo->rows, (const T *)coo->vals,        \
                             csc->colptr, csc->rowidxs, (T *)csc->vals, (hipStream_t)stream);                        \
        RBX_API_END                                                                                                  \
    }                                                                                                                \
    rbx_status rbx_##PFX##coo_sort(char order, rbx_coo_matrix *coo, void *stream) {                                  \
        RBX_API_BEGIN                                                                                                \
        check_coo(coo);                                                                                              \
        if (order != 'R' && order != 'C') RBX_FAIL(RBX_INVALID_ARG, "coo_sort: order must be 'R' (CSR) or 'C' (CSC)"); \
        if (order == 'R')                                                                                            \
            sort_coo<T>(coo->nnz, coo->n_rows, coo->n_cols, (const T *)coo->vals, coo->rows, coo->cols,             \
                        (T *)coo->vals, coo->rows, coo->cols, (hipStream_t)stream);                                  \
        else                                                                                                         \
            sort_coo<T>(coo->nnz, coo->n_cols, coo->n_rows, (const T *)coo->vals, coo->cols, coo->rows,             \
                        (T *)coo->vals, coo->cols, coo->rows, (hipStream_t)stream);                                  \
        RBX_API_END                                                                                                  \
    }

RBX_CONVERT_ENTRIES(s, float)
RBX_CONVERT_ENTRIES(d, double)

rbx_status rbx_coo_determine_sort(const rbx_coo_matrix *coo, char *order, void *stream) {
    RBX_API_BEGIN
    check_coo(coo);
    RBX_REQUIRE(order != nullptr);
    hipStream_t st = (hipStream_t)stream;
    int *flags = (int *)scratch(SCRATCH_SMALL, sizeof(int), st);
    hipLaunchKernelGGL(int_zero_kernel, dim3(1), dim3(1), 0, st, flags);
    if (coo->nnz > 1)
        hipLaunchKernelGGL(order_flags_kernel, dim3((unsigned)blocks_for(coo->nnz)), dim3(256), 0, st, coo->rows,
                           coo->cols, coo->nnz, flags);
    RBX_HIP(hipGetLastError());
    int h = 0;
    RBX_HIP(hipMemcpyAsync(&h, flags, sizeof(h), hipMemcpyDeviceToHost, st));
    RBX_HIP(hipStreamSynchronize(st));
    *order = !(h & 1) ? 'C' : (!(h & 2) ? 'R' : 'N');
    RBX_API_END
}

}  // extern "C"
