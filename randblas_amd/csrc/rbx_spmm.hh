// rbx_spmm.hh -- the sparse-operator apply path that reads the dense operand
// once (rbx_spmm_lds.hip), used by left_spmm_coo in rbx_sparse.hip.
#pragma once
#include "rbx_internal.hh"

namespace rbx {

// Summation forms of the reference's sparse kernels, per output entry c with
// stored entries (v, b) applied in order (beta = 0 starts from 0):
//   SPF_JKI  c = RN(beta*c); c = RN(c + RN(v*RN(alpha*b)))  (CSC column kernel,
//            csc_spmm_impl.hh:74-96)
//   SPF_KIB  c = RN(beta*c); c = RN(c + RN(RN(alpha*v)*b))  (row-major rank-1
//            updates, csc_spmm_impl.hh:154-209)
enum { SPF_JKI = 0, SPF_KIB = 1 };

// Most output rows the counting-sort builder handles (its per-row counters
// live in LDS).
constexpr int64_t CSRB_MAX_ROWS = 8192;

// Bytes of workspace csr_build_counting_ws needs (0 when the shape is
// outside what the builder takes).
size_t csr_build_counting_bytes(int64_t nnz, int64_t d, size_t val_size);
// The same build into a caller-provided workspace of ws_bytes bytes (every
// byte it reads is written first in the same call, so the workspace needs no
// clearing between calls).
template <typename T>
bool csr_build_counting_ws(const int64_t *rows, const int64_t *cols, const T *vals, int64_t nnz, int64_t ro,
                           int64_t co, int64_t d, int64_t m, char *ws, size_t ws_bytes, int64_t **row_ptr,
                           int32_t **col, T **val, hipStream_t st);

// COO entries inside the window [ro, ro + d) x [co, co + m) -> CSR sorted by
// (row, column, stored position): a stable counting sort by row, then rows
// whose columns are not ascending (entries not given in column order) sorted
// by (column, position).  Row pointers, 32-bit columns and values go to
// library scratch.  Returns false (nothing launched) when d > CSRB_MAX_ROWS,
// d == 0, or nnz / m do not fit 31 bits.
template <typename T>
bool csr_build_counting(const int64_t *rows, const int64_t *cols, const T *vals, int64_t nnz, int64_t ro, int64_t co,
                        int64_t d, int64_t m, int64_t **row_ptr, int32_t **col, T **val, hipStream_t st);

// C (d x n, element (i, j) at C[i*crs + j*ccs]) = alpha * A_csr * Bm + beta * C
// for a CSR A (d x m, nnz stored entries, columns ascending within each row)
// and a dense Bm (m x n) with unit stride along m (Bm[c, j] at B[c + j*ldb]).
// Each output entry is one sequential chain over its row's entries in stored
// order, in summation form `form`, so results equal the reference's bit for
// bit.  Bm is read once: a workgroup owns a block of output columns and 512
// output rows, and streams chunks of Bm's columns through LDS.  `ws` is a
// workspace of csr_lds_apply_ws_bytes(d, m, nnz, sizeof(T)) bytes (the
// entries laid out by chunk and slot; every byte read is written first in the
// same call).  Returns false when the shape is outside what the kernel takes
// (alpha == 0 and beta-only scaling are the caller's).
size_t csr_lds_apply_ws_bytes(int64_t d, int64_t m, int64_t nnz, size_t val_size);
// csr_lds_apply in two steps, for an operator applied more than once: the
// entry layout (chunk table and slots, depending on d, m and the entries only)
// into `ws`, then any number of applies that read it.  Each returns false,
// with nothing launched, when the shape is outside what the kernel takes.
// `unit` (the layout's: every value is exactly +1 or -1) lets the applies
// take each value's sign from its slot instead of loading the value.
template <typename T>
bool csr_lds_layout(const int64_t *row_ptr, const int32_t *col, const T *val, int64_t d, int64_t m, int64_t nnz,
                    char *ws, size_t ws_bytes, hipStream_t st, bool unit);
template <typename T>
bool csr_lds_apply_laid(const int64_t *row_ptr, const int32_t *col, const T *val, int64_t d, int64_t n, int64_t m,
                        int64_t nnz, T alpha, T beta, const T *B, int64_t ldb, T *C, int64_t crs, int64_t ccs,
                        int form, const char *ws, size_t ws_bytes, hipStream_t st, bool unit);
template <typename T>
bool csr_lds_apply(const int64_t *row_ptr, const int32_t *col, const T *val, int64_t d, int64_t n, int64_t m,
                   int64_t nnz, T alpha, T beta, const T *B, int64_t ldb, T *C, int64_t crs, int64_t ccs, int form,
                   char *ws, size_t ws_bytes, hipStream_t st, bool unit = false);

}  // namespace rbx
