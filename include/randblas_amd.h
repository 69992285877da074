/*
 * randblas_amd.h -- C ABI of the MI355X (gfx950) sketching library.
 *
 * This is the drop-in boundary for RandBLAS's sketching hot path.  RandBLAS
 * itself is a header-only C++ template library whose boundary is
 * namespace RandBLAS (RandBLAS.hh:33-41); every entry point below replaces one
 * template instantiation, named with the overload-free convention
 * [s|d][l|r]skge[3|s] / [s|d][l|r]sksp3 described in rtd/source/FAQ.rst:154-186.
 * A header-only C++ layer with the reference's own type and function names
 * (randblas_amd/cpp/RandBLAS.hh) forwards to these functions.
 *
 * Conventions
 *   - All matrix / array pointers are DEVICE pointers (hipMalloc'd or torch
 *     tensors on the GPU).  Descriptors (rbx_*_dist, rbx_*_skop, rbx_rng_state)
 *     are HOST structs passed by pointer.
 *   - layout is 'C' (blas::Layout::ColMajor) or 'R' (RowMajor);
 *     op is 'N' (blas::Op::NoTrans) or 'T' (Trans).
 *   - stream is a hipStream_t passed as void*; NULL = the default stream.
 *     Every call is stream-ordered and asynchronous unless stated otherwise.
 *   - No exceptions cross the ABI: every function returns an rbx_status and
 *     leaves a description in a thread-local slot read by
 *     rbx_last_error_message().  This replaces RandBLAS::Error and
 *     std::overflow_error (RandBLAS/exceptions.hh:57-95,152-162,
 *     RandBLAS/base.hh:260-274).
 *   - alpha == 0 leaves A unread; beta == 0 leaves B unread (BLAS semantics,
 *     RandBLAS/util.hh:55-62).
 *   - Results do not depend on the launch configuration: every operator entry
 *     is a pure function of (key, counter, row, column).
 */
#ifndef RANDBLAS_AMD_H
#define RANDBLAS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    RBX_OK = 0,
    RBX_INVALID_ARG = 1,  /* randblas_require / randblas_error_if failures */
    RBX_OVERFLOW = 2,     /* safe_int_product overflow (base.hh:260-274)   */
    RBX_HIP_ERROR = 3,    /* a HIP runtime call failed                      */
    RBX_UNSUPPORTED = 4,  /* valid in the reference, not offered here       */
    RBX_OUT_OF_MEMORY = 5 /* device allocation failed                       */
} rbx_status;

/* RNGState<Philox4x32> (RandBLAS/base.hh:64-164): a 128-bit counter in four
 * 32-bit words (word 0 least significant) and a 64-bit key in two words. */
typedef struct {
    uint32_t counter[4];
    uint32_t key[2];
} rbx_rng_state;

/* DenseDist (RandBLAS/dense_skops.hh:231-350).  family 'G' (Gaussian) or
 * 'U' (uniform on [-sqrt3, sqrt3]); major_axis 'L' (Long) or 'S' (Short). */
typedef struct {
    int64_t n_rows;
    int64_t n_cols;
    char family;
    char major_axis;
} rbx_dense_dist;

/* DenseSkOp (RandBLAS/dense_skops.hh:357-478).  buff == NULL means the operator
 * is never materialised: kernels regenerate its entries from (key, counter)
 * inside the GEMM tiles.  A non-NULL buff is a device buffer in `layout`
 * holding the full n_rows x n_cols operator. */
typedef struct {
    rbx_dense_dist dist;
    rbx_rng_state seed_state;
    rbx_rng_state next_state;
    void *buff;
    char layout;
    int own_memory; /* 1: buff was allocated by rbx_[sd]fill_dense_skop */
} rbx_dense_skop;

/* SparseDist (RandBLAS/sparse_skops.hh:131-246).  major_axis 'S' (SASO) or
 * 'L' (LASO). */
typedef struct {
    int64_t n_rows;
    int64_t n_cols;
    int64_t vec_nnz;
    char major_axis;
} rbx_sparse_dist;

/* SparseSkOp<T, Philox4x32, int64_t> (RandBLAS/sparse_skops.hh:289-450).
 * nnz < 0 means "not sampled yet": the sketch entry points then sample it
 * into library scratch for the duration of the call (skge.hh:483-487).
 * rows / cols / vals are device arrays of length >= dist.vec_nnz * dim_minor. */
typedef struct {
    rbx_sparse_dist dist;
    rbx_rng_state seed_state;
    rbx_rng_state next_state;
    int64_t nnz;
    void *vals;
    int64_t *rows;
    int64_t *cols;
    int own_memory; /* 1: arrays were allocated by rbx_[sd]fill_sparse_skop */
    /* Library-owned prepared copy of the sampled operator (its row-sorted CSR
     * and the LDS apply's entry layout), or NULL.  No reference counterpart.
     * Set by rbx_[sd]fill_sparse_skop and rbx_[sd]sparse_skop_prepare, freed by
     * rbx_sparse_skop_free; initialise the struct with rbx_sparse_skop_init (or
     * zero it) so that this field starts NULL. */
    void *prep;
} rbx_sparse_skop;

/* CSRMatrix<T, int64_t> (RandBLAS/sparse_data/csr_matrix.hh:73-245), zero-based,
 * device arrays. */
typedef struct {
    int64_t n_rows;
    int64_t n_cols;
    int64_t nnz;
    void *vals;
    int64_t *rowptr;
    int64_t *colidxs;
} rbx_csr_matrix;

/* CSCMatrix<T, int64_t> (RandBLAS/sparse_data/csc_matrix.hh:74-244), zero-based,
 * device arrays. */
typedef struct {
    int64_t n_rows;
    int64_t n_cols;
    int64_t nnz;
    void *vals;
    int64_t *rowidxs;
    int64_t *colptr;
} rbx_csc_matrix;

/* COOMatrix<T, int64_t> (RandBLAS/sparse_data/coo_matrix.hh:83-308), zero-based,
 * device arrays in any order (entries need not be sorted). */
typedef struct {
    int64_t n_rows;
    int64_t n_cols;
    int64_t nnz;
    void *vals;
    int64_t *rows;
    int64_t *cols;
} rbx_coo_matrix;

/* ---------------------------------------------------------------- errors */
const char *rbx_last_error_message(void);
/* Version string of this library. */
const char *rbx_version(void);

/* ---------------------------------------------------------------- RNG
 * Raw Philox4x32-10 words for counters state.counter + i, i in [0, n_counters):
 * out[4*i + w] = word w.  Replaces r123::Philox4x32::operator() as called at
 * RandBLAS/random_gen.hh:107,135. */
rbx_status rbx_philox4x32_10(const rbx_rng_state *state, int64_t n_counters, uint32_t *out,
                             void *stream);

/* Host-side helpers (synchronous, no device work).
 * RNGState(uint64 k) (base.hh:119): counter 0, key {k mod 2^32, k >> 32}. */
rbx_status rbx_rng_state_from_key(uint64_t key, rbx_rng_state *out);
/* counter.incr(n) (Random123 r123array::incr; semantics pinned by
 * test/test_basic_rng/test_r123.cc:706-797). */
rbx_status rbx_rng_state_incr(rbx_rng_state *state, uint64_t n);

/* ---------------------------------------------------------------- dense operators
 * DenseSkOp(dist, seed_state) (dense_skops.hh:396-406): fills in next_state via
 * dense::compute_next_state (dense_skops.hh:172-185); buff = NULL. */
rbx_status rbx_dense_skop_init(const rbx_dense_dist *dist, const rbx_rng_state *seed,
                               rbx_dense_skop *out);
/* dim_major / dim_minor / natural layout / isometry scale of a DenseDist
 * (dense_skops.hh:313-329, 187-199). */
rbx_status rbx_dense_dist_props(const rbx_dense_dist *dist, int64_t *dim_major,
                                int64_t *dim_minor, char *natural_layout,
                                double *isometry_scale);

/* fill_dense_unpacked (dense_skops.hh:563-606): writes the n_rows x n_cols
 * submatrix at (ro_s, co_s) of an operator drawn from D into buff (compact,
 * in `layout`) and returns the state the reference returns. */
rbx_status rbx_sfill_dense_unpacked(char layout, const rbx_dense_dist *D, int64_t n_rows,
                                    int64_t n_cols, int64_t ro_s, int64_t co_s, float *buff,
                                    const rbx_rng_state *seed, rbx_rng_state *next_state,
                                    void *stream);
rbx_status rbx_dfill_dense_unpacked(char layout, const rbx_dense_dist *D, int64_t n_rows,
                                    int64_t n_cols, int64_t ro_s, int64_t co_s, double *buff,
                                    const rbx_rng_state *seed, rbx_rng_state *next_state,
                                    void *stream);
/* fill_dense(S) (dense_skops.hh:649-658): allocates S->buff (device) if it is
 * NULL and marks the operator as owning it; fills in S->layout. */
rbx_status rbx_sfill_dense_skop(rbx_dense_skop *S, void *stream);
rbx_status rbx_dfill_dense_skop(rbx_dense_skop *S, void *stream);
/* ~DenseSkOp (dense_skops.hh:472-477): frees buff if owned. */
rbx_status rbx_dense_skop_free(rbx_dense_skop *S);

/* ---------------------------------------------------------------- dense sketches
 * lskge3 (skge.hh:154-202):  B = alpha * op(submat(S)) * op(A) + beta * B,
 *   B is d x n, op(A) is m x n, op(submat(S)) is d x m taken at (ro_s, co_s).
 * rskge3 (skge.hh:307-355):  B = alpha * op(A) * op(submat(S)) + beta * B,
 *   B is m x d, op(A) is m x n, op(submat(S)) is n x d.
 * With S->buff == NULL the operator is generated inside the GEMM tiles and
 * never written to memory (the reference materialises it, skge.hh:175-181). */
rbx_status rbx_slskge3(char layout, char opS, char opA, int64_t d, int64_t n, int64_t m,
                       float alpha, const rbx_dense_skop *S, int64_t ro_s, int64_t co_s,
                       const float *A, int64_t lda, float beta, float *B, int64_t ldb,
                       void *stream);
rbx_status rbx_dlskge3(char layout, char opS, char opA, int64_t d, int64_t n, int64_t m,
                       double alpha, const rbx_dense_skop *S, int64_t ro_s, int64_t co_s,
                       const double *A, int64_t lda, double beta, double *B, int64_t ldb,
                       void *stream);
rbx_status rbx_srskge3(char layout, char opA, char opS, int64_t m, int64_t d, int64_t n,
                       float alpha, const float *A, int64_t lda, const rbx_dense_skop *S,
                       int64_t ro_s, int64_t co_s, float beta, float *B, int64_t ldb,
                       void *stream);
rbx_status rbx_drskge3(char layout, char opA, char opS, int64_t m, int64_t d, int64_t n,
                       double alpha, const double *A, int64_t lda, const rbx_dense_skop *S,
                       int64_t ro_s, int64_t co_s, double beta, double *B, int64_t ldb,
                       void *stream);

/* Adjoints (no reference counterpart as a separate routine; the reference
 * obtains them as lskge3/rskge3 with the operator transposed, skge.hh:195-200).
 *   [sd]lskge3_adjoint: Abar = alpha * op(submat(S))^T * Bbar + beta * Abar,
 *     Bbar is d x n, Abar is m x n (the gradient of B = op(S)*A with respect to A).
 *   [sd]rskge3_adjoint: Abar = alpha * Bbar * op(submat(S))^T + beta * Abar,
 *     Bbar is m x d, Abar is m x n.
 * S is regenerated from its counters; nothing is stored between the forward
 * and the adjoint. */
rbx_status rbx_slskge3_adjoint(char layout, char opS, int64_t d, int64_t n, int64_t m,
                               float alpha, const rbx_dense_skop *S, int64_t ro_s, int64_t co_s,
                               const float *Bbar, int64_t ldbb, float beta, float *Abar,
                               int64_t ldab, void *stream);
rbx_status rbx_dlskge3_adjoint(char layout, char opS, int64_t d, int64_t n, int64_t m,
                               double alpha, const rbx_dense_skop *S, int64_t ro_s, int64_t co_s,
                               const double *Bbar, int64_t ldbb, double beta, double *Abar,
                               int64_t ldab, void *stream);
rbx_status rbx_srskge3_adjoint(char layout, char opS, int64_t m, int64_t d, int64_t n,
                               float alpha, const float *Bbar, int64_t ldbb,
                               const rbx_dense_skop *S, int64_t ro_s, int64_t co_s, float beta,
                               float *Abar, int64_t ldab, void *stream);
rbx_status rbx_drskge3_adjoint(char layout, char opS, int64_t m, int64_t d, int64_t n,
                               double alpha, const double *Bbar, int64_t ldbb,
                               const rbx_dense_skop *S, int64_t ro_s, int64_t co_s, double beta,
                               double *Abar, int64_t ldab, void *stream);

/* BF16-split form of slskge3 (no reference counterpart).  Computes
 *   B = alpha * op(bf16(submat(S))) * op(A) + beta * B
 * on the BF16 matrix cores: every entry of S is rounded to bfloat16
 * (round-to-nearest-even) and A is split into n_terms bfloat16 terms
 * (A = A_1 + ... + A_n_terms + residual, n_terms in {1,2,3}); products
 * accumulate in FP32 and are flushed to FP64 every K-chunk.  The operator
 * is a valid sketch in its own right (a rounded Gaussian); its distortion
 * and its error against the FP64 product are reported by the bench.
 * Only S->buff == NULL (generated operators) is supported. */
rbx_status rbx_slskge3_bf16split(char layout, char opS, char opA, int64_t d, int64_t n,
                                 int64_t m, float alpha, const rbx_dense_skop *S, int64_t ro_s,
                                 int64_t co_s, const float *A, int64_t lda, float beta,
                                 float *B, int64_t ldb, int n_terms, void *stream);

/* BF16-split form of srsksp3 for CSR data (no reference counterpart; the FP32
 * analogue is rsksp3, sparse_data/sksp.hh:277-326 -> csr_spmm_impl.hh:51-119).
 * Computes
 *   B = alpha * A_T * op(bf16(submat(S))) + beta * B
 * with A_T the n_terms-term bfloat16 split of A's values (n_terms in {1,2,3};
 * three terms reproduce FP32 values) and bf16(S) the operator of
 * rbx_slskge3_bf16split (every FP32 entry rounded to bfloat16, nearest-even).
 * The operator is written once per call as a bfloat16 matrix (half the bytes
 * the FP32 product gathers); each output row accumulates in FP32, flushed into
 * FP64 every 256 stored entries.  Only op(A) = NoTrans with the whole CSR
 * matrix (ro_a = co_a = 0) is supported; S may be generated or materialised
 * in FP32. */
rbx_status rbx_srsksp3_bf16split(char layout, char opA, char opS, int64_t m, int64_t d, int64_t n,
                                 float alpha, const rbx_csr_matrix *A, int64_t ro_a, int64_t co_a,
                                 const rbx_dense_skop *S, int64_t ro_s, int64_t co_s, float beta,
                                 float *B, int64_t ldb, int n_terms, void *stream);

/* Self-check of the BF16 operator's fast generator (no reference
 * counterpart).  The BF16-split kernels draw Gaussian entries with the
 * hardware sin / cos / log / sqrt and keep a value only when it provably
 * rounds to the same bfloat16 as the portable (exact) path, using the
 * per-factor error bounds assumed[0] (|s' - s|, |c' - c|, absolute) and
 * assumed[1] (|r' - r| / r).  This call sweeps all 2^32 input words on the
 * device and returns measured[0..3] = max |s' - s|, max |c' - c|,
 * max |r' - r| / r and the number of words where exactly one of r, r' is
 * zero.  The generator is exact when measured[0..1] <= assumed[0],
 * measured[2] <= assumed[1] and measured[3] == 0.  Synchronises `stream`. */
rbx_status rbx_bf16_fastgen_bounds(double *measured, double *assumed, void *stream);

/* ---------------------------------------------------------------- sparse operators
 * SparseSkOp(dist, seed_state) (sparse_skops.hh:356-365) with next_state from
 * compute_next_state (sparse_skops.hh:266-283); nnz = -1, arrays NULL. */
rbx_status rbx_sparse_skop_init(const rbx_sparse_dist *dist, const rbx_rng_state *seed,
                                rbx_sparse_skop *out);
/* dim_major / dim_minor / full_nnz / isometry scale (sparse_skops.hh:108-114,
 * 197-223). */
rbx_status rbx_sparse_dist_props(const rbx_sparse_dist *dist, int64_t *dim_major,
                                 int64_t *dim_minor, int64_t *full_nnz,
                                 double *isometry_scale);
/* fill_sparse_unpacked_nosub (sparse_skops.hh:515-565): writes nnz entries into
 * the device arrays (capacity full_nnz) and returns the next state.  *nnz is
 * written on the host, so this call synchronises `stream`.  SASO index arrays
 * match the reference element for element; LASO vectors that contain
 * duplicates are merged to sqrt(count)*first_sign in first-occurrence order
 * (the reference's order there is that of an std::unordered_map). */
rbx_status rbx_sfill_sparse_unpacked_nosub(const rbx_sparse_dist *D, int64_t *nnz, float *vals,
                                           int64_t *rows, int64_t *cols,
                                           const rbx_rng_state *seed, rbx_rng_state *next_state,
                                           void *stream);
rbx_status rbx_dfill_sparse_unpacked_nosub(const rbx_sparse_dist *D, int64_t *nnz, double *vals,
                                           int64_t *rows, int64_t *cols,
                                           const rbx_rng_state *seed, rbx_rng_state *next_state,
                                           void *stream);
/* fill_sparse(S) (sparse_skops.hh:587-603): allocates the arrays if NULL, and
 * rebuilds S->prep for the new entries (rbx_[sd]sparse_skop_prepare). */
rbx_status rbx_sfill_sparse_skop(rbx_sparse_skop *S, void *stream);
rbx_status rbx_dfill_sparse_skop(rbx_sparse_skop *S, void *stream);
/* ~SparseSkOp (sparse_skops.hh:442-449); also frees S->prep (for operators
 * whose arrays the caller owns, only S->prep). */
rbx_status rbx_sparse_skop_free(rbx_sparse_skop *S);
/* No reference counterpart (the reference re-sorts the operator's COO view in
 * every call, coo_spmm_impl.hh:69-91).  For callers that apply one sampled
 * operator many times: builds S->prep from S's current arrays (replacing any
 * earlier copy), so that later lskges calls over the whole operator
 * (ro_s = co_s = 0, d x m = the operator's shape, opS = 'N') with op(A) of unit
 * stride along the contraction skip the per-call sort; results are unchanged
 * bit for bit.  Leaves S->prep NULL, and returns RBX_OK, when the operator is
 * outside that path (more than 8192 rows, rows of more than 2048 entries on
 * average, a copy above four times the operator's own bytes) or the memory
 * cannot be had.  It finds whether every value is +1 or -1 (as a SASO's
 * are), so that the applies take the signs from the copy's layout and read no
 * values; that check synchronises `stream` once.  The copy is complete when
 * `stream` reaches this call; it stays valid until S's arrays change, which
 * callers that write them in place signal by calling this again or
 * rbx_sparse_skop_free.  rbx_[sd]fill_sparse_skop calls it itself. */
rbx_status rbx_ssparse_skop_prepare(rbx_sparse_skop *S, void *stream);
rbx_status rbx_dsparse_skop_prepare(rbx_sparse_skop *S, void *stream);

/* ---------------------------------------------------------------- sparse sketches
 * lskges (skge.hh:465-492) and rskges (skge.hh:598-626): as lskge3/rskge3 with a
 * SparseSkOp.  Each output entry is accumulated in the reference's order
 * (ascending column of the operator, csc_spmm_impl.hh:74-96), so results with
 * +-1 operator values are bit-identical to the reference's fallback kernels. */
rbx_status rbx_slskges(char layout, char opS, char opA, int64_t d, int64_t n, int64_t m,
                       float alpha, const rbx_sparse_skop *S, int64_t ro_s, int64_t co_s,
                       const float *A, int64_t lda, float beta, float *B, int64_t ldb,
                       void *stream);
rbx_status rbx_dlskges(char layout, char opS, char opA, int64_t d, int64_t n, int64_t m,
                       double alpha, const rbx_sparse_skop *S, int64_t ro_s, int64_t co_s,
                       const double *A, int64_t lda, double beta, double *B, int64_t ldb,
                       void *stream);
rbx_status rbx_srskges(char layout, char opA, char opS, int64_t m, int64_t d, int64_t n,
                       float alpha, const float *A, int64_t lda, const rbx_sparse_skop *S,
                       int64_t ro_s, int64_t co_s, float beta, float *B, int64_t ldb,
                       void *stream);
rbx_status rbx_drskges(char layout, char opA, char opS, int64_t m, int64_t d, int64_t n,
                       double alpha, const double *A, int64_t lda, const rbx_sparse_skop *S,
                       int64_t ro_s, int64_t co_s, double beta, double *B, int64_t ldb,
                       void *stream);

/* (no reference counterpart) lskges without device allocation, for callers
 * that cannot allocate: code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  Same arguments and results as
 * rbx_[sd]lskges (bit-identical), restricted to: S holding its sampled arrays
 * (filled; S->nnz >= 0), op(A) with unit stride along the contraction
 * (layout 'C' with opA 'N', or layout 'R' with opA 'T'), d <= 8192, at most
 * 2048 stored entries per sketch row on average, and a workspace (the
 * operator window's CSR and its chunked layout) that fits the library's fixed
 * per-GPU scratch array of rbx_lskges_noalloc_scratch_bytes() bytes (64 MiB;
 * the 2000 x 100000, vec_nnz 8 SASO operator needs 48 MB in FP64); larger or
 * other shapes return RBX_INVALID_ARG or RBX_UNSUPPORTED with a message and
 * launch nothing.  The array is one per GPU and shared by every call on that
 * GPU: calls whose execution overlaps on one GPU (other streams, other host
 * threads) race.  Over a whole operator that holds its prepared copy (S->prep)
 * the call reads that copy and uses no scratch at all, whatever its size. */
rbx_status rbx_slskges_noalloc(char layout, char opS, char opA, int64_t d, int64_t n, int64_t m,
                               float alpha, const rbx_sparse_skop *S, int64_t ro_s, int64_t co_s,
                               const float *A, int64_t lda, float beta, float *B, int64_t ldb,
                               void *stream);
rbx_status rbx_dlskges_noalloc(char layout, char opS, char opA, int64_t d, int64_t n, int64_t m,
                               double alpha, const rbx_sparse_skop *S, int64_t ro_s, int64_t co_s,
                               const double *A, int64_t lda, double beta, double *B, int64_t ldb,
                               void *stream);
size_t rbx_lskges_noalloc_scratch_bytes(void);

/* ---------------------------------------------------------------- sparse data
 * rsksp3 (sparse_data/sksp.hh:277-326): B = alpha * op(A) * op(submat(S)) + beta * B
 * with A a whole CSR matrix (ro_a = co_a = 0) and S a DenseSkOp.  op(submat(S))
 * is materialised once into library scratch (generated from its counters, as
 * the reference does at sksp.hh:298-305, or gathered from S->buff) and the
 * product accumulates each output entry in the reference's order
 * (csr_spmm_impl.hh:51-119; opA = 'T' takes the CSC order). */
rbx_status rbx_srsksp3(char layout, char opA, char opS, int64_t m, int64_t d, int64_t n,
                       float alpha, const rbx_csr_matrix *A, int64_t ro_a, int64_t co_a,
                       const rbx_dense_skop *S, int64_t ro_s, int64_t co_s, float beta,
                       float *B, int64_t ldb, void *stream);
rbx_status rbx_drsksp3(char layout, char opA, char opS, int64_t m, int64_t d, int64_t n,
                       double alpha, const rbx_csr_matrix *A, int64_t ro_a, int64_t co_a,
                       const rbx_dense_skop *S, int64_t ro_s, int64_t co_s, double beta,
                       double *B, int64_t ldb, void *stream);

/* lsksp3 (sparse_data/sksp.hh:132-180): B = alpha * op(submat(S)) * op(A) + beta * B
 * with S a DenseSkOp (op(submat(S)) d x m at (ro_s, co_s)) and A a whole CSR
 * matrix (op(A) m x n, ro_a = co_a = 0); each output entry is accumulated in
 * the order of the reference's reduction through right_spmm
 * (spmm_dispatch.hh:182-219). */
rbx_status rbx_slsksp3(char layout, char opS, char opA, int64_t d, int64_t n, int64_t m, float alpha,
                       const rbx_dense_skop *S, int64_t ro_s, int64_t co_s, const rbx_csr_matrix *A, int64_t ro_a,
                       int64_t co_a, float beta, float *B, int64_t ldb, void *stream);
rbx_status rbx_dlsksp3(char layout, char opS, char opA, int64_t d, int64_t n, int64_t m, double alpha,
                       const rbx_dense_skop *S, int64_t ro_s, int64_t co_s, const rbx_csr_matrix *A, int64_t ro_a,
                       int64_t co_a, double beta, double *B, int64_t ldb, void *stream);

/* rsksp3 / lsksp3 for CSC and COO data matrices (the reference's templates
 * over SparseMatrix, sksp.hh:132-180, 277-326).  A CSC matrix is used whole
 * (ro_a = co_a = 0); a COO matrix may be a submatrix at (ro_a, co_a), as in
 * spmm_dispatch.hh:95-98.  Summation order per output entry is the
 * reference's for that format (spmm_dispatch.hh:139-176). */
rbx_status rbx_srsksp3_csc(char layout, char opA, char opS, int64_t m, int64_t d, int64_t n,
                           float alpha, const rbx_csc_matrix *A, int64_t ro_a, int64_t co_a,
                           const rbx_dense_skop *S, int64_t ro_s, int64_t co_s, float beta,
                           float *B, int64_t ldb, void *stream);
rbx_status rbx_drsksp3_csc(char layout, char opA, char opS, int64_t m, int64_t d, int64_t n,
                           double alpha, const rbx_csc_matrix *A, int64_t ro_a, int64_t co_a,
                           const rbx_dense_skop *S, int64_t ro_s, int64_t co_s, double beta,
                           double *B, int64_t ldb, void *stream);
rbx_status rbx_srsksp3_coo(char layout, char opA, char opS, int64_t m, int64_t d, int64_t n,
                           float alpha, const rbx_coo_matrix *A, int64_t ro_a, int64_t co_a,
                           const rbx_dense_skop *S, int64_t ro_s, int64_t co_s, float beta,
                           float *B, int64_t ldb, void *stream);
rbx_status rbx_drsksp3_coo(char layout, char opA, char opS, int64_t m, int64_t d, int64_t n,
                           double alpha, const rbx_coo_matrix *A, int64_t ro_a, int64_t co_a,
                           const rbx_dense_skop *S, int64_t ro_s, int64_t co_s, double beta,
                           double *B, int64_t ldb, void *stream);
rbx_status rbx_slsksp3_csc(char layout, char opS, char opA, int64_t d, int64_t n, int64_t m, float alpha,
                           const rbx_dense_skop *S, int64_t ro_s, int64_t co_s, const rbx_csc_matrix *A,
                           int64_t ro_a, int64_t co_a, float beta, float *B, int64_t ldb, void *stream);
rbx_status rbx_dlsksp3_csc(char layout, char opS, char opA, int64_t d, int64_t n, int64_t m, double alpha,
                           const rbx_dense_skop *S, int64_t ro_s, int64_t co_s, const rbx_csc_matrix *A,
                           int64_t ro_a, int64_t co_a, double beta, double *B, int64_t ldb, void *stream);
rbx_status rbx_slsksp3_coo(char layout, char opS, char opA, int64_t d, int64_t n, int64_t m, float alpha,
                           const rbx_dense_skop *S, int64_t ro_s, int64_t co_s, const rbx_coo_matrix *A,
                           int64_t ro_a, int64_t co_a, float beta, float *B, int64_t ldb, void *stream);
rbx_status rbx_dlsksp3_coo(char layout, char opS, char opA, int64_t d, int64_t n, int64_t m, double alpha,
                           const rbx_dense_skop *S, int64_t ro_s, int64_t co_s, const rbx_coo_matrix *A,
                           int64_t ro_a, int64_t co_a, double beta, double *B, int64_t ldb, void *stream);

/* Adjoint of rsksp3 (opA = 'N') with respect to the stored values of A (no
 * reference counterpart; it is the gradient of B = alpha * A * op(submat(S))
 * with respect to A's nonzeros):
 *   dvals[k] = alpha * sum_j Bbar[i_k, j] * op(submat(S))[c_k, j]
 * for every stored entry k = (i_k, c_k) of A (a sampled dense-dense product).
 * A is m x n, Bbar is m x d in `layout`, dvals has A->nnz entries.  S may be
 * generated (buff == NULL) or materialised. */
rbx_status rbx_srsksp3_adjoint(char layout, char opS, int64_t m, int64_t d, int64_t n, float alpha,
                               const rbx_csr_matrix *A, const rbx_dense_skop *S, int64_t ro_s, int64_t co_s,
                               const float *Bbar, int64_t ldbb, float *dvals, void *stream);
rbx_status rbx_drsksp3_adjoint(char layout, char opS, int64_t m, int64_t d, int64_t n, double alpha,
                               const rbx_csr_matrix *A, const rbx_dense_skop *S, int64_t ro_s, int64_t co_s,
                               const double *Bbar, int64_t ldbb, double *dvals, void *stream);

/* ---------------------------------------------------------------- utilities
 * util::require_symmetric (RandBLAS/util.hh:127-148), the check that
 * sketch_symmetric runs first (sksy.hh:175,310): fails with RBX_INVALID_ARG
 * and the reference's message for the first pair (i < j, row-major loop
 * order) with |A(i,j) - A(j,i)| > (|A(i,j)| + |A(j,i)| + 1) * tol, computed in
 * the matrix's precision.  tol < 0 skips the check.  Synchronises `stream`. */
rbx_status rbx_srequire_symmetric(char layout, const float *A, int64_t n, int64_t lda, float tol, void *stream);
rbx_status rbx_drequire_symmetric(char layout, const double *A, int64_t n, int64_t lda, double tol, void *stream);

/* ---------------------------------------------------------------- format conversions
 * csr_to_coo / csc_to_coo (sparse_data/conversions.hh:49-75): the COO arrays
 * (caller-allocated, nnz entries) receive the stored entries in the compressed
 * order; coo->nnz is set.  coo_to_csr / coo_to_csc (conversions.hh:79-126):
 * the entries sorted by (row, column) / (column, row) into the caller's
 * index and value arrays (nnz entries) and pointer array (n_rows + 1 /
 * n_cols + 1); the sort is stable, so equal (row, column) pairs keep their
 * stored order.  coo_sort(order, coo) (coo_arrays_apply_sort,
 * sparse_data/base.hh:169-220) sorts a COO matrix in place, order 'R' (CSR)
 * or 'C' (CSC).  coo_to_csr, coo_to_csc and coo_sort first find how the
 * entries already lie (one pass and a 4-byte read, so they synchronise the
 * stream) and copy entries already in the target order, or sort entries in
 * the other order on the major index alone (same result).
 * coo_determine_sort (base.hh:142-167) sets *order to 'C' if the entries are
 * in CSC order, else 'R' if in CSR order, else 'N', and synchronises the
 * stream.  Zero-based indices. */
rbx_status rbx_scsr_to_coo(const rbx_csr_matrix *csr, rbx_coo_matrix *coo, void *stream);
rbx_status rbx_dcsr_to_coo(const rbx_csr_matrix *csr, rbx_coo_matrix *coo, void *stream);
rbx_status rbx_scsc_to_coo(const rbx_csc_matrix *csc, rbx_coo_matrix *coo, void *stream);
rbx_status rbx_dcsc_to_coo(const rbx_csc_matrix *csc, rbx_coo_matrix *coo, void *stream);
rbx_status rbx_scoo_to_csr(const rbx_coo_matrix *coo, rbx_csr_matrix *csr, void *stream);
rbx_status rbx_dcoo_to_csr(const rbx_coo_matrix *coo, rbx_csr_matrix *csr, void *stream);
rbx_status rbx_scoo_to_csc(const rbx_coo_matrix *coo, rbx_csc_matrix *csc, void *stream);
rbx_status rbx_dcoo_to_csc(const rbx_coo_matrix *coo, rbx_csc_matrix *csc, void *stream);
rbx_status rbx_scoo_sort(char order, rbx_coo_matrix *coo, void *stream);
rbx_status rbx_dcoo_sort(char order, rbx_coo_matrix *coo, void *stream);
rbx_status rbx_coo_determine_sort(const rbx_coo_matrix *coo, char *order, void *stream);

/* ---------------------------------------------------------------- index sampling
 * sample_indices_iid_uniform(n, k, samples, state) (RandBLAS/util.hh:552-560):
 * k iid indices in [0, n), samples[i] = (int64) (n * (uneg11(w) + 1) / 2) in
 * double for word w = i mod 4 of counter state + floor(i / 4).  *next_state
 * (may be NULL) receives the state the reference returns: the seed advanced
 * by ceil(k / 4) counters.  Device array `samples` of k entries. */
rbx_status rbx_sample_indices_iid_uniform(int64_t n, int64_t k, int64_t *samples, const rbx_rng_state *state,
                                          rbx_rng_state *next_state, void *stream);
/* The same with Rademacher signs (sample_indices_iid_uniform<T, sint_t, true>,
 * util.hh:516-550, the sampler of LASO operators): sample i uses words
 * 2(i mod 2) and 2(i mod 2) + 1 of counter state + floor(i / 2);
 * rademachers[i] = +1 if uneg11 of the second word is >= 0, else -1;
 * the state advances by ceil(k / 2). */
rbx_status rbx_ssample_indices_iid_uniform_signed(int64_t n, int64_t k, int64_t *samples, float *rademachers,
                                                  const rbx_rng_state *state, rbx_rng_state *next_state,
                                                  void *stream);
rbx_status rbx_dsample_indices_iid_uniform_signed(int64_t n, int64_t k, int64_t *samples, double *rademachers,
                                                  const rbx_rng_state *state, rbx_rng_state *next_state,
                                                  void *stream);
/* sample_indices_iid(n, cdf, k, samples, state) (util.hh:493-514): samples[i]
 * is the first index j with cdf[j] >= u (std::lower_bound), u = ((T)
 * uneg11(w) + 1) / 2 in T for word w = i mod 4 of counter state + floor(i / 4);
 * the state advances by ceil(k / 4).  `cdf` is a device array of n entries. */
rbx_status rbx_ssample_indices_iid(int64_t n, const float *cdf, int64_t k, int64_t *samples,
                                   const rbx_rng_state *state, rbx_rng_state *next_state, void *stream);
rbx_status rbx_dsample_indices_iid(int64_t n, const double *cdf, int64_t k, int64_t *samples,
                                   const rbx_rng_state *state, rbx_rng_state *next_state, void *stream);
/* weights_to_cdf(n, w, error_if_below) (util.hh:459-472), in place on the
 * device, summed in the reference's sequential order (one thread, so the
 * partial sums round as the reference's do); the reference's default
 * error_if_below is -sqrt(epsilon of T).  Fails with RBX_INVALID_ARG, leaving
 * w as the reference leaves it when it throws, if a weight is below
 * error_if_below or the weights sum to less than sqrt(n) * epsilon.
 * Synchronises the stream. */
rbx_status rbx_sweights_to_cdf(int64_t n, float *w, float error_if_below, void *stream);
rbx_status rbx_dweights_to_cdf(int64_t n, double *w, double error_if_below, void *stream);

/* Power-law sparse test data (no reference counterpart: the reference's
 * generators, sparse_data/random_matrix.hh:136-358, draw uniform-density
 * matrices from one sequential PhiloxStream and cannot run in parallel bit
 * for bit).  An m x n CSR matrix built from this library's counter-addressed
 * samplers: k column draws by sample_indices_iid over the Zipf weights
 * w[c] = (c+1)^(-1/(exponent-1)) (host libm pow, rounded to T; CDF by
 * weights_to_cdf), starting at `state`; k row draws by
 * sample_indices_iid_uniform(m, k) from the state after them; the pairs
 * sorted by (row, column) with repeated pairs kept once; the nnz values the
 * Gaussian entries 0..nnz-1 of a DenseDist(nnz, 1) operator from the state
 * after the row draws.  next_state = that state + ceil(nnz / 4).  Column
 * degrees follow a power law with the given exponent (> 1).  A->rowptr
 * (m + 1), A->colidxs and A->vals (k entries each) are caller-allocated device
 * arrays; A->n_rows, A->n_cols and A->nnz are set.  Synchronises the stream. */
rbx_status rbx_spower_law_csr(int64_t m, int64_t n, int64_t k, double exponent, const rbx_rng_state *state,
                              rbx_csr_matrix *A, rbx_rng_state *next_state, void *stream);
rbx_status rbx_dpower_law_csr(int64_t m, int64_t n, int64_t k, double exponent, const rbx_rng_state *state,
                              rbx_csr_matrix *A, rbx_rng_state *next_state, void *stream);

/* ---------------------------------------------------------------- sketch-and-solve
 * The downstream use of a left sketch: examples/total-least-squares/
 * tls_dense_skop.cc:138-171 sketches [A b] with sketch_general and solves the
 * small sketched problem on the host.  These do the sketched least-squares
 * solve on the device:
 *   x = argmin_x || S (A x - b) ||_2,   S = submat(S) at (ro_s, co_s), d x m,
 * A m x n in `layout` (leading dimension lda), b and x device vectors of m and
 * n entries, d >= n.  SA and Sb come from the fused sketch (lskge3, opS = opA
 * = 'N'; S is never stored), the QR factorisation of the d x n sketched system
 * and the triangular solve run in FP64 for both precisions.  A rank-deficient
 * SA gives non-finite x.  The adjoint returns the gradient of <xbar, x(A, b)>:
 *   Abar = S^T (r z^T - w x^T),  bbar = S^T w,
 * with SA = QR, u = R^{-T} xbar, w = Q u, z = R^{-1} u, r = Sb - SA x; Abar is
 * m x n in `layout` (leading dimension ldab), bbar has m entries.  Both keep
 * their workspace between calls. */
rbx_status rbx_ssketch_solve(char layout, int64_t d, int64_t m, int64_t n, const rbx_dense_skop *S,
                             int64_t ro_s, int64_t co_s, const float *A, int64_t lda, const float *b,
                             float *x, void *stream);
rbx_status rbx_dsketch_solve(char layout, int64_t d, int64_t m, int64_t n, const rbx_dense_skop *S,
                             int64_t ro_s, int64_t co_s, const double *A, int64_t lda, const double *b,
                             double *x, void *stream);
rbx_status rbx_ssketch_solve_adjoint(char layout, int64_t d, int64_t m, int64_t n, const rbx_dense_skop *S,
                                     int64_t ro_s, int64_t co_s, const float *A, int64_t lda,
                                     const float *b, const float *xbar, float *Abar, int64_t ldab,
                                     float *bbar, void *stream);
rbx_status rbx_dsketch_solve_adjoint(char layout, int64_t d, int64_t m, int64_t n, const rbx_dense_skop *S,
                                     int64_t ro_s, int64_t co_s, const double *A, int64_t lda,
                                     const double *b, const double *xbar, double *Abar, int64_t ldab,
                                     double *bbar, void *stream);

/* ---------------------------------------------------------------- workspaces
 * (no reference counterpart: the reference allocates its temporaries per call,
 * skge.hh:175-181)  The library keeps grow-only device scratch and a rocBLAS
 * handle per (device, stream), so calls on different streams are independent
 * and may run concurrently; host threads that use one stream take turns.
 * rbx_release_workspace frees the calling device's workspace of `stream`
 * (stream-ordered: after the work already enqueued there); call it before
 * destroying a stream the library has used.  rbx_release_all_workspaces
 * synchronises every device the library used and frees everything; call it
 * with no library work in flight. */
rbx_status rbx_release_workspace(void *stream);
rbx_status rbx_release_all_workspaces(void);

/* ---------------------------------------------------------------- diagnostics
 * (no reference counterpart) Kernel timing for benchmarks: while enabled,
 * every kernel the library launches is bracketed by HIP events on its stream.
 * rbx_profile_collect synchronises those events and returns, per kernel name,
 * the number of launches and their total duration in milliseconds (at most
 * max_entries rows; names are NUL-terminated, name_len bytes apart), then
 * clears the record.  Returns the number of rows written, or -1 on error. */
rbx_status rbx_profile_enable(int on);
int rbx_profile_collect(char *names, int name_len, int *counts, double *total_ms, int max_entries);

#ifdef __cplusplus
}
#endif
#endif /* RANDBLAS_AMD_H */
