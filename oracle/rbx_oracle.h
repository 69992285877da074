/*
 * rbx_oracle.h -- CPU restatement of RandBLAS's sketching hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in randblas_amd/.  Only the test suite, the smoke check and
 * the CPU-baseline leg of bench.py may load it.  The product never links it.
 *
 * Each function names the reference file:line it restates (paths relative to
 * the RandBLAS source tree).  Conventions shared by all entry points:
 *   - counters are 4 x uint32 (word 0 least significant), keys 2 x uint32;
 *   - layout is 'C' (column-major) or 'R' (row-major); op is 'N' or 'T';
 *   - dtype is 'd' (double) or 's' (float);
 *   - mathmode 0 = the portable Box-Muller math (sin/cos/log evaluated in
 *     double with explicit fma, identical on CPU and GPU); mathmode 1 = the
 *     host C library's sincosf/logf/sqrtf, i.e. what the reference computes
 *     when built against Random123 on this platform.
 * Return value 0 = success, nonzero = invalid argument.
 */
#ifndef RBX_ORACLE_H
#define RBX_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Philox4x32-10 and counters ------------------------------------- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void orc_philox_stream(const uint32_t ctr[4], const uint32_t key[2], int64_t n_ctr, uint32_t *out);
void orc_ctr_incr(uint32_t ctr[4], uint64_t n);

/* ---- scalar maps ----------------------------------------------------- */
float orc_uneg11(uint32_t w);
float orc_u01(uint32_t w);
float orc_sinpi_arg(uint32_t w);              /* fl(pi_f * uneg11(w)) */
void  orc_sincos_t(float t, int mathmode, float *s, float *c);
float orc_log_u(float u, int mathmode);
void  orc_boxmuller(uint32_t w0, uint32_t w1, int mathmode, float out[2]);
/* Bulk versions over arrays of words (OpenMP). */
void  orc_boxmuller_bulk(int64_t n_pairs, const uint32_t *w, int mathmode, float *out);
void  orc_factor_bulk(int64_t n, const uint32_t *w, int which, int mathmode, float *out);

/* ---- dense operators ------------------------------------------------- */
int orc_dense_next_state(int64_t n_rows, int64_t n_cols, char major_axis,
                         const uint32_t ctr[4], uint32_t next_ctr[4]);
int orc_fill_dense_unpacked(char layout, int64_t D_rows, int64_t D_cols, char family,
                            char major_axis, int64_t n_rows, int64_t n_cols,
                            int64_t ro_s, int64_t co_s, char dtype, void *buff,
                            const uint32_t ctr[4], const uint32_t key[2],
                            uint32_t next_ctr[4], int mathmode, int nthreads);

/* General strided GEMM in double: C = alpha*X*Y + beta*C, X is M x K, Y is K x N.
 * Element (i,j) of a matrix with strides (rs, cs) lives at base[i*rs + j*cs]. */
void orc_gemm_d(int64_t M, int64_t N, int64_t K, double alpha,
                const double *X, int64_t xrs, int64_t xcs,
                const double *Y, int64_t yrs, int64_t ycs,
                double beta, double *C, int64_t crs, int64_t ccs, int nthreads);

/* lskge3 / rskge3 with an unmaterialised DenseSkOp (skge.hh:154-202, 307-355).
 * Computed in double; for dtype 's' the float inputs are promoted and the
 * result is rounded to float once. */
int orc_lskge3(char dtype, char layout, char opS, char opA, int64_t d, int64_t n, int64_t m,
               double alpha, int64_t S_rows, int64_t S_cols, char family, char major_axis,
               const uint32_t ctr[4], const uint32_t key[2], int64_t ro_s, int64_t co_s,
               const void *A, int64_t lda, double beta, void *B, int64_t ldb,
               int mathmode, int nthreads);
int orc_rskge3(char dtype, char layout, char opA, char opS, int64_t m, int64_t d, int64_t n,
               double alpha, const void *A, int64_t lda,
               int64_t S_rows, int64_t S_cols, char family, char major_axis,
               const uint32_t ctr[4], const uint32_t key[2], int64_t ro_s, int64_t co_s,
               double beta, void *B, int64_t ldb, int mathmode, int nthreads);

/* ---- sparse operators ------------------------------------------------ */
int orc_sparse_next_state(int64_t n_rows, int64_t n_cols, int64_t vec_nnz, char major_axis,
                          const uint32_t ctr[4], uint32_t next_ctr[4]);
/* fill_sparse_unpacked_nosub (sparse_skops.hh:515-565).  Arrays hold full_nnz
 * entries.  LASO vectors with duplicates are written in first-occurrence order. */
int orc_fill_sparse(int64_t n_rows, int64_t n_cols, int64_t vec_nnz, char major_axis,
                    char dtype, int64_t *nnz_out, void *vals, int64_t *rows, int64_t *cols,
                    const uint32_t ctr[4], const uint32_t key[2], uint32_t next_ctr[4]);
/* left_spmm on a COO matrix (spmm_dispatch.hh:52-179, coo_spmm_impl.hh:52-105,
 * csc_spmm_impl.hh:74-209): C = alpha*op(submat(A))*op(B) + beta*C. */
int orc_left_spmm_coo(char dtype, char layout, char opA, char opB, int64_t d, int64_t n, int64_t m,
                      double alpha, int64_t A_rows, int64_t A_cols, int64_t nnz,
                      const void *vals, const int64_t *rows, const int64_t *cols,
                      int64_t ro_a, int64_t co_a, const void *B, int64_t ldb,
                      double beta, void *C, int64_t ldc, int nthreads);
/* left_spmm on a CSR matrix (csr_spmm_impl.hh:51-155), no offsets. */
int orc_left_spmm_csr(char dtype, char layout, char opB, int64_t d, int64_t n, int64_t m,
                      double alpha, const void *vals, const int64_t *rowptr,
                      const int64_t *colidxs, const void *B, int64_t ldb,
                      double beta, void *C, int64_t ldc, int nthreads);

/* Gradient of alpha * A * Sm with respect to A's stored values (SDDMM),
 * double accumulation in ascending j. */
int orc_sddmm_csr(char dtype, int64_t m, int64_t d, double alpha, const int64_t *rowptr,
                  const int64_t *colidxs, const void *Bbar, int64_t brs, int64_t bcs,
                  const void *Sm, int64_t srs, int64_t scs, void *dvals, int nthreads);

/* index sampling (util.hh:459-560) */
void orc_sample_indices_iid_uniform(int64_t n, int64_t k, int64_t *samples, double *rademachers,
                                    const uint32_t ctr[4], const uint32_t key[2], uint32_t next_ctr[4]);
void orc_sample_indices_iid(char dtype, int64_t n, const void *cdf, int64_t k, int64_t *samples,
                            const uint32_t ctr[4], const uint32_t key[2], uint32_t next_ctr[4]);
int64_t orc_weights_to_cdf(char dtype, int64_t n, void *w, double error_if_below);
/* Zipf weights of the power-law sparse test data (no reference counterpart):
 * w[c] = pow(c + 1, -1 / (exponent - 1)) by libm. */
void orc_power_law_weights(int64_t n, double exponent, double *w);

#ifdef __cplusplus
}
#endif
#endif
