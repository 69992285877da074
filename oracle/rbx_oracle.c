/*
 * rbx_oracle.c -- CPU restatement of RandBLAS's sketching hot path.
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker for the HIP kernels in
 * randblas_amd/csrc.  It is loaded by tests/, by __graft_entry__.smoke() and
 * by the CPU-baseline leg of bench.py, never by the product.
 *
 * Every routine cites the RandBLAS source it restates (file:line, paths
 * relative to the RandBLAS tree).  Third-party arithmetic that RandBLAS takes
 * from Random123 (un-vendored, unpinned version) is restated from its
 * published description:
 *   - Philox4x32-10: Salmon, Moraes, Dror, Shaw, "Parallel random numbers:
 *     as easy as 1, 2, 3", SC'11.  Pinned by the Random123 known-answer rows
 *     held in test/test_basic_rng/r123_kat_vectors.txt:19-21.
 *   - u01 / uneg11: (float)x * 2^-32 + 2^-33 and (float)(int32)x * 2^-31 + 2^-32.
 *   - boxmuller(u0,u1): (s,c) = sincos(pi_f * uneg11(u0)); r = sqrtf(-2 logf(u01(u1)));
 *     return (r*s, r*c).
 * The portable math (mathmode 0) evaluates sin, cos and log of a float
 * argument in double precision (Cody-Waite reduction by pi/2 with a two-part
 * constant, Taylor polynomials in Horner form; log via the atanh series
 * 2*atanh(f/(2+f))), then rounds once to float.  The HIP generator performs
 * exactly the same IEEE double operations (explicit fma, no contraction), so
 * its Gaussians are bit-identical to mathmode 0.  mathmode 1 calls the host
 * C library, which is what the reference computes on Linux/glibc.
 */
#define _GNU_SOURCE
#include "rbx_oracle.h"
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

/* ======================================================================
 * Philox4x32-10 (Random123 philox.h as used at random_gen.hh:107,135,
 * sparse_skops.hh:77, util.hh:521,540)
 * ====================================================================== */
#define PH_M0 0xD2511F53u
#define PH_M1 0xCD9E8D57u
#define PH_W0 0x9E3779B9u
#define PH_W1 0xBB67AE85u

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t x0 = ctr[0], x1 = ctr[1], x2 = ctr[2], x3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += PH_W0; k1 += PH_W1; }
        uint64_t p0 = (uint64_t)PH_M0 * x0;
        uint64_t p1 = (uint64_t)PH_M1 * x2;
        uint32_t y0 = (uint32_t)(p1 >> 32) ^ x1 ^ k0;
        uint32_t y1 = (uint32_t)p1;
        uint32_t y2 = (uint32_t)(p0 >> 32) ^ x3 ^ k1;
        uint32_t y3 = (uint32_t)p0;
        x0 = y0; x1 = y1; x2 = y2; x3 = y3;
    }
    out[0] = x0; out[1] = x1; out[2] = x2; out[3] = x3;
}

/* r123array::incr(uint64) semantics, pinned by test_r123.cc:706-797:
 * the counter is a 128-bit little-endian integer in 32-bit words. */
void orc_ctr_incr(uint32_t c[4], uint64_t n) {
    uint64_t s = (uint64_t)c[0] + (n & 0xffffffffull);
    c[0] = (uint32_t)s;
    s = (uint64_t)c[1] + (n >> 32) + (s >> 32);
    c[1] = (uint32_t)s;
    s = (uint64_t)c[2] + (s >> 32);
    c[2] = (uint32_t)s;
    s = (uint64_t)c[3] + (s >> 32);
    c[3] = (uint32_t)s;
}

void orc_philox_stream(const uint32_t ctr[4], const uint32_t key[2], int64_t n_ctr, uint32_t *out) {
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n_ctr; ++i) {
        uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
        orc_ctr_incr(c, (uint64_t)i);
        orc_philox4x32_10(c, key, out + 4 * i);
    }
}

/* ======================================================================
 * Uniform maps (Random123 uniform.hpp, used via r123::uneg11all at
 * random_gen.hh:135 and inside r123::boxmuller).
 * ====================================================================== */
float orc_uneg11(uint32_t w) {
    volatile float a = (float)(int32_t)w * 0x1p-31f;   /* exact: power-of-two scale */
    return a + 0x1p-32f;
}
float orc_u01(uint32_t w) {
    volatile float a = (float)w * 0x1p-32f;
    return a + 0x1p-33f;
}
float orc_sinpi_arg(uint32_t w) {
    const float PIf = 3.1415926535897932f;
    return PIf * orc_uneg11(w);
}

/* ======================================================================
 * Portable Box-Muller math (mathmode 0).  Same operation sequence as
 * randblas_amd/csrc/rb_philox.hh on the GPU.
 * ====================================================================== */
static const double RBM_TWO_OVER_PI = 0x1.45f306dc9c883p-1;
static const double RBM_PIO2_HI = 0x1.921fb54442d18p+0;
static const double RBM_PIO2_LO = 0x1.1a62633145c07p-54;
/* (-1)^k / (2k+1)! for k = 1..8 */
static const double RBM_S3 = -0x1.5555555555555p-3, RBM_S5 = 0x1.1111111111111p-7,
    RBM_S7 = -0x1.a01a01a01a01ap-13, RBM_S9 = 0x1.71de3a556c734p-19,
    RBM_S11 = -0x1.ae64567f544e4p-26, RBM_S13 = 0x1.6124613a86d09p-33,
    RBM_S15 = -0x1.ae7f3e733b81fp-41, RBM_S17 = 0x1.952c77030ad4ap-49;
/* (-1)^k / (2k)! for k = 1..8 */
static const double RBM_C2 = -0x1.0000000000000p-1, RBM_C4 = 0x1.5555555555555p-5,
    RBM_C6 = -0x1.6c16c16c16c17p-10, RBM_C8 = 0x1.a01a01a01a01ap-16,
    RBM_C10 = -0x1.27e4fb7789f5cp-22, RBM_C12 = 0x1.1eed8eff8d898p-29,
    RBM_C14 = -0x1.93974a8c07c9dp-37, RBM_C16 = 0x1.ae7f3e733b81fp-45;
/* 2/(2k+1) for k = 1..10 */
static const double RBM_L1 = 0x1.5555555555555p-1, RBM_L2 = 0x1.999999999999ap-2,
    RBM_L3 = 0x1.2492492492492p-2, RBM_L4 = 0x1.c71c71c71c71cp-3,
    RBM_L5 = 0x1.745d1745d1746p-3, RBM_L6 = 0x1.3b13b13b13b14p-3,
    RBM_L7 = 0x1.1111111111111p-3, RBM_L8 = 0x1.e1e1e1e1e1e1ep-4,
    RBM_L9 = 0x1.af286bca1af28p-4, RBM_L10 = 0x1.8618618618618p-4;
static const double RBM_LN2_HI = 0x1.62e42fee00000p-1;
static const double RBM_LN2_LO = 0x1.a39ef35793c76p-33;
static const double RBM_SQRT2 = 0x1.6a09e667f3bcdp+0;

static void rbm_sincos(float t, float *s_out, float *c_out) {
    double x = (double)t;
    double kd = rint(x * RBM_TWO_OVER_PI);
    double r = fma(-kd, RBM_PIO2_HI, x);
    r = fma(-kd, RBM_PIO2_LO, r);
    double r2 = r * r;
    double ps = RBM_S17;
    ps = fma(ps, r2, RBM_S15); ps = fma(ps, r2, RBM_S13); ps = fma(ps, r2, RBM_S11);
    ps = fma(ps, r2, RBM_S9);  ps = fma(ps, r2, RBM_S7);  ps = fma(ps, r2, RBM_S5);
    ps = fma(ps, r2, RBM_S3);
    double sv = fma(r * r2, ps, r);
    double pc = RBM_C16;
    pc = fma(pc, r2, RBM_C14); pc = fma(pc, r2, RBM_C12); pc = fma(pc, r2, RBM_C10);
    pc = fma(pc, r2, RBM_C8);  pc = fma(pc, r2, RBM_C6);  pc = fma(pc, r2, RBM_C4);
    pc = fma(pc, r2, RBM_C2);
    double cv = fma(r2, pc, 1.0);
    int q = ((int)kd) & 3;
    double S, C;
    switch (q) {
        case 0: S = sv;  C = cv;  break;
        case 1: S = cv;  C = -sv; break;
        case 2: S = -sv; C = -cv; break;
        default: S = -cv; C = sv; break;
    }
    *s_out = (float)S;
    *c_out = (float)C;
}

static float rbm_logf(float u) {
    uint32_t bits;
    memcpy(&bits, &u, 4);
    int e = (int)(bits >> 23) - 127;
    uint32_t mb = (bits & 0x7fffffu) | 0x3f800000u;
    float mf;
    memcpy(&mf, &mb, 4);
    double m = (double)mf;
    if (m > RBM_SQRT2) { m = m * 0.5; e += 1; }
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double s2 = s * s;
    double p = RBM_L10;
    p = fma(p, s2, RBM_L9); p = fma(p, s2, RBM_L8); p = fma(p, s2, RBM_L7);
    p = fma(p, s2, RBM_L6); p = fma(p, s2, RBM_L5); p = fma(p, s2, RBM_L4);
    p = fma(p, s2, RBM_L3); p = fma(p, s2, RBM_L2); p = fma(p, s2, RBM_L1);
    double lm = fma(s * s2, p, 2.0 * s);
    double ed = (double)e;
    double res = fma(ed, RBM_LN2_HI, fma(ed, RBM_LN2_LO, lm));
    return (float)res;
}

void orc_sincos_t(float t, int mathmode, float *s, float *c) {
    if (mathmode == 1) {
        sincosf(t, s, c);
    } else {
        rbm_sincos(t, s, c);
    }
}

float orc_log_u(float u, int mathmode) {
    return (mathmode == 1) ? logf(u) : rbm_logf(u);
}

void orc_boxmuller(uint32_t w0, uint32_t w1, int mathmode, float out[2]) {
    float s, c;
    orc_sincos_t(orc_sinpi_arg(w0), mathmode, &s, &c);
    volatile float m2l = -2.0f * orc_log_u(orc_u01(w1), mathmode);
    float r = sqrtf(m2l);
    out[0] = s * r;
    out[1] = c * r;
}

void orc_boxmuller_bulk(int64_t n_pairs, const uint32_t *w, int mathmode, float *out) {
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n_pairs; ++i)
        orc_boxmuller(w[2 * i], w[2 * i + 1], mathmode, out + 2 * i);
}

/* which: 0 = sin factor, 1 = cos factor, 2 = radius sqrtf(-2 log u01) */
void orc_factor_bulk(int64_t n, const uint32_t *w, int which, int mathmode, float *out) {
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        if (which == 2) {
            volatile float m2l = -2.0f * orc_log_u(orc_u01(w[i]), mathmode);
            out[i] = sqrtf(m2l);
        } else {
            float s, c;
            orc_sincos_t(orc_sinpi_arg(w[i]), mathmode, &s, &c);
            out[i] = (which == 0) ? s : c;
        }
    }
}

/* r123ext::boxmul::generate / uneg11::generate (random_gen.hh:85-137):
 * one Philox call -> 4 floats. */
static void gen4(int gauss, const uint32_t c[4], const uint32_t k[2], int mathmode, float v[4]) {
    uint32_t w[4];
    orc_philox4x32_10(c, k, w);
    if (gauss) {
        orc_boxmuller(w[0], w[1], mathmode, v);
        orc_boxmuller(w[2], w[3], mathmode, v + 2);
    } else {
        for (int i = 0; i < 4; ++i) v[i] = orc_uneg11(w[i]);
    }
}

/* ======================================================================
 * Dense operators (dense_skops.hh)
 * ====================================================================== */

/* DenseDist dim_major / natural layout (dense_skops.hh:313-329, 187-199). */
static int64_t dd_dim_major(int64_t r, int64_t c, char ax) {
    return (ax == 'L') ? (r > c ? r : c) : (r < c ? r : c);
}
static int64_t dd_dim_minor(int64_t r, int64_t c, char ax) {
    return (ax == 'L') ? (r < c ? r : c) : (r > c ? r : c);
}
static char dd_natural_layout(int64_t r, int64_t c, char ax) {
    int wide = r < c, flong = ax == 'L';
    if (wide && flong) return 'R';
    if (wide) return 'C';
    if (flong) return 'C';
    return 'R';
}

/* dense::compute_next_state (dense_skops.hh:172-185) */
int orc_dense_next_state(int64_t n_rows, int64_t n_cols, char major_axis,
                         const uint32_t ctr[4], uint32_t next_ctr[4]) {
    if (n_rows <= 0 || n_cols <= 0) return 1;
    int64_t maj = dd_dim_major(n_rows, n_cols, major_axis);
    int64_t mnr = dd_dim_minor(n_rows, n_cols, major_axis);
    int64_t pad = (maj % 4) ? 4 - maj % 4 : 0;
    int64_t stride = (maj + pad) / 4;
    memcpy(next_ctr, ctr, 16);
    orc_ctr_incr(next_ctr, (uint64_t)stride * (uint64_t)mnr);
    return 0;
}

/* dense::fill_dense_submat_impl (dense_skops.hh:96-170), row-major output,
 * OpenMP over rows with schedule(static) as at :133-164. */
#define DEF_FILL_SUBMAT(T, NAME)                                                        \
static void NAME(int64_t n_cols, T *smat, int64_t n_srows, int64_t n_scols, int64_t ptr, \
                 const uint32_t seed_ctr[4], const uint32_t key[2], int gauss,           \
                 int mathmode, uint32_t next_ctr[4], int nthreads) {                     \
    const int64_t cs = 4;                                                                \
    int64_t pad = (n_cols % cs != 0) ? cs - n_cols % cs : 0;                              \
    const int64_t ptr_padded = ptr + ptr / n_cols * pad;                                 \
    const int64_t ctr_mat_start = ptr_padded / cs;                                       \
    const int64_t first_block_start = ptr_padded % cs;                                   \
    const int64_t ctr_mat_row_end = (ptr_padded + n_scols - 1) / cs;                     \
    const int64_t last_block_stop = ((ptr_padded + n_scols - 1) % cs) + 1;               \
    const int64_t ctr_inter_row_stride = (n_cols + pad) / cs;                            \
    const int one_block_per_row = ctr_mat_start == ctr_mat_row_end;                      \
    const int64_t first_block_len = (one_block_per_row ? last_block_stop : cs)           \
                                    - first_block_start;                                 \
    uint32_t c[4] = {seed_ctr[0], seed_ctr[1], seed_ctr[2], seed_ctr[3]};                \
    orc_ctr_incr(c, (uint64_t)ctr_mat_start);                                            \
    _Pragma("omp parallel for schedule(static) num_threads(nthreads)")                   \
    for (int64_t row = 0; row < n_srows; row++) {                                        \
        uint32_t cr[4] = {c[0], c[1], c[2], c[3]};                                       \
        orc_ctr_incr(cr, (uint64_t)(ctr_inter_row_stride * row));                        \
        float rv[4];                                                                     \
        gen4(gauss, cr, key, mathmode, rv);                                              \
        T *srow = smat + row * n_scols;                                                  \
        for (int64_t i = 0; i < first_block_len; i++) srow[i] = (T)rv[i + first_block_start]; \
        if (one_block_per_row) continue;                                                 \
        int64_t ind = first_block_len;                                                   \
        for (int64_t i = 0; i < (ctr_mat_row_end - ctr_mat_start - 1); ++i) {            \
            orc_ctr_incr(cr, 1);                                                         \
            gen4(gauss, cr, key, mathmode, rv);                                          \
            for (int t = 0; t < 4; ++t) srow[ind + t] = (T)rv[t];                        \
            ind += cs;                                                                   \
        }                                                                                \
        orc_ctr_incr(cr, 1);                                                             \
        gen4(gauss, cr, key, mathmode, rv);                                              \
        for (int64_t t = 0; t < last_block_stop; ++t) srow[ind + t] = (T)rv[t];          \
    }                                                                                    \
    memcpy(next_ctr, c, 16);                                                             \
    orc_ctr_incr(next_ctr, (uint64_t)(n_srows * ctr_inter_row_stride));                  \
}
DEF_FILL_SUBMAT(double, fill_submat_d)
DEF_FILL_SUBMAT(float, fill_submat_s)

/* fill_dense_unpacked (dense_skops.hh:563-606) */
int orc_fill_dense_unpacked(char layout, int64_t D_rows, int64_t D_cols, char family,
                            char major_axis, int64_t n_rows, int64_t n_cols,
                            int64_t ro_s, int64_t co_s, char dtype, void *buff,
                            const uint32_t ctr[4], const uint32_t key[2],
                            uint32_t next_ctr[4], int mathmode, int nthreads) {
    if (D_rows <= 0 || D_cols <= 0 || n_rows <= 0 || n_cols <= 0) return 1;
    if (ro_s < 0 || co_s < 0) return 1;
    if (D_rows < n_rows + ro_s || D_cols < n_cols + co_s) return 1;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    char nat = dd_natural_layout(D_rows, D_cols, major_axis);
    int64_t ma_len = dd_dim_major(D_rows, D_cols, major_axis);
    int64_t nr_, nc_, ptr;
    if (nat == 'C') { nr_ = n_cols; nc_ = n_rows; ptr = ro_s + co_s * ma_len; }
    else            { nr_ = n_rows; nc_ = n_cols; ptr = ro_s * ma_len + co_s; }
    int gauss = (family == 'G');
    if (dtype == 'd') {
        double *b = (double *)buff;
        fill_submat_d(ma_len, b, nr_, nc_, ptr, ctr, key, gauss, mathmode, next_ctr, nthreads);
        if (!gauss) {
            const double s3 = sqrt(3.0);
            for (int64_t i = 0; i < nr_ * nc_; ++i) b[i] = b[i] * s3;
        }
    } else {
        float *b = (float *)buff;
        fill_submat_s(ma_len, b, nr_, nc_, ptr, ctr, key, gauss, mathmode, next_ctr, nthreads);
        if (!gauss) {
            const float s3 = (float)sqrt(3.0);
            for (int64_t i = 0; i < nr_ * nc_; ++i) b[i] = b[i] * s3;
        }
    }
    if (layout != nat) {
        /* util::omatcopy from the natural layout into the requested one (:597-604) */
        size_t esz = dtype == 'd' ? 8 : 4;
        int64_t total = n_rows * n_cols;
        char *work = (char *)malloc((size_t)total * esz);
        memcpy(work, buff, (size_t)total * esz);
        int64_t irs_n = (nat == 'C') ? 1 : n_cols, ics_n = (nat == 'C') ? n_rows : 1;
        int64_t irs_r = (layout == 'C') ? 1 : n_cols, ics_r = (layout == 'C') ? n_rows : 1;
        for (int64_t i = 0; i < n_rows; ++i)
            for (int64_t j = 0; j < n_cols; ++j)
                memcpy((char *)buff + (size_t)(i * irs_r + j * ics_r) * esz,
                       work + (size_t)(i * irs_n + j * ics_n) * esz, esz);
        free(work);
    }
    return 0;
}

/* ======================================================================
 * Dense GEMM (stands in for blas::gemm at skge.hh:200,353).
 * ====================================================================== */
void orc_gemm_d(int64_t M, int64_t N, int64_t K, double alpha,
                const double *X, int64_t xrs, int64_t xcs,
                const double *Y, int64_t yrs, int64_t ycs,
                double beta, double *C, int64_t crs, int64_t ccs, int nthreads) {
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    #pragma omp parallel for collapse(2) schedule(static) num_threads(nthreads)
    for (int64_t j = 0; j < N; ++j) {
        for (int64_t i = 0; i < M; ++i) {
            double acc = 0.0;
            if (alpha != 0.0) {
                for (int64_t k = 0; k < K; ++k)
                    acc += X[i * xrs + k * xcs] * Y[k * yrs + j * ycs];
            }
            double *cij = &C[i * crs + j * ccs];
            *cij = (beta == 0.0) ? alpha * acc : alpha * acc + beta * (*cij);
        }
    }
}

static void strides_of(char layout, int64_t ld, int64_t *rs, int64_t *cs) {
    if (layout == 'C') { *rs = 1; *cs = ld; } else { *rs = ld; *cs = 1; }
}

/* Materialise submat(S) of the given shape in S's natural layout as doubles
 * (submatrix_as_blackbox, dense_skops.hh:677-688). */
static double *materialise_submat(int64_t S_rows, int64_t S_cols, char family, char major_axis,
                                  const uint32_t ctr[4], const uint32_t key[2],
                                  int64_t r, int64_t c, int64_t ro_s, int64_t co_s,
                                  int mathmode, int nthreads, char *lay_out) {
    char nat = dd_natural_layout(S_rows, S_cols, major_axis);
    double *buf = (double *)malloc((size_t)(r * c) * sizeof(double));
    uint32_t nxt[4];
    orc_fill_dense_unpacked(nat, S_rows, S_cols, family, major_axis, r, c, ro_s, co_s, 'd',
                            buf, ctr, key, nxt, mathmode, nthreads);
    *lay_out = nat;
    return buf;
}

/* Promote a strided float matrix view into a dense double copy, returned with
 * the same strides.  Only the referenced region is touched. */
static double *promote(const void *A, char dtype, int64_t rows, int64_t cols, int64_t rs, int64_t cs) {
    int64_t span = (rows - 1) * rs + (cols - 1) * cs + 1;
    double *out = (double *)malloc((size_t)span * sizeof(double));
    for (int64_t i = 0; i < rows; ++i)
        for (int64_t j = 0; j < cols; ++j) {
            int64_t p = i * rs + j * cs;
            out[p] = (dtype == 'd') ? ((const double *)A)[p] : (double)((const float *)A)[p];
        }
    return out;
}

/* lskge3 (skge.hh:154-202): B = alpha*op(submat(S))*op(A) + beta*B */
int orc_lskge3(char dtype, char layout, char opS, char opA, int64_t d, int64_t n, int64_t m,
               double alpha, int64_t S_rows, int64_t S_cols, char family, char major_axis,
               const uint32_t ctr[4], const uint32_t key[2], int64_t ro_s, int64_t co_s,
               const void *A, int64_t lda, double beta, void *B, int64_t ldb,
               int mathmode, int nthreads) {
    int64_t rS = (opS == 'N') ? d : m, cS = (opS == 'N') ? m : d;
    if (S_rows < rS + ro_s || S_cols < cS + co_s) return 1;
    int64_t rA = (opA == 'N') ? m : n, cA = (opA == 'N') ? n : m;
    if (layout == 'C') { if (lda < rA || ldb < d) return 1; }
    else               { if (lda < cA || ldb < n) return 1; }
    char slay;
    double *Sb = materialise_submat(S_rows, S_cols, family, major_axis, ctr, key, rS, cS,
                                    ro_s, co_s, mathmode, nthreads, &slay);
    int64_t srs, scs, ars, acs, brs, bcs;
    strides_of(slay, slay == 'C' ? rS : cS, &srs, &scs);
    strides_of(layout, lda, &ars, &acs);
    strides_of(layout, ldb, &brs, &bcs);
    /* op(S) strides */
    int64_t xrs = (opS == 'N') ? srs : scs, xcs = (opS == 'N') ? scs : srs;
    int64_t yrs = (opA == 'N') ? ars : acs, ycs = (opA == 'N') ? acs : ars;
    double *Ad = promote(A, dtype, rA, cA, ars, acs);
    if (dtype == 'd') {
        orc_gemm_d(d, n, m, alpha, Sb, xrs, xcs, Ad, yrs, ycs, beta, (double *)B, brs, bcs, nthreads);
    } else {
        double *Bd = promote(B, dtype, d, n, brs, bcs);
        orc_gemm_d(d, n, m, alpha, Sb, xrs, xcs, Ad, yrs, ycs, beta, Bd, brs, bcs, nthreads);
        for (int64_t i = 0; i < d; ++i)
            for (int64_t j = 0; j < n; ++j)
                ((float *)B)[i * brs + j * bcs] = (float)Bd[i * brs + j * bcs];
        free(Bd);
    }
    free(Ad);
    free(Sb);
    return 0;
}

/* rskge3 (skge.hh:307-355): B = alpha*op(A)*op(submat(S)) + beta*B */
int orc_rskge3(char dtype, char layout, char opA, char opS, int64_t m, int64_t d, int64_t n,
               double alpha, const void *A, int64_t lda,
               int64_t S_rows, int64_t S_cols, char family, char major_axis,
               const uint32_t ctr[4], const uint32_t key[2], int64_t ro_s, int64_t co_s,
               double beta, void *B, int64_t ldb, int mathmode, int nthreads) {
    int64_t rS = (opS == 'N') ? n : d, cS = (opS == 'N') ? d : n;
    if (S_rows < rS + ro_s || S_cols < cS + co_s) return 1;
    int64_t rA = (opA == 'N') ? m : n, cA = (opA == 'N') ? n : m;
    if (layout == 'C') { if (lda < rA || ldb < m) return 1; }
    else               { if (lda < cA || ldb < d) return 1; }
    char slay;
    double *Sb = materialise_submat(S_rows, S_cols, family, major_axis, ctr, key, rS, cS,
                                    ro_s, co_s, mathmode, nthreads, &slay);
    int64_t srs, scs, ars, acs, brs, bcs;
    strides_of(slay, slay == 'C' ? rS : cS, &srs, &scs);
    strides_of(layout, lda, &ars, &acs);
    strides_of(layout, ldb, &brs, &bcs);
    int64_t xrs = (opA == 'N') ? ars : acs, xcs = (opA == 'N') ? acs : ars;
    int64_t yrs = (opS == 'N') ? srs : scs, ycs = (opS == 'N') ? scs : srs;
    double *Ad = promote(A, dtype, rA, cA, ars, acs);
    if (dtype == 'd') {
        orc_gemm_d(m, d, n, alpha, Ad, xrs, xcs, Sb, yrs, ycs, beta, (double *)B, brs, bcs, nthreads);
    } else {
        double *Bd = promote(B, dtype, m, d, brs, bcs);
        orc_gemm_d(m, d, n, alpha, Ad, xrs, xcs, Sb, yrs, ycs, beta, Bd, brs, bcs, nthreads);
        for (int64_t i = 0; i < m; ++i)
            for (int64_t j = 0; j < d; ++j)
                ((float *)B)[i * brs + j * bcs] = (float)Bd[i * brs + j * bcs];
        free(Bd);
    }
    free(Ad);
    free(Sb);
    return 0;
}

/* ======================================================================
 * Sparse operators (sparse_skops.hh)
 * ====================================================================== */

/* compute_next_state(SparseDist, ...) (sparse_skops.hh:266-283) */
int orc_sparse_next_state(int64_t n_rows, int64_t n_cols, int64_t vec_nnz, char major_axis,
                          const uint32_t ctr[4], uint32_t next_ctr[4]) {
    if (n_rows <= 0 || n_cols <= 0 || vec_nnz <= 0) return 1;
    int64_t num_mavec, incrs;
    if (major_axis == 'S') {
        num_mavec = n_rows > n_cols ? n_rows : n_cols;
        incrs = vec_nnz;
    } else {
        num_mavec = n_rows < n_cols ? n_rows : n_cols;
        incrs = (int64_t)ceil((double)vec_nnz / 2.0);
    }
    memcpy(next_ctr, ctr, 16);
    orc_ctr_incr(next_ctr, (uint64_t)(num_mavec * incrs));
    return 0;
}

/* sparse::repeated_fisher_yates (sparse_skops.hh:51-106).  vec_work is
 * restored after each vector (:92-102), so each vector is a partial
 * Fisher-Yates shuffle of the identity; this restatement keeps the same
 * per-vector work array and restore loop. */
static void fisher_yates(const uint32_t ctr[4], const uint32_t key[2], int64_t vec_nnz,
                         int64_t dim_major, int64_t dim_minor, int64_t *idxs_major,
                         int64_t *idxs_minor, double *vals_d, float *vals_s,
                         uint32_t next_ctr[4]) {
    int64_t *vec_work = (int64_t *)malloc((size_t)dim_major * sizeof(int64_t));
    int64_t *pivots = (int64_t *)malloc((size_t)vec_nnz * sizeof(int64_t));
    for (int64_t j = 0; j < dim_major; ++j) vec_work[j] = j;
    for (int64_t i = 0; i < dim_minor; ++i) {
        int64_t offset = i * vec_nnz;
        uint32_t cw[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
        orc_ctr_incr(cw, (uint64_t)offset);
        for (int64_t j = 0; j < vec_nnz; ++j) {
            uint32_t rv[4];
            orc_philox4x32_10(cw, key, rv);
            int64_t ell = j + (int64_t)((int64_t)rv[0] % (dim_major - j));
            pivots[j] = ell;
            int64_t swap = vec_work[ell];
            vec_work[ell] = vec_work[j];
            vec_work[j] = swap;
            idxs_major[j + offset] = swap;
            double v = (rv[1] % 2 == 0) ? 1.0 : -1.0;
            if (vals_d) vals_d[j + offset] = v;
            if (vals_s) vals_s[j + offset] = (float)v;
            if (idxs_minor) idxs_minor[j + offset] = i;
            orc_ctr_incr(cw, 1);
        }
        for (int64_t j = 1; j <= vec_nnz; ++j) {
            int64_t jj = vec_nnz - j;
            int64_t swap = idxs_major[jj + offset];
            int64_t ell = pivots[jj];
            vec_work[jj] = vec_work[ell];
            vec_work[ell] = swap;
        }
    }
    free(pivots);
    free(vec_work);
    memcpy(next_ctr, ctr, 16);
    orc_ctr_incr(next_ctr, (uint64_t)(dim_minor * vec_nnz));
}

/* sample_indices_iid_uniform (util.hh:516-546) + uneg11_to_u01 (util.hh:474-477)
 * + laso_merge_long_axis_vector_coo_data (sparse_skops.hh:453-491).  Returns
 * the number of entries kept for this vector.  Duplicates are merged to
 * sqrt(count)*first_sign; the merged vector is written in first-occurrence
 * order (the reference iterates an std::unordered_map, whose order is
 * unspecified). */
static int64_t laso_vector(uint32_t c[4], const uint32_t key[2], int64_t n, int64_t k,
                           int64_t *idx, double *val) {
    uint32_t w[4];
    orc_philox4x32_10(c, key, w);
    int rv_index = 0;
    const double dN = (double)n;
    for (int64_t i = 0; i < k; ++i) {
        double u = ((double)orc_uneg11(w[rv_index]) + 1.0) / 2.0;
        idx[i] = (int64_t)(dN * u);
        rv_index += 1;
        val[i] = (orc_uneg11(w[rv_index]) >= 0.0f) ? 1.0 : -1.0;
        rv_index += 1;
        if (rv_index == 4) {
            orc_ctr_incr(c, 1);
            orc_philox4x32_10(c, key, w);
            rv_index = 0;
        }
    }
    if (rv_index > 0) orc_ctr_incr(c, 1);
    /* merge duplicates */
    int64_t out = 0;
    int any_dup = 0;
    for (int64_t i = 0; i < k && !any_dup; ++i)
        for (int64_t j = 0; j < i; ++j)
            if (idx[j] == idx[i]) { any_dup = 1; break; }
    if (!any_dup) return k;
    int64_t *ui = (int64_t *)malloc((size_t)k * sizeof(int64_t));
    double *cnt = (double *)malloc((size_t)k * sizeof(double));
    double *sg = (double *)malloc((size_t)k * sizeof(double));
    for (int64_t i = 0; i < k; ++i) {
        int64_t f = -1;
        for (int64_t j = 0; j < out; ++j) if (ui[j] == idx[i]) { f = j; break; }
        if (f >= 0) cnt[f] += 1.0;
        else { ui[out] = idx[i]; cnt[out] = 1.0; sg[out] = val[i]; out++; }
    }
    for (int64_t j = 0; j < out; ++j) { idx[j] = ui[j]; val[j] = sqrt(cnt[j]) * sg[j]; }
    free(ui); free(cnt); free(sg);
    return out;
}

/* fill_sparse_unpacked_nosub (sparse_skops.hh:515-565) */
int orc_fill_sparse(int64_t n_rows, int64_t n_cols, int64_t vec_nnz, char major_axis,
                    char dtype, int64_t *nnz_out, void *vals, int64_t *rows, int64_t *cols,
                    const uint32_t ctr[4], const uint32_t key[2], uint32_t next_ctr[4]) {
    if (n_rows <= 0 || n_cols <= 0 || vec_nnz <= 0) return 1;
    int64_t dim_major = (major_axis == 'S') ? (n_rows < n_cols ? n_rows : n_cols)
                                            : (n_rows > n_cols ? n_rows : n_cols);
    int64_t dim_minor = n_rows + n_cols - dim_major;
    if (vec_nnz > dim_major) return 1;
    int64_t *idxs_short = (n_rows <= n_cols) ? rows : cols;
    int64_t *idxs_long = (n_rows <= n_cols) ? cols : rows;
    if (major_axis == 'S') {
        fisher_yates(ctr, key, vec_nnz, dim_major, dim_minor, idxs_short, idxs_long,
                     dtype == 'd' ? (double *)vals : NULL, dtype == 's' ? (float *)vals : NULL,
                     next_ctr);
        *nnz_out = vec_nnz * dim_minor;
        return 0;
    }
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    int64_t *idx = (int64_t *)malloc((size_t)vec_nnz * sizeof(int64_t));
    double *val = (double *)malloc((size_t)vec_nnz * sizeof(double));
    int64_t total = 0;
    for (int64_t i = 0; i < dim_minor; ++i) {
        int64_t cnt = laso_vector(c, key, dim_major, vec_nnz, idx, val);
        for (int64_t j = 0; j < cnt; ++j) {
            idxs_long[total + j] = idx[j];
            idxs_short[total + j] = i;
            if (dtype == 'd') ((double *)vals)[total + j] = val[j];
            else ((float *)vals)[total + j] = (float)val[j];
        }
        total += cnt;
    }
    free(idx); free(val);
    memcpy(next_ctr, c, 16);
    *nnz_out = total;
    return 0;
}

/* ---- left_spmm for COO (spmm_dispatch.hh:52-179 -> coo_spmm_impl.hh:52-105) ---- */
typedef struct { int64_t r, c, p; } coo_ent;
static int cmp_csc(const void *a, const void *b) {
    const coo_ent *x = (const coo_ent *)a, *y = (const coo_ent *)b;
    if (x->c != y->c) return x->c < y->c ? -1 : 1;
    if (x->r != y->r) return x->r < y->r ? -1 : 1;
    return x->p < y->p ? -1 : (x->p > y->p);
}

#define DEF_LEFT_SPMM_COO(T, NAME)                                                         \
static void NAME(char layout, char opB, int64_t d, int64_t n, int64_t m, T alpha,          \
                 int64_t nnz, const T *vals, const int64_t *rows, const int64_t *cols,     \
                 int64_t ro_a, int64_t co_a, const T *B, int64_t ldb, T beta, T *C,         \
                 int64_t ldc, int nthreads) {                                               \
    int64_t crs, ccs; strides_of(layout, ldc, &crs, &ccs);                                 \
    /* beta scaling with util::safe_scal semantics (util.hh:55-62) */                      \
    for (int64_t i = 0; i < d; ++i)                                                        \
        for (int64_t j = 0; j < n; ++j) {                                                  \
            T *p = &C[i * crs + j * ccs];                                                  \
            *p = (beta == (T)0) ? (T)0 : beta * (*p);                                      \
        }                                                                                  \
    if (alpha == (T)0) return;                                                             \
    char layout_opB = (opB == 'N') ? layout : (layout == 'C' ? 'R' : 'C');                \
    int64_t brs, bcs; strides_of(layout_opB, ldb, &brs, &bcs);                             \
    coo_ent *e = (coo_ent *)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(coo_ent));         \
    int64_t ne = 0;                                                                        \
    for (int64_t i = 0; i < nnz; ++i) {                                                    \
        int64_t r = rows[i] - ro_a, c = cols[i] - co_a;                                    \
        if (0 <= r && r < d && 0 <= c && c < m) { e[ne].r = r; e[ne].c = c; e[ne].p = i; ne++; } \
    }                                                                                      \
    qsort(e, (size_t)ne, sizeof(coo_ent), cmp_csc);                                        \
    int kib = (layout_opB == 'R' && layout == 'R');                                        \
    _Pragma("omp parallel for schedule(static) num_threads(nthreads)")                     \
    for (int64_t j = 0; j < n; ++j) {                                                      \
        for (int64_t q = 0; q < ne; ++q) {                                                 \
            T v = vals[e[q].p];                                                            \
            T bval = B[e[q].c * brs + j * bcs];                                            \
            volatile T prod;                                                               \
            if (kib) { volatile T sc = alpha * v; prod = sc * bval; }                      \
            else     { volatile T sc = alpha * bval; prod = v * sc; }                      \
            T *p = &C[e[q].r * crs + j * ccs];                                             \
            *p = *p + prod;                                                                \
        }                                                                                  \
    }                                                                                      \
    free(e);                                                                               \
}
DEF_LEFT_SPMM_COO(double, left_spmm_coo_d)
DEF_LEFT_SPMM_COO(float, left_spmm_coo_s)

int orc_left_spmm_coo(char dtype, char layout, char opA, char opB, int64_t d, int64_t n, int64_t m,
                      double alpha, int64_t A_rows, int64_t A_cols, int64_t nnz,
                      const void *vals, const int64_t *rows, const int64_t *cols,
                      int64_t ro_a, int64_t co_a, const void *B, int64_t ldb,
                      double beta, void *C, int64_t ldc, int nthreads) {
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    if (opA == 'T') {
        /* transpose view: swap index arrays and offsets (spmm_dispatch.hh:72-90) */
        const int64_t *t = rows; rows = cols; cols = t;
        int64_t ts = A_rows; A_rows = A_cols; A_cols = ts;
        ts = ro_a; ro_a = co_a; co_a = ts;
    }
    if (A_rows < d || A_cols < m) return 1;
    int64_t rB = (opB == 'N') ? m : n, cB = (opB == 'N') ? n : m;
    if (layout == 'C') { if (ldb < rB || ldc < d) return 1; }
    else               { if (ldc < n || ldb < cB) return 1; }
    if (dtype == 'd')
        left_spmm_coo_d(layout, opB, d, n, m, alpha, nnz, (const double *)vals, rows, cols,
                        ro_a, co_a, (const double *)B, ldb, beta, (double *)C, ldc, nthreads);
    else
        left_spmm_coo_s(layout, opB, d, n, m, (float)alpha, nnz, (const float *)vals, rows, cols,
                        ro_a, co_a, (const float *)B, ldb, (float)beta, (float *)C, ldc, nthreads);
    return 0;
}

/* ---- left_spmm for CSR (csr_spmm_impl.hh:51-155) ---- */
#define DEF_LEFT_SPMM_CSR(T, NAME)                                                         \
static void NAME(char layout, char opB, int64_t d, int64_t n, T alpha, const T *vals,      \
                 const int64_t *rowptr, const int64_t *colidxs, const T *B, int64_t ldb,    \
                 T beta, T *C, int64_t ldc, int nthreads) {                                 \
    int64_t crs, ccs; strides_of(layout, ldc, &crs, &ccs);                                 \
    for (int64_t i = 0; i < d; ++i)                                                        \
        for (int64_t j = 0; j < n; ++j) {                                                  \
            T *p = &C[i * crs + j * ccs];                                                  \
            *p = (beta == (T)0) ? (T)0 : beta * (*p);                                      \
        }                                                                                  \
    if (alpha == (T)0) return;                                                             \
    char layout_opB = (opB == 'N') ? layout : (layout == 'C' ? 'R' : 'C');                \
    int64_t brs, bcs; strides_of(layout_opB, ldb, &brs, &bcs);                             \
    int ikb = (layout_opB == 'R' && layout == 'R');                                        \
    _Pragma("omp parallel for schedule(static) num_threads(nthreads)")                     \
    for (int64_t i = 0; i < d; ++i) {                                                      \
        for (int64_t j = 0; j < n; ++j) {                                                  \
            T *p = &C[i * crs + j * ccs];                                                  \
            if (ikb) {                                                                     \
                for (int64_t l = rowptr[i]; l < rowptr[i + 1]; ++l) {                      \
                    volatile T sc = alpha * vals[l];                                       \
                    volatile T prod = sc * B[colidxs[l] * brs + j * bcs];                  \
                    *p = *p + prod;                                                        \
                }                                                                          \
            } else {                                                                       \
                T diff = 0;                                                                \
                for (int64_t l = rowptr[i]; l < rowptr[i + 1]; ++l) {                      \
                    volatile T prod = vals[l] * B[colidxs[l] * brs + j * bcs];             \
                    diff = diff + prod;                                                    \
                }                                                                          \
                volatile T sd = alpha * diff;                                              \
                *p = *p + sd;                                                              \
            }                                                                              \
        }                                                                                  \
    }                                                                                      \
}
DEF_LEFT_SPMM_CSR(double, left_spmm_csr_d)
DEF_LEFT_SPMM_CSR(float, left_spmm_csr_s)

int orc_left_spmm_csr(char dtype, char layout, char opB, int64_t d, int64_t n, int64_t m,
                      double alpha, const void *vals, const int64_t *rowptr,
                      const int64_t *colidxs, const void *B, int64_t ldb,
                      double beta, void *C, int64_t ldc, int nthreads) {
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    int64_t rB = (opB == 'N') ? m : n, cB = (opB == 'N') ? n : m;
    if (layout == 'C') { if (ldb < rB || ldc < d) return 1; }
    else               { if (ldc < n || ldb < cB) return 1; }
    if (dtype == 'd')
        left_spmm_csr_d(layout, opB, d, n, alpha, (const double *)vals, rowptr, colidxs,
                        (const double *)B, ldb, beta, (double *)C, ldc, nthreads);
    else
        left_spmm_csr_s(layout, opB, d, n, (float)alpha, (const float *)vals, rowptr, colidxs,
                        (const float *)B, ldb, (float)beta, (float *)C, ldc, nthreads);
    return 0;
}

/* Sampled dense-dense product: the gradient of B = alpha * A * Sm (A CSR m x n,
 * Sm dense n x d) with respect to A's stored values,
 *   dvals[k] = alpha * sum_j Bbar[i_k, j] * Sm[c_k, j],
 * accumulated in double in ascending j.  No reference counterpart (the
 * reference has no adjoints); it is the definition of that gradient.
 * Bbar(i, j) at Bbar[i*brs + j*bcs], Sm(c, j) at Sm[c*srs + j*scs]. */
int orc_sddmm_csr(char dtype, int64_t m, int64_t d, double alpha, const int64_t *rowptr,
                  const int64_t *colidxs, const void *Bbar, int64_t brs, int64_t bcs,
                  const void *Sm, int64_t srs, int64_t scs, void *dvals, int nthreads) {
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    #pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
    for (int64_t i = 0; i < m; ++i) {
        for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k) {
            const int64_t c = colidxs[k];
            double acc = 0.0;
            for (int64_t j = 0; j < d; ++j) {
                double b, sv;
                if (dtype == 'd') {
                    b = ((const double *)Bbar)[i * brs + j * bcs];
                    sv = ((const double *)Sm)[c * srs + j * scs];
                } else {
                    b = ((const float *)Bbar)[i * brs + j * bcs];
                    sv = ((const float *)Sm)[c * srs + j * scs];
                }
                acc += b * sv;
            }
            if (dtype == 'd') ((double *)dvals)[k] = alpha * acc;
            else ((float *)dvals)[k] = (float)(alpha * acc);
        }
    }
    return 0;
}

/* ---- index sampling (util.hh:459-560), sequential as the reference ---- */

/* sample_indices_iid_uniform (util.hh:516-560): one rv array of 4 uneg11
 * values per counter, regenerated when exhausted; with signs the array length
 * is rounded down to even (2 here: no change) and each sample takes two
 * values.  rademachers == NULL selects the unsigned form (util.hh:552-560).
 * Writes the returned state's counter to next_ctr. */
void orc_sample_indices_iid_uniform(int64_t n, int64_t k, int64_t *samples, double *rademachers,
                                    const uint32_t ctr[4], const uint32_t key[2], uint32_t next_ctr[4]) {
    uint32_t c[4], w[4];
    memcpy(c, ctr, 16);
    orc_philox4x32_10(c, key, w);
    const int len_c = 4;
    int rv_index = 0;
    const double dN = (double)n;
    for (int64_t i = 0; i < k; ++i) {
        const double u = ((double)orc_uneg11(w[rv_index]) + 1.0) / 2.0;   /* uneg11_to_u01<double> */
        samples[i] = (int64_t)((double)(int64_t)dN * u);                 /* (sint_t) dN * u, util.hh:533 */
        rv_index += 1;
        if (rademachers) {
            rademachers[i] = (orc_uneg11(w[rv_index]) >= 0.0f) ? 1.0 : -1.0;
            rv_index += 1;
        }
        if (rv_index == len_c) {
            orc_ctr_incr(c, 1);
            orc_philox4x32_10(c, key, w);
            rv_index = 0;
        }
    }
    if (rv_index > 0) orc_ctr_incr(c, 1);
    memcpy(next_ctr, c, 16);
}

/* sample_indices_iid (util.hh:493-514): u = uneg11_to_u01<T>(rv) in T, then
 * std::lower_bound(cdf, cdf + n, u) - cdf. */
#define DEF_IID_CDF(T, NAME)                                                                   \
static void NAME(int64_t n, const T *cdf, int64_t k, int64_t *samples, const uint32_t ctr[4], \
                 const uint32_t key[2], uint32_t next_ctr[4]) {                                \
    uint32_t c[4], w[4];                                                                       \
    memcpy(c, ctr, 16);                                                                        \
    orc_philox4x32_10(c, key, w);                                                              \
    int rv_index = 0;                                                                          \
    for (int64_t i = 0; i < k; ++i) {                                                          \
        volatile T s = (T)orc_uneg11(w[rv_index]) + (T)1.0;                                    \
        const T u = s / (T)2.0;                                                                \
        int64_t lo = 0, len = n;                                                               \
        while (len > 0) {                                                                      \
            const int64_t half = len / 2, mid = lo + half;                                     \
            if (cdf[mid] < u) { lo = mid + 1; len -= half + 1; }                               \
            else len = half;                                                                   \
        }                                                                                      \
        samples[i] = lo;                                                                       \
        rv_index += 1;                                                                         \
        if (rv_index == 4) {                                                                   \
            orc_ctr_incr(c, 1);                                                                \
            orc_philox4x32_10(c, key, w);                                                      \
            rv_index = 0;                                                                      \
        }                                                                                      \
    }                                                                                          \
    if (rv_index > 0) orc_ctr_incr(c, 1);                                                      \
    memcpy(next_ctr, c, 16);                                                                   \
}
DEF_IID_CDF(double, iid_cdf_d)
DEF_IID_CDF(float, iid_cdf_s)

void orc_sample_indices_iid(char dtype, int64_t n, const void *cdf, int64_t k, int64_t *samples,
                            const uint32_t ctr[4], const uint32_t key[2], uint32_t next_ctr[4]) {
    if (dtype == 'd') iid_cdf_d(n, (const double *)cdf, k, samples, ctr, key, next_ctr);
    else iid_cdf_s(n, (const float *)cdf, k, samples, ctr, key, next_ctr);
}

/* weights_to_cdf (util.hh:459-472).  Returns 0, or 1 + i when w[i] is below
 * error_if_below (w[0..i) then hold the partial sums), or -1 when the total is
 * below sqrt(n) * epsilon (w holds the unscaled partial sums). */
#define DEF_W2CDF(T, NAME, EPS)                                                                \
static int64_t NAME(int64_t n, T *w, T error_if_below) {                                       \
    volatile T sum = (T)0.0;                                                                   \
    for (int64_t i = 0; i < n; ++i) {                                                          \
        T val = w[i];                                                                          \
        if (!(val >= error_if_below)) return 1 + i;                                            \
        val = val > (T)0.0 ? val : (T)0.0;                                                     \
        sum = sum + val;                                                                       \
        w[i] = sum;                                                                            \
    }                                                                                          \
    if (!(sum >= (T)sqrt((double)n) * (EPS))) return -1;                                       \
    const T inv = (T)1.0 / sum;                                                                \
    for (int64_t i = 0; i < n; ++i) w[i] = w[i] * inv;                                         \
    return 0;                                                                                  \
}
DEF_W2CDF(double, w2cdf_d, DBL_EPSILON)
DEF_W2CDF(float, w2cdf_s, FLT_EPSILON)

/* Zipf column weights of the library's power-law sparse test data (no
 * reference counterpart; the weights feed weights_to_cdf, util.hh:459-472,
 * and sample_indices_iid, util.hh:493-514).  libm pow, as the library's host
 * side evaluates them. */
void orc_power_law_weights(int64_t n, double exponent, double *w) {
    const double e = -1.0 / (exponent - 1.0);
    for (int64_t c = 0; c < n; ++c) w[c] = pow((double)(c + 1), e);
}

int64_t orc_weights_to_cdf(char dtype, int64_t n, void *w, double error_if_below) {
    if (dtype == 'd') return w2cdf_d(n, (double *)w, error_if_below);
    return w2cdf_s(n, (float *)w, (float)error_if_below);
}
