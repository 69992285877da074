// rbx_solve.hip -- sketch-and-solve least squares on the device, the
// downstream use of a left sketch (the reference's pattern in
// examples/total-least-squares/tls_dense_skop.cc:139-186, with a QR solve of
// the sketched system in place of its SVD):
//   x = argmin_x || S (A x - b) ||_2,   S = submat(S) (d x m), A m x n, b m,
// SA and Sb by the fused sketch (the operator is never stored), then the QR
// of the d x n sketched system and the triangular solve in FP64 with
// rocSOLVER / rocBLAS for both precisions.  The QR is CholeskyQR2 (Yamamoto,
// Nakatsukasa, Yanagisawa, Fukaya 2015: R1 = chol(Y^T Y), Q1 = Y R1^{-1},
// R2 = chol(Q1^T Q1), Q = Q1 R2^{-1}, R = R2 R1; explicit Q, a few large
// BLAS-3 calls) when the second Gram matrix shows Q1 orthonormal to 1e-3
// (|R2_ii - 1| <= 1e-3, both Cholesky factorizations succeeded), i.e. a
// well-conditioned sketched system; otherwise Householder QR (geqrf), which
// is stable for any conditioning.  RBX_SOLVE_QR=householder in the
// environment selects Householder QR for every call.  The adjoint gives the
// gradient of
// <xbar, x(A, b)>:
//   u = R^{-T} xbar,  w = Q [u; 0],  z = R^{-1} u,  r = Sb - SA x,
//   bbar = S^T w,     Abar = S^T (r z^T - w x^T).
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "rbx_gemm_common.hh"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

namespace rbx {

// one rocBLAS handle per (device, stream), kept in the stream's workspace and
// bound to that stream once
static void destroy_blas_handle(void *h) { (void)rocblas_destroy_handle((rocblas_handle)h); }

static rocblas_handle blas_handle(hipStream_t st) {
    void (*destroy)(void *) = nullptr;
    void *&h = stream_blas_handle(st, &destroy);
    if (h == nullptr) {
        rocblas_handle nh = nullptr;
        if (rocblas_create_handle(&nh) != rocblas_status_success)
            RBX_FAIL(RBX_HIP_ERROR, "sketch_solve: rocblas_create_handle failed");
        if (rocblas_set_stream(nh, st) != rocblas_status_success) {
            (void)rocblas_destroy_handle(nh);
            RBX_FAIL(RBX_HIP_ERROR, "sketch_solve: rocblas_set_stream failed");
        }
        h = nh;
        set_stream_blas_destroy(st, destroy_blas_handle);
    }
    return (rocblas_handle)h;
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

// zero the strictly lower triangle of an n x n column-major matrix
__global__ __launch_bounds__(256) void zero_lower_kernel(double *R, int64_t n) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n * n; e += (int64_t)gridDim.x * blockDim.x)
        if (e % n > e / n) R[e] = 0.0;
}

// v from lane `lane` of the wave (a uniform, compile-time lane index)
__device__ __forceinline__ double readlane_f64(double v, int lane) {
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, lane);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), lane);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Upper Cholesky A = R^T R of an n x n column-major FP64 matrix, n <= 256,
// in one 1024-thread workgroup (the upper triangle is read and R written over
// it; the strict lower triangle is untouched), with 16 x 16 tiles in the
// v_mfma_f64_16x16x4f64 accumulator layout (lane l holds rows l / 16 + 4 r,
// r = 0..3, of column l % 16; rows and columns past n are identity padding).
// Wave 0 keeps the diagonal tiles in LDS and works one step ahead: while the
// other waves update the trailing tiles of step s, it updates diagonal tile
// s + 1 and factors it, R_{s+1} and W_{s+1} = R_{s+1}^{-T} (pivot k read from
// lane k, 1 / sqrt by the hardware estimate and two Newton steps, row k
// broadcast with v_readlane; then W by substitution).  Waves 1..15 hold the
// off-diagonal tiles in registers (slot q of wave w: tile q * 15 + w - 1 of
// the row-major strict upper triangle) and per step s
//   1. form the tile row X_b = W_s B_b (an MFMA product) and publish it to LDS;
//   2. update the diagonal tiles past s + 1 in LDS, and their trailing tiles
//      T_ab -= X_a^T X_b, with v_mfma_f64_16x16x4f64.
// Global memory is read once and written once; two barriers per step (the
// tile row's LDS buffer alternates between steps; W of every step is kept).
// info = 0, or k + 1 when pivot k is not positive (R then incomplete).
// When Rp / Wp are given, the kernel also writes, for trsm_tiles_kernel, the
// strictly upper tiles of R (tile (k, j), k < j, at Rp + 256 (j (j - 1) / 2 +
// k)) and W_s = R_ss^{-T} (at Wp + 256 s), each as 4 x 64 doubles: element
// [ks][lane] is what lane `lane` passes to v_mfma_f64_16x16x4f64 at K-step ks
// (R tiles in the accumulator layout, which is the A-operand layout of R^T;
// W in the A-operand layout).
constexpr int CH_MAX = 256, CH_THR = 1024, CH_WAVES = CH_THR / 64, CH_T = CH_MAX / 16;
constexpr int CH_OFF = CH_T * (CH_T - 1) / 2;                        // 120 off-diagonal tiles
constexpr int CH_SLOTS = (CH_OFF + CH_WAVES - 2) / (CH_WAVES - 1);  // 8 per off-diagonal wave
constexpr int CH_PLD = CH_MAX + 16;                                 // LDS row of the tile row (16 doubles of pad)
typedef double chf64x4 __attribute__((ext_vector_type(4)));

// wave 0: factor the diagonal tile D (LDS, [row][col], upper part used) in
// place, R over its upper triangle (zeros below), and W = R^{-T} into Wl
// ([row][k]); returns the bad pivot index + 1, or 0
__device__ __forceinline__ int chol_factor_tile(double (*D)[17], double (*Wl)[17], int lane) {
    const int l = lane;
    int bad = 0;
    double c[16];  // lane l < 16: column l of the tile
#pragma unroll
    for (int i = 0; i < 16; ++i) c[i] = (l < 16 && i <= l) ? D[i][l] : (i == l ? 1.0 : 0.0);
    double dinv = 1.0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const double p = readlane_f64(c[k], k);
        if (bad == 0 && !(p > 0.0)) bad = k + 1;
        double y = __builtin_amdgcn_rsq(p);
        y = fma(0.5 * y, fma(-p, y * y, 1.0), y);
        y = fma(0.5 * y, fma(-p, y * y, 1.0), y);
        if (l == k) dinv = y;
        c[k] = l == k ? p * y : c[k] * y;
#pragma unroll
        for (int i = 0; i < 16; ++i)
            if (i > k) c[i] -= readlane_f64(c[k], i) * c[k];
    }
    // column l of R^{-1}: x = R^{-1} e_l, from the bottom (R(i, k) = lane
    // k's c[i], 1 / R(k, k) = lane k's dinv)
    double x[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) x[i] = i == l ? 1.0 : 0.0;
#pragma unroll
    for (int k = 15; k >= 0; --k) {
        x[k] *= readlane_f64(dinv, k);
#pragma unroll
        for (int i = 0; i < 16; ++i)
            if (i < k) x[i] -= readlane_f64(c[i], k) * x[k];
    }
    if (l < 16)
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            D[i][l] = i <= l ? c[i] : 0.0;  // R
            Wl[l][i] = x[i];                // W[l][i] = (R^{-1})[i][l]
        }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    return bad;
}

struct CholShared {
    double Dg[CH_T][16][17];  // the diagonal tiles (wave 0's)
    double Wl[CH_T][16][17];  // W_s = R_ss^{-T}, kept for every s (the tile solve's operands)
    double Pn[2][16][CH_PLD];
    int failed;
};

// wave 0's part: the diagonal tiles in LDS, updated and factored one step ahead
__device__ __forceinline__ void chol_diag_role(CholShared &sh, double *A, int n, int *info, int T, double *Wp) {
    const int lane = threadIdx.x & 63, lr = lane >> 4, lc = lane & 15;
    auto put_w = [&](int t) {
        if (Wp != nullptr)
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) Wp[t * 256 + ks * 64 + lane] = sh.Wl[t][lc][4 * ks + lr];
    };
    for (int a = 0; a < T; ++a)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 16 * a + lr + 4 * r, col = 16 * a + lc;
            double v = row == col ? 1.0 : 0.0;  // identity padding past n
            if (row < n && col < n && row <= col) v = A[row + (int64_t)col * n];
            sh.Dg[a][lr + 4 * r][lc] = v;
        }
    if (lane == 0) sh.failed = 0;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    {
        const int bad = chol_factor_tile(sh.Dg[0], sh.Wl[0], lane);
        if (bad && lane == 0) {
            sh.failed = 1;
            *info = bad;
        }
    }
    __syncthreads();
    for (int s = 0; s < T; ++s) {
        if (sh.failed) return;  // every wave reads the flag after the same barrier
        __syncthreads();        // the tile row published
        double(*P)[CH_PLD] = sh.Pn[s & 1];
        const int c0 = 16 * (s + 1);
        if (s + 1 < T) {
            // the critical path: ahead of the trailing updates on this SIMD
            __builtin_amdgcn_s_setprio(3);
            chf64x4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) {
                const double v = P[4 * ks + lr][16 * (s + 1) - c0 + lc];
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(v, v, acc, 0, 0, 0);
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) sh.Dg[s + 1][lr + 4 * r][lc] -= acc[r];
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            const int bad = chol_factor_tile(sh.Dg[s + 1], sh.Wl[s + 1], lane);
            if (bad && lane == 0) {
                sh.failed = 1;
                *info = 16 * (s + 1) + bad;
            }
            __builtin_amdgcn_s_setprio(0);
        }
        __syncthreads();
    }
    for (int a = 0; a < T; ++a) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 16 * a + lr + 4 * r, col = 16 * a + lc;
            if (row < n && col < n && row <= col) A[row + (int64_t)col * n] = sh.Dg[a][lr + 4 * r][lc];
        }
        put_w(a);
    }
    if (lane == 0) *info = 0;
}

// waves 1..15: the off-diagonal tiles in registers
__device__ __forceinline__ void chol_offdiag_role(CholShared &sh, double *A, int n, int T, int wave, double *Rp) {
    const int lane = threadIdx.x & 63, lr = lane >> 4, lc = lane & 15;
    const int noff = T * (T - 1) / 2;
    int ti_s[CH_SLOTS], tj_s[CH_SLOTS];
    double tile[CH_SLOTS][4];
#pragma unroll
    for (int q = 0; q < CH_SLOTS; ++q) {
        int a = -1, b = -1;
        const int e = q * (CH_WAVES - 1) + wave - 1;  // off-diagonal tile e, row-major
        if (e < noff) {
            int aa = 0, rem = e;
            while (rem >= T - 1 - aa) {  // row aa holds T - 1 - aa off-diagonal tiles
                rem -= T - 1 - aa;
                ++aa;
            }
            a = aa;
            b = aa + 1 + rem;
        }
        ti_s[q] = a;
        tj_s[q] = b;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 16 * a + lr + 4 * r, col = 16 * b + lc;
            tile[q][r] = (a >= 0 && row < n && col < n) ? A[row + (int64_t)col * n] : 0.0;
        }
    }
    __syncthreads();
    for (int s = 0; s < T; ++s) {
        if (sh.failed) return;
        double(*P)[CH_PLD] = sh.Pn[s & 1];
        const int c0 = 16 * (s + 1);
        // the tile row: X = W_s B, published
        double wa[4];
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) wa[ks] = sh.Wl[s][lc][4 * ks + lr];
#pragma unroll
        for (int q = 0; q < CH_SLOTS; ++q)
            if (ti_s[q] == s) {
                chf64x4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int ks = 0; ks < 4; ++ks)
                    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(wa[ks], tile[q][ks], acc, 0, 0, 0);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    tile[q][r] = acc[r];
                    P[lr + 4 * r][16 * tj_s[q] - c0 + lc] = acc[r];
                }
            }
        __syncthreads();
        // the diagonal tiles past the next one (in LDS; wave 0 takes tile s + 1)
        for (int a = s + 2; a < T; ++a)
            if (1 + a % (CH_WAVES - 1) == wave) {
                chf64x4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int ks = 0; ks < 4; ++ks) {
                    const double v = P[4 * ks + lr][16 * a - c0 + lc];
                    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(v, v, acc, 0, 0, 0);
                }
#pragma unroll
                for (int r = 0; r < 4; ++r) sh.Dg[a][lr + 4 * r][lc] -= acc[r];
            }
        // trailing tiles -= X_a^T X_b
#pragma unroll
        for (int q = 0; q < CH_SLOTS; ++q) {
            if (ti_s[q] > s) {
                const int a = ti_s[q], b = tj_s[q];
                chf64x4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int ks = 0; ks < 4; ++ks) {
                    const double av = P[4 * ks + lr][16 * a - c0 + lc];
                    const double bv = P[4 * ks + lr][16 * b - c0 + lc];
                    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
                }
#pragma unroll
                for (int r = 0; r < 4; ++r) tile[q][r] -= acc[r];
            }
            __builtin_amdgcn_sched_barrier(0);  // one slot's accumulators live at a time
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < CH_SLOTS; ++q) {
        if (ti_s[q] < 0) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 16 * ti_s[q] + lr + 4 * r, col = 16 * tj_s[q] + lc;
            if (row < n && col < n) A[row + (int64_t)col * n] = tile[q][r];
        }
        if (Rp != nullptr) {
            double *rp = Rp + 256 * (tj_s[q] * (tj_s[q] - 1) / 2 + ti_s[q]);
#pragma unroll
            for (int r = 0; r < 4; ++r) rp[r * 64 + lane] = tile[q][r];
        }
    }
}

__global__ __launch_bounds__(CH_THR) void chol_tiles_kernel(double *A, int n, int *info, double *Rp, double *Wp) {
    __shared__ __attribute__((aligned(16))) CholShared sh;
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const int T = (n + 15) / 16;
    if (wave == 0) chol_diag_role(sh, A, n, info, T, Wp);
    else chol_offdiag_role(sh, A, n, T, wave, Rp);
}

// X R = B in place (X = B R^{-1}) for the d x n column-major B (leading
// dimension ldb), with R the factor chol_tiles_kernel just formed, given as
// its Rp / Wp tiles (n <= CH_MAX).  One wave per 16 rows of B, independent of
// the others: with Y = X^T (that wave's 16 columns of it), R^T Y = B^T is
// solved by tile rows, Y_j = W_j (B^T_j - sum_{k<j} R_kj^T Y_k), every product
// a v_mfma_f64_16x16x4f64 and every Y_k held in registers in the accumulator
// layout, which is also the B-operand layout the later products need.  The
// sum over k runs in two chains (even and odd k) added at the end.
// (TT = (n + 15) / 16 tiles, a template parameter so that every Y index is
// a constant and Y stays in registers)
template <int TT>
__global__ __launch_bounds__(64) void trsm_tiles_kernel(double *B, int64_t ldb, int d, int n, const double *Rp,
                                                        const double *Wp) {
    const int lane = threadIdx.x, lr = lane >> 4, lc = lane & 15;
    const int64_t row = (int64_t)blockIdx.x * 16 + lc;  // this lane's row of B
    const bool rok = row < d;
    double Y[TT][4];
#pragma unroll
    for (int j = 0; j < TT; ++j) {
        chf64x4 acc0, acc1 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int col = 16 * j + lr + 4 * r;
            acc0[r] = (rok && col < n) ? B[row + (int64_t)col * ldb] : 0.0;
        }
        // tile row j of R^T: all of its operands are loaded before the first
        // product (one memory latency per tile row, not one per product)
        double ra[TT > 1 ? TT - 1 : 1][4];
        double w4[4];
#pragma unroll
        for (int k = 0; k < j; ++k) {
            const double *rp = Rp + 256 * (j * (j - 1) / 2 + k);
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) ra[k][ks] = rp[ks * 64 + lane];
        }
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) w4[ks] = Wp[256 * j + ks * 64 + lane];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int k = 0; k < j; ++k) {
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) {
                if (k & 1) acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(-ra[k][ks], Y[k][ks], acc1, 0, 0, 0);
                else acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(-ra[k][ks], Y[k][ks], acc0, 0, 0, 0);
            }
        }
        chf64x4 acc = acc0 + acc1;
        chf64x4 y = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) y = __builtin_amdgcn_mfma_f64_16x16x4f64(w4[ks], acc[ks], y, 0, 0, 0);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            Y[j][r] = y[r];
            const int col = 16 * j + lr + 4 * r;
            if (rok && col < n) B[row + (int64_t)col * ldb] = y[r];
        }
    }
}

template <int TT>
static void launch_trsm_tiles(double *B, int64_t d, int64_t n, const double *Rp, const double *Wp, hipStream_t st) {
    if ((n + 15) / 16 == TT) {
        hipLaunchKernelGGL(trsm_tiles_kernel<TT>, dim3((unsigned)((d + 15) / 16)), dim3(64), 0, st, B, d, (int)d, (int)n,
                           Rp, Wp);
        RBX_HIP(hipGetLastError());
    } else if constexpr (TT > 1) {
        launch_trsm_tiles<TT - 1>(B, d, n, Rp, Wp, st);
    }
}

// R x = y (TRANS false) or R^T x = y (TRANS true) in place, R n x n upper
// triangular (column-major, leading dimension ldr, non-unit diagonal),
// n <= TS_MAX, one 256-thread workgroup: y in LDS, 16-row blocks in the
// order of the substitution.  Per block, lane i < 16 of wave 0 holds row i
// of the block's triangle (R's row, or R's column when transposed) and
// substitutes with x_k broadcast by v_readlane (1 / R(k, k) precomputed per
// lane); then every thread of a remaining row subtracts the block's 16
// products from its y, the R values loaded before the block's solve.
constexpr int TS_MAX = 256;
template <bool TRANS>
__global__ __launch_bounds__(256) void trsv_upper_kernel(const double *R, int64_t ldr, double *y, int n) {
    __shared__ double ys[TS_MAX];
    __shared__ double xs[16];
    const int tid = threadIdx.x;
    if (tid < n) ys[tid] = y[tid];
    __syncthreads();
    const int nblk = (n + 15) / 16;
    for (int bi = 0; bi < nblk; ++bi) {
        const int kb = TRANS ? bi : nblk - 1 - bi;
        const int k0 = 16 * kb, nk = n - k0 < 16 ? n - k0 : 16;
        // rows still to update: above the block (back substitution), below it (forward)
        const bool upd = TRANS ? (tid >= k0 + nk && tid < n) : (tid < k0);
        double rv[16];
#pragma unroll
        for (int j = 0; j < 16; ++j)
            rv[j] = (upd && j < nk) ? (TRANS ? R[(k0 + j) + (int64_t)tid * ldr] : R[tid + (int64_t)(k0 + j) * ldr]) : 0.0;
        if (tid < 64) {
            const int i = tid;
            double r[16];  // row i of the block's triangle T (T = R_bb, or R_bb^T)
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const bool in = i < nk && j < nk && (TRANS ? j <= i : j >= i);
                r[j] = in ? (TRANS ? R[(k0 + j) + (int64_t)(k0 + i) * ldr] : R[(k0 + i) + (int64_t)(k0 + j) * ldr])
                          : (i == j ? 1.0 : 0.0);
            }
            double rii = 1.0;
#pragma unroll
            for (int j = 0; j < 16; ++j)
                if (j == i) rii = r[j];
            const double dinv = 1.0 / rii;
            double b = i < nk ? ys[k0 + i] : 0.0, x = 0.0;
#pragma unroll
            for (int t = 0; t < 16; ++t) {
                const int k = TRANS ? t : 15 - t;
                const double xk = readlane_f64(b * dinv, k);  // lane k's b is final here
                if (i == k) x = xk;
                b -= r[k] * xk;  // rows not yet solved; finished lanes keep x
            }
            if (i < nk) {
                ys[k0 + i] = x;
                xs[i] = x;
            }
        }
        __syncthreads();
        if (upd) {
            double v = ys[tid];
#pragma unroll
            for (int j = 0; j < 16; ++j) v -= rv[j] * xs[j];
            ys[tid] = v;
        }
        __syncthreads();
    }
    if (tid < n) y[tid] = ys[tid];
}

// x = op(R)^{-1} x in place for the solve's triangular factor: the kernel
// above for n <= TS_MAX, rocBLAS above that.
static void trsv_upper(rocblas_handle h, bool trans, int64_t n, const double *R, int64_t ldr, double *x,
                       hipStream_t st) {
    if (n <= TS_MAX) {
        if (trans)
            hipLaunchKernelGGL(trsv_upper_kernel<true>, dim3(1), dim3(256), 0, st, R, ldr, x, (int)n);
        else
            hipLaunchKernelGGL(trsv_upper_kernel<false>, dim3(1), dim3(256), 0, st, R, ldr, x, (int)n);
        RBX_HIP(hipGetLastError());
        return;
    }
    blas_ok(rocblas_dtrsv(h, rocblas_fill_upper, trans ? rocblas_operation_transpose : rocblas_operation_none,
                          rocblas_diagonal_non_unit, (rocblas_int)n, R, (rocblas_int)ldr, x, 1),
            "trsv");
}

// ok[0] = 1 when both Cholesky factorizations succeeded and every diagonal
// entry of R2 is within 1e-3 of 1 (one workgroup)
__global__ __launch_bounds__(256) void cholqr_check_kernel(const double *R2, int64_t n, const int *info, int *ok) {
    __shared__ int bad;
    if (threadIdx.x == 0) bad = (info[0] != 0) || (info[1] != 0);
    __syncthreads();
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x)
        if (!(fabs(R2[i * n + i] - 1.0) <= 1e-3)) atomicOr(&bad, 1);
    __syncthreads();
    if (threadIdx.x == 0) ok[0] = !bad;
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
    double *G, *R2, *Rc;             // n x n each: R1 (CholeskyQR2), R2, R = R2 R1
    double *Rp, *Wp;                 // the factor's tiles for trsm_tiles_kernel (n <= CH_MAX)
    int *info;                       // potrf status of the two passes, and the check flag
    T *tmpT;                         // d x (n + 1) in T (SA, then Sb: the sketches before promotion)
    // the factorization: explicit Q in QR (CholeskyQR2) or Householder
    // vectors in QR with tau; R upper triangular at R with leading dim ldr
    bool explicit_q = false;
    const double *R = nullptr;
    int64_t ldr = 0;
};

template <typename T>
static SolveWork<T> solve_work(int64_t d, int64_t n, hipStream_t st) {
    auto a = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t dn = (size_t)safe_mul(d, n), nn = (size_t)safe_mul(n, n);
    const size_t rp_bytes = (size_t)CH_T * (CH_T - 1) / 2 * 256 * 8, wp_bytes = (size_t)CH_T * 256 * 8;
    const size_t bytes = 2 * a(dn * 8) + 2 * a((size_t)d * 8) + a((size_t)n * 8) + 3 * a(nn * 8) + a(16) +
                         a((dn + (size_t)d) * sizeof(T)) + a(rp_bytes) + a(wp_bytes);
    char *ws = (char *)scratch(SCRATCH_SOLVE, bytes, st);
    SolveWork<T> w;
    size_t off = 0;
    auto take = [&](size_t b) {
        char *p = ws + off;
        off += a(b);
        return p;
    };
    w.SA = (double *)take(dn * 8);
    w.QR = (double *)take(dn * 8);
    w.Sb = (double *)take((size_t)d * 8);
    w.y = (double *)take((size_t)d * 8);
    w.tau = (double *)take((size_t)n * 8);
    w.G = (double *)take(nn * 8);
    w.R2 = (double *)take(nn * 8);
    w.Rc = (double *)take(nn * 8);
    w.info = (int *)take(16);
    w.tmpT = (T *)take((dn + (size_t)d) * sizeof(T));
    w.Rp = (double *)take(rp_bytes);
    w.Wp = (double *)take(wp_bytes);
    return w;
}

// CholeskyQR2 of the d x n FP64 matrix in w.QR (overwritten by Q), R in w.Rc,
// enqueued without waiting: w.info[2] = 1 when it is accepted (see the file
// comment); otherwise Householder QR must be used instead (w.QR and w.Rc then
// hold no usable factor).
template <typename T>
static void cholqr2(int64_t d, int64_t n, SolveWork<T> &w, hipStream_t st) {
    ProfScope prof("solve_cholqr2", st);
    rocblas_handle h = blas_handle(st);
    const rocblas_int D = (rocblas_int)d, Nn = (rocblas_int)n;
    const double one = 1.0, zero = 0.0;
    double *Rk[2] = {w.G, w.R2};
    for (int pass = 0; pass < 2; ++pass) {
        // the full Gram matrix by GEMM (faster than rocBLAS's SYRK at these
        // shapes); potrf reads its upper triangle
        blas_ok(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, Nn, Nn, D, &one, w.QR, D, w.QR,
                              D, &zero, Rk[pass], Nn),
                "gemm");
        if (n <= CH_MAX) {
            // Q = QR R^{-1} by the factor's tiles, one wave per 16 rows
            hipLaunchKernelGGL(chol_tiles_kernel, dim3(1), dim3(CH_THR), 0, st, Rk[pass], (int)n, w.info + pass,
                               w.Rp, w.Wp);
            RBX_HIP(hipGetLastError());
            launch_trsm_tiles<CH_T>(w.QR, d, n, (const double *)w.Rp, (const double *)w.Wp, st);
        } else {
            blas_ok(rocsolver_dpotrf(h, rocblas_fill_upper, Nn, Rk[pass], Nn, w.info + pass), "potrf");
            blas_ok(rocblas_dtrsm(h, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                                  rocblas_diagonal_non_unit, D, Nn, &one, Rk[pass], Nn, w.QR, D),
                    "trsm");
        }
    }
    hipLaunchKernelGGL(cholqr_check_kernel, dim3(1), dim3(256), 0, st, (const double *)w.R2, n, (const int *)w.info,
                       w.info + 2);
    RBX_HIP(hipGetLastError());
    // R = R2 R1 (R1's strictly lower triangle still holds the Gram matrix)
    hipLaunchKernelGGL(zero_lower_kernel, dim3(grid_of(n * n)), dim3(256), 0, st, w.G, n);
    RBX_HIP(hipGetLastError());
    blas_ok(rocblas_dtrmm(h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none, rocblas_diagonal_non_unit,
                          Nn, Nn, &one, w.R2, Nn, w.G, Nn, w.Rc, Nn),
            "trmm");
}

// SA, Sb (FP64), the QR of SA (CholeskyQR2, else Householder), y = Q^T Sb,
// x64 = R^{-1} y[0:n] (left in y)
template <typename T>
static void factor_and_solve(char layout, int64_t d, int64_t m, int64_t n, const rbx_dense_skop *S, int64_t ro_s,
                             int64_t co_s, const T *A, int64_t lda, const T *b, SolveWork<T> &w,
                             hipStream_t st) {
    RBX_REQUIRE(d <= INT32_MAX && n <= INT32_MAX);
    // SA in T, in `layout` with leading dimension d (C) or n (R), and Sb after
    // it; column-major [A b] in one pass over the operator where the wide
    // kernels take the shape, else two sketches
    const int64_t ldsa = layout == 'C' ? d : n;
    // FP64 column-major sketches land in SA / Sb directly; others go through
    // tmpT and are promoted
    const bool direct = sizeof(T) == 8 && layout == 'C';
    T *SAo = direct ? (T *)w.SA : w.tmpT;
    T *Sbo = direct ? (T *)w.Sb : w.tmpT + d * n;
    if (layout != 'C' || !left_sketch_with_column<T>('N', d, n, m, S, ro_s, co_s, A, lda, b, SAo, d, Sbo, st)) {
        inner_ok(lsk<T>(layout, d, n, m, S, ro_s, co_s, A, lda, SAo, ldsa, st));
        inner_ok(lsk<T>('C', d, 1, m, S, ro_s, co_s, b, m, Sbo, d, st));
    }
    if (!direct) {
        hipLaunchKernelGGL((to_f64_kernel<T>), dim3(grid_of(d * n)), dim3(256), 0, st, (const T *)SAo,
                           layout == 'C' ? (int64_t)1 : ldsa, layout == 'C' ? ldsa : (int64_t)1, d, n, w.SA);
        hipLaunchKernelGGL((to_f64_kernel<T>), dim3(grid_of(d)), dim3(256), 0, st, (const T *)Sbo, (int64_t)1, d, d,
                           (int64_t)1, w.Sb);
        RBX_HIP(hipGetLastError());
    }
    RBX_HIP(hipMemcpyAsync(w.QR, w.SA, (size_t)(d * n) * 8, hipMemcpyDeviceToDevice, st));
    rocblas_handle h = blas_handle(st);
    const char *qr_env = std::getenv("RBX_SOLVE_QR");
    const bool householder_only = qr_env != nullptr && std::strcmp(qr_env, "householder") == 0;
    bool accepted = false;
    if (!householder_only) {
        // the whole CholeskyQR2 path is enqueued, and its acceptance read
        // once at the end (no idle gap in the middle of the solve)
        cholqr2<T>(d, n, w, st);
        w.explicit_q = true;
        w.R = w.Rc;
        w.ldr = n;
        // y[0:n] = Q^T Sb
        const double one = 1.0, zero = 0.0;
        blas_ok(rocblas_dgemv(h, rocblas_operation_transpose, (rocblas_int)d, (rocblas_int)n, &one, w.QR,
                              (rocblas_int)d, w.Sb, 1, &zero, w.y, 1),
                "gemv");
        trsv_upper(h, false, n, w.R, w.ldr, w.y, st);
        int ok = 0;
        RBX_HIP(hipMemcpyAsync(&ok, w.info + 2, sizeof(int), hipMemcpyDeviceToHost, st));
        RBX_HIP(hipStreamSynchronize(st));
        accepted = ok != 0;
    }
    if (!accepted) {
        ProfScope prof("solve_householder", st);
        w.explicit_q = false;
        w.R = w.QR;
        w.ldr = d;
        RBX_HIP(hipMemcpyAsync(w.QR, w.SA, (size_t)(d * n) * 8, hipMemcpyDeviceToDevice, st));
        RBX_HIP(hipMemcpyAsync(w.y, w.Sb, (size_t)d * 8, hipMemcpyDeviceToDevice, st));
        blas_ok(rocsolver_dgeqrf(h, (rocblas_int)d, (rocblas_int)n, w.QR, (rocblas_int)d, w.tau), "geqrf");
        blas_ok(rocsolver_dormqr(h, rocblas_side_left, rocblas_operation_transpose, (rocblas_int)d, 1, (rocblas_int)n,
                                 w.QR, (rocblas_int)d, w.tau, w.y, (rocblas_int)d),
                "ormqr");
        trsv_upper(h, false, n, w.R, w.ldr, w.y, st);
    }
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
    SolveWork<T> w = solve_work<T>(d, n, st);
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
    SolveWork<T> w = solve_work<T>(d, n, st);
    factor_and_solve<T>(layout, d, m, n, S, ro_s, co_s, A, lda, b, w, st);
    // second workspace: x64 (n), u (n), z (n), r (d), wv (d), M (d x n) in FP64, [M w] in T
    auto a = [](size_t bb) { return (bb + 255) / 256 * 256; };
    const size_t dn = (size_t)(d * n);
    char *ws = (char *)scratch(SCRATCH_SOLVE2,
                               3 * a((size_t)n * 8) + 2 * a((size_t)d * 8) + a(dn * 8) + a((dn + (size_t)d) * sizeof(T)),
                               st);
    double *x64 = (double *)ws, *u = (double *)(ws + a((size_t)n * 8)), *z = (double *)(ws + 2 * a((size_t)n * 8));
    double *r = (double *)(ws + 3 * a((size_t)n * 8)), *wv = (double *)(ws + 3 * a((size_t)n * 8) + a((size_t)d * 8));
    double *M = (double *)(ws + 3 * a((size_t)n * 8) + 2 * a((size_t)d * 8));
    T *MT = (T *)(ws + 3 * a((size_t)n * 8) + 2 * a((size_t)d * 8) + a(dn * 8));
    rocblas_handle h = blas_handle(st);
    RBX_HIP(hipMemcpyAsync(x64, w.y, (size_t)n * 8, hipMemcpyDeviceToDevice, st));
    // u = R^{-T} xbar
    hipLaunchKernelGGL((to_f64_kernel<T>), dim3(grid_of(n)), dim3(256), 0, st, xbar, (int64_t)1, n, n, (int64_t)1, u);
    RBX_HIP(hipGetLastError());
    trsv_upper(h, true, n, w.R, w.ldr, u, st);
    const double one = 1.0, mone = -1.0, zero = 0.0;
    // w = Q [u; 0]
    if (w.explicit_q) {
        blas_ok(rocblas_dgemv(h, rocblas_operation_none, (rocblas_int)d, (rocblas_int)n, &one, w.QR, (rocblas_int)d, u,
                              1, &zero, wv, 1),
                "gemv");
    } else {
        RBX_HIP(hipMemcpyAsync(wv, u, (size_t)n * 8, hipMemcpyDeviceToDevice, st));
        hipLaunchKernelGGL(zero_tail_kernel, dim3(grid_of(d - n)), dim3(256), 0, st, wv, n, d);
        RBX_HIP(hipGetLastError());
        blas_ok(rocsolver_dormqr(h, rocblas_side_left, rocblas_operation_none, (rocblas_int)d, 1, (rocblas_int)n, w.QR,
                                 (rocblas_int)d, w.tau, wv, (rocblas_int)d),
                "ormqr");
    }
    // z = R^{-1} u
    RBX_HIP(hipMemcpyAsync(z, u, (size_t)n * 8, hipMemcpyDeviceToDevice, st));
    trsv_upper(h, false, n, w.R, w.ldr, z, st);
    // r = Sb - SA x
    RBX_HIP(hipMemcpyAsync(r, w.Sb, (size_t)d * 8, hipMemcpyDeviceToDevice, st));
    blas_ok(rocblas_dgemv(h, rocblas_operation_none, (rocblas_int)d, (rocblas_int)n, &mone, w.SA, (rocblas_int)d, x64, 1,
                          &one, r, 1),
            "gemv");
    // M = r z^T - w x^T, in `layout` and T for the adjoint sketch
    hipLaunchKernelGGL(rank2_kernel, dim3(grid_of(d * n)), dim3(256), 0, st, (const double *)r, (const double *)z,
                       (const double *)wv, (const double *)x64, d, n, M);
    const int64_t ldm = layout == 'C' ? d : n;
    hipLaunchKernelGGL((from_f64_kernel<T>), dim3(grid_of(d * n)), dim3(256), 0, st, (const double *)M, d, n, MT,
                       layout == 'C' ? (int64_t)1 : ldm, layout == 'C' ? ldm : (int64_t)1);
    T *wT = MT + dn;
    hipLaunchKernelGGL((from_f64_kernel<T>), dim3(grid_of(d)), dim3(256), 0, st, (const double *)wv, d, (int64_t)1, wT,
                       (int64_t)1, d);
    RBX_HIP(hipGetLastError());
    // Abar = S^T M and bbar = S^T w: column-major [M w] in one pass over the
    // operator where the wide kernels take the shape, else two sketches
    if (layout != 'C' || !left_sketch_with_column<T>('T', m, n, d, S, ro_s, co_s, MT, ldm, wT, Abar, ldab, bbar, st)) {
        inner_ok(lsk_adj<T>(layout, d, n, m, S, ro_s, co_s, MT, ldm, Abar, ldab, st));
        inner_ok(lsk_adj<T>('C', d, 1, m, S, ro_s, co_s, wT, d, bbar, m, st));
    }
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
