// rbx_spmm_lds.hip -- sparse operator times dense matrix with the dense
// operand read once, for the left_spmm cases whose dense operand has unit
// stride along the contraction (the SASO / LASO lskges of a column-major A,
// coo_spmm_impl.hh:52-105 -> csc_spmm_impl.hh:74-152).
//
// 1. csr_build_counting: the operator's entries inside the window become a
//    CSR sorted by (row, column, stored position) with a stable counting sort
//    by row (block histograms in LDS, one scan, a block-local stable radix
//    sort that ranks each entry among the earlier entries of its row), then a
//    per-row check: rows whose columns are not ascending -- entries not stored
//    in column order -- are sorted by (column, position) in LDS, or by LDS
//    tiles and merge-path passes when longer than a tile.  For an operator
//    stored column by column (the sampled SASO / LASO arrays of a wide
//    operator) no row needs that second sort.
// 2. csr_lds_apply: a 1024-thread workgroup owns J output columns and 1024
//    output rows (one per thread).  It streams the J columns of the dense
//    operand through LDS in 1024-deep chunks (double-buffered, coalesced
//    column reads); each thread walks its row's entries in order and applies
//    those whose column falls in the current chunk.  Every output entry is
//    one sequential chain in the reference's order and summation form, so
//    the result is the reference's bit for bit, while the dense operand is
//    read from HBM once (the row groups of a column block are dispatched to
//    one XCD back to back, so they share it through that XCD's L2).
#include <rocprim/block/block_radix_sort.hpp>

#include "rbx_spmm.hh"

namespace rbx {

static size_t al256(size_t x) { return (x + 255) / 256 * 256; }

static int bits_of(uint64_t x) {
    int b = 0;
    while (b < 64 && (x >> b) != 0) ++b;
    return b < 1 ? 1 : b;
}

// ---------------------------------------------------------------- CSR by counting sort
constexpr int CB_THR = 256, CB_IPT = 16, CB_TILE = CB_THR * CB_IPT;  // 4096 entries per tile
constexpr int CB_MAX_BLOCKS = 256;

// H[b * d + r] = number of window entries of row r among block b's entries
// [b * span, (b + 1) * span).
__global__ __launch_bounds__(CB_THR) void csrb_hist_kernel(const int64_t *rows, const int64_t *cols, int64_t nnz,
                                                           int64_t ro, int64_t co, int d, int64_t m, int64_t span,
                                                           int32_t *H) {
    __shared__ int32_t h[CSRB_MAX_ROWS];
    for (int r = threadIdx.x; r < d; r += CB_THR) h[r] = 0;
    __syncthreads();
    const int64_t e0 = (int64_t)blockIdx.x * span, e1 = e0 + span < nnz ? e0 + span : nnz;
    for (int64_t e = e0 + threadIdx.x; e < e1; e += CB_THR) {
        const int64_t r = rows[e] - ro, c = cols[e] - co;
        if (r >= 0 && r < d && c >= 0 && c < m) atomicAdd(&h[r], 1);
    }
    __syncthreads();
    for (int r = threadIdx.x; r < d; r += CB_THR) H[(int64_t)blockIdx.x * d + r] = h[r];
}

// One workgroup: H[b * d + r] becomes the number of row-r entries in blocks
// before b, and row_ptr the exclusive scan of the row totals.
__global__ __launch_bounds__(1024) void csrb_scan_kernel(int32_t *H, int nb, int d, int64_t *row_ptr) {
    __shared__ int32_t tot[CSRB_MAX_ROWS];
    __shared__ int32_t part[1024];
    const int t = threadIdx.x;
    for (int r = t; r < d; r += 1024) {
        int32_t run = 0;
#pragma unroll 8
        for (int b = 0; b < nb; ++b) {
            const int64_t i = (int64_t)b * d + r;
            const int32_t x = H[i];
            H[i] = run;
            run += x;
        }
        tot[r] = run;
    }
    __syncthreads();
    const int per = (d + 1023) / 1024;  // rows per thread in the scan (<= 8)
    int32_t s = 0;
    for (int q = 0; q < per; ++q) {
        const int r = t * per + q;
        if (r < d) s += tot[r];
    }
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int32_t v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int64_t run = t > 0 ? part[t - 1] : 0;
    for (int q = 0; q < per; ++q) {
        const int r = t * per + q;
        if (r < d) {
            row_ptr[r] = run;
            run += tot[r];
        }
    }
    if (t == 1023) row_ptr[d] = part[1023];
}

// Each block places its entries, tile by tile in stored order: a stable block
// radix sort on the row orders a tile's entries by (row, position); an
// entry's slot is row_ptr[r] + (row-r entries of earlier blocks) + (row-r
// entries of earlier tiles of this block) + (its rank in the tile's run).
// The slot receives (column << 32) | position.
__global__ __launch_bounds__(CB_THR) void csrb_scatter_kernel(const int64_t *rows, const int64_t *cols, int64_t nnz,
                                                              int64_t ro, int64_t co, int d, int64_t m, int64_t span,
                                                              const int32_t *H, const int64_t *row_ptr, int end_bit,
                                                              uint64_t *keys) {
    using Sort = rocprim::block_radix_sort<uint32_t, CB_THR, CB_IPT, uint32_t>;
    __shared__ typename Sort::storage_type sst;
    __shared__ int32_t run[CSRB_MAX_ROWS];
    __shared__ int32_t first[CSRB_MAX_ROWS];
    __shared__ uint32_t skey[CB_TILE];
    const int t = threadIdx.x;
    for (int r = t; r < d; r += CB_THR) run[r] = 0;
    const int64_t e0 = (int64_t)blockIdx.x * span, e1 = e0 + span < nnz ? e0 + span : nnz;
    __syncthreads();
    for (int64_t base = e0; base < e1; base += CB_TILE) {
        uint32_t key[CB_IPT], pos[CB_IPT];
#pragma unroll
        for (int k = 0; k < CB_IPT; ++k) {
            const int li = t * CB_IPT + k;
            const int64_t e = base + li;
            key[k] = (uint32_t)d;  // outside the window or past the end: sorts last, ignored
            pos[k] = (uint32_t)li;
            if (e < e1) {
                const int64_t r = rows[e] - ro, c = cols[e] - co;
                if (r >= 0 && r < d && c >= 0 && c < m) key[k] = (uint32_t)r;
            }
        }
        Sort().sort(key, pos, sst, 0, (unsigned)end_bit);
#pragma unroll
        for (int k = 0; k < CB_IPT; ++k) skey[t * CB_IPT + k] = key[k];
        __syncthreads();
#pragma unroll
        for (int k = 0; k < CB_IPT; ++k) {
            const int i = t * CB_IPT + k;
            const uint32_t r = key[k];
            if (r < (uint32_t)d && (i == 0 || skey[i - 1] != r)) first[r] = i;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < CB_IPT; ++k) {
            const int i = t * CB_IPT + k;
            const uint32_t r = key[k];
            if (r < (uint32_t)d) {
                const int64_t e = base + pos[k];
                const int64_t dest = row_ptr[r] + H[(int64_t)blockIdx.x * d + r] + run[r] + (i - first[r]);
                keys[dest] = ((uint64_t)(cols[e] - co) << 32) | (uint64_t)e;
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < CB_IPT; ++k) {
            const int i = t * CB_IPT + k;
            const uint32_t r = key[k];
            if (r < (uint32_t)d && (i == CB_TILE - 1 || skey[i + 1] != r)) run[r] += i - first[r] + 1;
        }
        __syncthreads();
    }
}

// Ascending bitonic sort of s[0, N), N a power of two <= CB_TILE.
__device__ void csrb_bitonic(uint64_t *s, int N) {
    for (int k = 2; k <= N; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < N; i += CB_THR) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const uint64_t x = s[i], y = s[ixj];
                    const bool up = (i & k) == 0;
                    if ((x > y) == up) {
                        s[i] = y;
                        s[ixj] = x;
                    }
                }
            }
            __syncthreads();
        }
}

// Number of elements of sorted a (na) among the first k of the stable merge
// of a and b (nb): elements of a precede equal elements of b.
__device__ int64_t csrb_corank(int64_t k, const uint64_t *a, int64_t na, const uint64_t *b, int64_t nb) {
    int64_t lo = k - nb > 0 ? k - nb : 0, hi = k < na ? k : na;
    while (lo < hi) {
        const int64_t i = (lo + hi) >> 1;
        if (a[i] <= b[k - i - 1]) lo = i + 1;
        else hi = i;
    }
    return lo;
}

// Sorts a[0, len) (global); b is a buffer of the same length.  Returns the
// buffer holding the result.  Called by the whole workgroup with uniform
// arguments.
__device__ const uint64_t *csrb_sort_row(uint64_t *a, uint64_t *b, int64_t len, uint64_t *s) {
    for (int64_t t0 = 0; t0 < len; t0 += CB_TILE) {
        const int n = (int)(len - t0 < CB_TILE ? len - t0 : CB_TILE);
        int N = 1;
        while (N < n) N <<= 1;
        for (int i = threadIdx.x; i < N; i += CB_THR) s[i] = i < n ? a[t0 + i] : ~0ull;
        __syncthreads();
        csrb_bitonic(s, N);
        for (int i = threadIdx.x; i < n; i += CB_THR) a[t0 + i] = s[i];
        __syncthreads();
    }
    uint64_t *src = a, *dst = b;
    for (int64_t w = CB_TILE; w < len; w *= 2) {
        for (int64_t s0 = 0; s0 < len; s0 += 2 * w) {
            const int64_t amid = s0 + w < len ? s0 + w : len, aend = s0 + 2 * w < len ? s0 + 2 * w : len;
            const int64_t na = amid - s0, nb = aend - amid, tot = na + nb;
            const int64_t k0 = tot * threadIdx.x / CB_THR, k1 = tot * (threadIdx.x + 1) / CB_THR;
            int64_t i = csrb_corank(k0, src + s0, na, src + amid, nb), j = k0 - i;
            for (int64_t k = k0; k < k1; ++k) {
                uint64_t x;
                if (j >= nb || (i < na && src[s0 + i] <= src[amid + j])) x = src[s0 + i++];
                else x = src[amid + j++];
                dst[s0 + k] = x;
            }
        }
        __syncthreads();
        uint64_t *tmp = src;
        src = dst;
        dst = tmp;
    }
    return src;
}

// One workgroup per row (grid-stride): sort the row by (column, position) if
// its columns are not ascending, then write the 32-bit columns and gather the
// values.
template <typename T>
__global__ __launch_bounds__(CB_THR) void csrb_finish_kernel(const int64_t *row_ptr, int d, uint64_t *keys,
                                                             uint64_t *keys2, const T *vals, int32_t *col, T *val) {
    __shared__ uint64_t s[CB_TILE];
    __shared__ int unsorted;
    for (int r = blockIdx.x; r < d; r += gridDim.x) {
        const int64_t p0 = row_ptr[r], len = row_ptr[r + 1] - p0;
        if (threadIdx.x == 0) unsorted = 0;
        __syncthreads();
        int bad = 0;
        for (int64_t q = threadIdx.x; q + 1 < len; q += CB_THR) bad |= keys[p0 + q] > keys[p0 + q + 1];
        if (bad) unsorted = 1;
        __syncthreads();
        const uint64_t *src = keys + p0;
        if (unsorted) src = csrb_sort_row(keys + p0, keys2 + p0, len, s);
        for (int64_t q = threadIdx.x; q < len; q += CB_THR) {
            const uint64_t k = src[q];
            col[p0 + q] = (int32_t)(k >> 32);
            val[p0 + q] = vals[(uint32_t)k];
        }
        __syncthreads();
    }
}

template <typename T>
bool csr_build_counting(const int64_t *rows, const int64_t *cols, const T *vals, int64_t nnz, int64_t ro, int64_t co,
                        int64_t d, int64_t m, int64_t **row_ptr_out, int32_t **col_out, T **val_out, hipStream_t st) {
    if (d <= 0 || d > CSRB_MAX_ROWS || nnz <= 0 || nnz >= (1ll << 31) || m >= (1ll << 31)) return false;
    const int64_t tiles = (nnz + CB_TILE - 1) / CB_TILE;
    const int64_t tpb = (tiles + CB_MAX_BLOCKS - 1) / CB_MAX_BLOCKS;  // tiles per block
    const int nb = (int)((tiles + tpb - 1) / tpb);
    const int64_t span = tpb * CB_TILE;
    const size_t bH = al256((size_t)nb * d * 4), bR = al256((size_t)(d + 1) * 8), bK = al256((size_t)nnz * 8),
                 bC = al256((size_t)nnz * 4), bV = al256((size_t)nnz * sizeof(T));
    char *ws = (char *)scratch(SCRATCH_SPARSE_CSR, bH + bR + 2 * bK + bC + bV, st);
    int32_t *H = (int32_t *)ws;
    int64_t *row_ptr = (int64_t *)(ws + bH);
    uint64_t *keys = (uint64_t *)(ws + bH + bR), *keys2 = (uint64_t *)(ws + bH + bR + bK);
    int32_t *col = (int32_t *)(ws + bH + bR + 2 * bK);
    T *val = (T *)(ws + bH + bR + 2 * bK + bC);
    const int di = (int)d;
    hipLaunchKernelGGL(csrb_hist_kernel, dim3(nb), dim3(CB_THR), 0, st, rows, cols, nnz, ro, co, di, m, span, H);
    RBX_HIP(hipGetLastError());
    hipLaunchKernelGGL(csrb_scan_kernel, dim3(1), dim3(1024), 0, st, H, nb, di, row_ptr);
    RBX_HIP(hipGetLastError());
    hipLaunchKernelGGL(csrb_scatter_kernel, dim3(nb), dim3(CB_THR), 0, st, rows, cols, nnz, ro, co, di, m, span,
                       (const int32_t *)H, (const int64_t *)row_ptr, bits_of((uint64_t)d), keys);
    RBX_HIP(hipGetLastError());
    const int fg = (int)(d < 4 * (int64_t)num_cus() ? d : 4 * (int64_t)num_cus());
    hipLaunchKernelGGL((csrb_finish_kernel<T>), dim3(fg), dim3(CB_THR), 0, st, (const int64_t *)row_ptr, di, keys,
                       keys2, vals, col, val);
    RBX_HIP(hipGetLastError());
    *row_ptr_out = row_ptr;
    *col_out = col;
    *val_out = val;
    return true;
}

// ---------------------------------------------------------------- apply through LDS
constexpr int LA_THR = 1024, LA_CK = 1024, LA_WIN = 8;

template <typename T, int J, int FORM>
__global__ __launch_bounds__(LA_THR) void csr_lds_apply_kernel(const int64_t *row_ptr, const int32_t *col,
                                                               const T *val, int64_t d, int64_t n, int64_t m, T alpha,
                                                               T beta, const T *B, int64_t ldb, T *C, int64_t crs,
                                                               int64_t ccs, int RW, int ncb) {
    __shared__ __attribute__((aligned(16))) T As[2][LA_CK * J];
    const int tid = threadIdx.x;
    // XCD-aware order: the RW row groups of a column block are consecutive
    // tiles on one XCD, so its columns of B stream from HBM once
    const int W = ncb * RW, L = blockIdx.x;
    const int tt = (W % 8 == 0) ? (L % 8) * (W / 8) + L / 8 : L;
    const int cb = tt / RW, rg = tt % RW;
    const int64_t j0 = (int64_t)cb * J;
    const int64_t r = (int64_t)rg * LA_THR + tid;
    const bool active = r < d;

    T acc[J];
#pragma unroll
    for (int jj = 0; jj < J; ++jj) {
        acc[jj] = (T)0;
        if (active && beta != (T)0 && j0 + jj < n) acc[jj] = beta * C[r * crs + (j0 + jj) * ccs];
    }
    int64_t p = active ? row_ptr[r] : 0;
    const int64_t pend = active ? row_ptr[r + 1] : 0;

    T ld[J];
    auto load_chunk = [&](int64_t k) {
        const int64_t c = k * LA_CK + tid;
#pragma unroll
        for (int jj = 0; jj < J; ++jj) ld[jj] = (c < m && j0 + jj < n) ? B[c + (j0 + jj) * ldb] : (T)0;
    };
    auto store_chunk = [&](int buf) {
#pragma unroll
        for (int jj = 0; jj < J; ++jj) As[buf][tid * J + jj] = ld[jj];
    };
    const int64_t nk = (m + LA_CK - 1) / LA_CK;
    load_chunk(0);
    store_chunk(0);
    __syncthreads();
    for (int64_t k = 0; k < nk; ++k) {
        if (k + 1 < nk) load_chunk(k + 1);
        const int64_t c0 = k * LA_CK;
        const int32_t c1 = (int32_t)((k + 1) * LA_CK < m ? (k + 1) * LA_CK : m);
        const T *Ab = As[k & 1];
        // windows of LA_WIN entries from p; stop at the chunk's end
        while (p < pend) {
            int32_t wc[LA_WIN];
            T wv[LA_WIN];
#pragma unroll
            for (int q = 0; q < LA_WIN; ++q) {
                const bool in = p + q < pend;
                wc[q] = in ? col[p + q] : INT32_MAX;
                wv[q] = in ? val[p + q] : (T)0;
            }
            int used = 0;
#pragma unroll
            for (int q = 0; q < LA_WIN; ++q) {
                if (used == q && wc[q] < c1) {
                    const T v = wv[q];
                    const T *b = Ab + (wc[q] - c0) * J;
#pragma unroll
                    for (int jj = 0; jj < J; ++jj) {
                        if (FORM == SPF_KIB) acc[jj] = acc[jj] + (alpha * v) * b[jj];
                        else acc[jj] = acc[jj] + v * (alpha * b[jj]);
                    }
                    used = q + 1;
                }
            }
            p += used;
            if (used < LA_WIN) break;
        }
        if (k + 1 < nk) store_chunk((int)((k + 1) & 1));
        __syncthreads();
    }
    if (active) {
#pragma unroll
        for (int jj = 0; jj < J; ++jj)
            if (j0 + jj < n) C[r * crs + (j0 + jj) * ccs] = acc[jj];
    }
}

template <typename T, int J>
static void launch_lds_apply(int form, dim3 grid, const int64_t *row_ptr, const int32_t *col, const T *val, int64_t d,
                             int64_t n, int64_t m, T alpha, T beta, const T *B, int64_t ldb, T *C, int64_t crs,
                             int64_t ccs, int RW, int ncb, hipStream_t st) {
    if (form == SPF_KIB)
        hipLaunchKernelGGL((csr_lds_apply_kernel<T, J, SPF_KIB>), grid, dim3(LA_THR), 0, st, row_ptr, col, val, d, n,
                           m, alpha, beta, B, ldb, C, crs, ccs, RW, ncb);
    else
        hipLaunchKernelGGL((csr_lds_apply_kernel<T, J, SPF_JKI>), grid, dim3(LA_THR), 0, st, row_ptr, col, val, d, n,
                           m, alpha, beta, B, ldb, C, crs, ccs, RW, ncb);
    RBX_HIP(hipGetLastError());
}

template <typename T>
bool csr_lds_apply(const int64_t *row_ptr, const int32_t *col, const T *val, int64_t d, int64_t n, int64_t m,
                   T alpha, T beta, const T *B, int64_t ldb, T *C, int64_t crs, int64_t ccs, int form,
                   hipStream_t st) {
    if (d <= 0 || n <= 0 || m <= 0 || m >= (1ll << 31)) return false;
    const int64_t RW = (d + LA_THR - 1) / LA_THR;
    // widest column block (16 bytes of B per contraction index, i.e. 2 FP64
    // or 4 FP32 columns, up to twice that) that still gives one workgroup per CU
    constexpr int JW = 16 / (int)sizeof(T);
    const int64_t cus = num_cus();
    int J = 2 * JW;
    if (((n + J - 1) / J) * RW < cus) J = JW;
    if (((n + J - 1) / J) * RW < cus) J = 1;
    const int64_t ncb = (n + J - 1) / J;
    if (ncb * RW >= (1ll << 31)) return false;
    const dim3 grid((unsigned)(ncb * RW));
    const int rw = (int)RW, nc = (int)ncb;
    if (J == 2 * JW) launch_lds_apply<T, 2 * JW>(form, grid, row_ptr, col, val, d, n, m, alpha, beta, B, ldb, C, crs, ccs, rw, nc, st);
    else if (J == JW) launch_lds_apply<T, JW>(form, grid, row_ptr, col, val, d, n, m, alpha, beta, B, ldb, C, crs, ccs, rw, nc, st);
    else launch_lds_apply<T, 1>(form, grid, row_ptr, col, val, d, n, m, alpha, beta, B, ldb, C, crs, ccs, rw, nc, st);
    return true;
}

template bool csr_build_counting<float>(const int64_t *, const int64_t *, const float *, int64_t, int64_t, int64_t,
                                        int64_t, int64_t, int64_t **, int32_t **, float **, hipStream_t);
template bool csr_build_counting<double>(const int64_t *, const int64_t *, const double *, int64_t, int64_t, int64_t,
                                         int64_t, int64_t, int64_t **, int32_t **, double **, hipStream_t);
template bool csr_lds_apply<float>(const int64_t *, const int32_t *, const float *, int64_t, int64_t, int64_t, float,
                                   float, const float *, int64_t, float *, int64_t, int64_t, int, hipStream_t);
template bool csr_lds_apply<double>(const int64_t *, const int32_t *, const double *, int64_t, int64_t, int64_t,
                                    double, double, const double *, int64_t, double *, int64_t, int64_t, int,
                                    hipStream_t);

}  // namespace rbx
