// rbx_sparse.hip -- sparse sketching operators (SASO / LASO) and the sparse
// products of RandBLAS's hot path:
//   fill_sparse_unpacked_nosub  (sparse_skops.hh:515-565)
//   lskges / rskges             (skge.hh:465-626, spmm_dispatch.hh:52-219)
//   rsksp3                      (sparse_data/sksp.hh:277-326, csr_spmm_impl.hh:51-155)
//
// Sampling.  SASO: vector i (of dim_minor) is a partial Fisher-Yates shuffle
// of the identity on dim_major points using counters seed + i*vec_nnz + j
// (sparse_skops.hh:71-103; the reference restores its work array after every
// vector, so vectors are independent) -- one thread per vector keeps the few
// displaced positions in registers.  LASO: vec_nnz iid samples per vector from
// two Philox words each (util.hh:516-546), duplicates merged to
// sqrt(count)*first_sign in first-occurrence order, then compacted.
//
// Products.  Each output entry of op(S)*op(B) is accumulated by one thread in
// ascending operator column, which is the order of the reference's fallback
// kernels (csc_spmm_impl.hh:74-96, 154-209), so +-1 operators give
// bit-identical results.  The COO operator is transposed once into row-major
// CSR (rocPRIM radix sort on the (row, column) key); the apply kernel stages a
// chunk of op(B) in LDS and lets every thread walk its rows' sorted entries.
#include <type_traits>

#include <mutex>

#include "rbx_internal.hh"
#include "rbx_spmm.hh"

#include <math.h>
#include <string.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

namespace rbx {

// ---------------------------------------------------------------- geometry
struct SparseGeom {
    int64_t dim_major, dim_minor, full_nnz;
};

static SparseGeom sparse_geom(const rbx_sparse_dist &D) {
    if (D.n_rows <= 0 || D.n_cols <= 0) RBX_FAIL(RBX_INVALID_ARG, "SparseDist: n_rows and n_cols must be positive");
    if (D.vec_nnz <= 0) RBX_FAIL(RBX_INVALID_ARG, "SparseDist: vec_nnz must be positive");
    if (D.major_axis != 'S' && D.major_axis != 'L') RBX_FAIL(RBX_INVALID_ARG, "SparseDist: major_axis must be 'S' or 'L'");
    SparseGeom g;
    int64_t mx = D.n_rows > D.n_cols ? D.n_rows : D.n_cols;
    int64_t mn = D.n_rows < D.n_cols ? D.n_rows : D.n_cols;
    g.dim_major = D.major_axis == 'S' ? mn : mx;
    g.dim_minor = D.n_rows + D.n_cols - g.dim_major;
    if (D.vec_nnz > g.dim_major) RBX_FAIL(RBX_INVALID_ARG, "SparseDist: requirement failed: vec_nnz <= dim_major");
    g.full_nnz = safe_mul(D.vec_nnz, g.dim_minor);
    return g;
}

// compute_next_state(SparseDist, ...) (sparse_skops.hh:266-283)
static int64_t sparse_incr(const rbx_sparse_dist &D) {
    if (D.major_axis == 'S') {
        int64_t mx = D.n_rows > D.n_cols ? D.n_rows : D.n_cols;
        return safe_mul(mx, D.vec_nnz);
    }
    int64_t mn = D.n_rows < D.n_cols ? D.n_rows : D.n_cols;
    return safe_mul(mn, (D.vec_nnz + 1) / 2);
}

// ---------------------------------------------------------------- SASO
template <typename T, int MAXK>
__global__ __launch_bounds__(256) void saso_kernel(Ctr128 ctr, uint32_t k0, uint32_t k1, int64_t vec_nnz,
                                                   int64_t dim_major, int64_t dim_minor, int64_t *idx_major,
                                                   int64_t *idx_minor, T *vals) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= dim_minor) return;
    // displaced positions (> current step) and the values they hold
    int64_t pos[MAXK], val[MAXK];
    int nm = 0;
    const int64_t base = i * vec_nnz;
    for (int64_t j = 0; j < vec_nnz; ++j) {
        uint4 w = philox4x32_10(ctr_add(ctr, (uint64_t)(base + j)), k0, k1);
        const int64_t ell = j + (int64_t)((int64_t)w.x % (dim_major - j));
        int64_t a = j, b = ell;
        int ia = -1, ib = -1;
        for (int t = 0; t < nm; ++t) {
            if (pos[t] == j) { a = val[t]; ia = t; }
            if (pos[t] == ell) { b = val[t]; ib = t; }
        }
        int64_t out = a;
        if (ell != j) {
            out = b;
            if (ib >= 0) val[ib] = a;
            else if (ia >= 0) { pos[ia] = ell; val[ia] = a; ia = -1; }  // reuse j's slot: j is final now
            else { pos[nm] = ell; val[nm] = a; ++nm; }
        }
        if (ia >= 0) {  // drop position j (never read again)
            pos[ia] = pos[nm - 1];
            val[ia] = val[nm - 1];
            --nm;
        }
        idx_major[base + j] = out;
        if (idx_minor) idx_minor[base + j] = i;
        vals[base + j] = (w.y % 2u == 0u) ? (T)1 : (T)-1;
    }
}

// ---------------------------------------------------------------- LASO
template <typename T, int MAXK>
__global__ __launch_bounds__(256) void laso_kernel(Ctr128 ctr, uint32_t k0, uint32_t k1, int64_t vec_nnz,
                                                   int64_t dim_major, int64_t dim_minor, int64_t *tmp_idx, T *tmp_val,
                                                   int64_t *counts) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= dim_minor) return;
    const int64_t per = (vec_nnz + 1) / 2;
    const double dN = (double)dim_major;
    int64_t uidx[MAXK];
    int cnt[MAXK];
    float sgn[MAXK];
    int nu = 0;
    uint4 w = make_uint4(0, 0, 0, 0);
    for (int64_t s = 0; s < vec_nnz; ++s) {
        if ((s & 1) == 0) w = philox4x32_10(ctr_add(ctr, (uint64_t)(i * per + s / 2)), k0, k1);
        const uint32_t wi = (s & 1) ? w.z : w.x;
        const uint32_t ws = (s & 1) ? w.w : w.y;
        const double u = __dmul_rn(__dadd_rn((double)uneg11(wi), 1.0), 0.5);
        const int64_t idx = (int64_t)__dmul_rn(dN, u);
        const float sg = (uneg11(ws) >= 0.0f) ? 1.0f : -1.0f;
        int f = -1;
        for (int t = 0; t < nu; ++t)
            if (uidx[t] == idx) { f = t; break; }
        if (f >= 0) cnt[f] += 1;
        else { uidx[nu] = idx; cnt[nu] = 1; sgn[nu] = sg; ++nu; }
    }
    for (int t = 0; t < nu; ++t) {
        tmp_idx[i * vec_nnz + t] = uidx[t];
        T c = (T)cnt[t];
        tmp_val[i * vec_nnz + t] = (cnt[t] == 1) ? (T)sgn[t] : (T)sqrt(c) * (T)sgn[t];
    }
    counts[i] = nu;
}

template <typename T>
__global__ __launch_bounds__(256) void laso_compact_kernel(const int64_t *tmp_idx, const T *tmp_val,
                                                           const int64_t *counts, const int64_t *offsets,
                                                           int64_t vec_nnz, int64_t dim_minor, int64_t *idx_long,
                                                           int64_t *idx_short, T *vals) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= dim_minor) return;
    const int64_t o = offsets[i];
    for (int64_t t = 0; t < counts[i]; ++t) {
        idx_long[o + t] = tmp_idx[i * vec_nnz + t];
        idx_short[o + t] = i;
        vals[o + t] = tmp_val[i * vec_nnz + t];
    }
}

template <typename T>
static void launch_saso(Ctr128 c, uint32_t k0, uint32_t k1, int64_t k, int64_t dmaj, int64_t dmin, int64_t *imaj,
                        int64_t *imin, T *vals, hipStream_t st) {
    dim3 grid((unsigned)((dmin + 255) / 256));
    ProfScope prof("saso_kernel", st);
    if (k <= 8) hipLaunchKernelGGL((saso_kernel<T, 8>), grid, dim3(256), 0, st, c, k0, k1, k, dmaj, dmin, imaj, imin, vals);
    else if (k <= 32) hipLaunchKernelGGL((saso_kernel<T, 32>), grid, dim3(256), 0, st, c, k0, k1, k, dmaj, dmin, imaj, imin, vals);
    else if (k <= 128) hipLaunchKernelGGL((saso_kernel<T, 128>), grid, dim3(256), 0, st, c, k0, k1, k, dmaj, dmin, imaj, imin, vals);
    else RBX_FAIL(RBX_UNSUPPORTED, "fill_sparse: SASO vec_nnz > 128 is not supported on the device");
    RBX_HIP(hipGetLastError());
}

template <typename T>
static void launch_laso(Ctr128 c, uint32_t k0, uint32_t k1, int64_t k, int64_t dmaj, int64_t dmin, int64_t *tidx,
                        T *tval, int64_t *counts, hipStream_t st) {
    dim3 grid((unsigned)((dmin + 255) / 256));
    ProfScope prof("laso_kernel", st);
    if (k <= 8) hipLaunchKernelGGL((laso_kernel<T, 8>), grid, dim3(256), 0, st, c, k0, k1, k, dmaj, dmin, tidx, tval, counts);
    else if (k <= 32) hipLaunchKernelGGL((laso_kernel<T, 32>), grid, dim3(256), 0, st, c, k0, k1, k, dmaj, dmin, tidx, tval, counts);
    else if (k <= 128) hipLaunchKernelGGL((laso_kernel<T, 128>), grid, dim3(256), 0, st, c, k0, k1, k, dmaj, dmin, tidx, tval, counts);
    else RBX_FAIL(RBX_UNSUPPORTED, "fill_sparse: LASO vec_nnz > 128 is not supported on the device");
    RBX_HIP(hipGetLastError());
}

static size_t align256(size_t x) { return (x + 255) / 256 * 256; }

// fill_sparse_unpacked_nosub; returns nnz (synchronises the stream for LASO).
template <typename T>
static int64_t fill_sparse_impl(const rbx_sparse_dist &D, T *vals, int64_t *rows, int64_t *cols,
                                const rbx_rng_state &seed, rbx_rng_state *next_state, hipStream_t st) {
    SparseGeom g = sparse_geom(D);
    RBX_REQUIRE(vals != nullptr && rows != nullptr && cols != nullptr);
    const Ctr128 c = ctr_from_words(seed.counter);
    int64_t *idxs_short = (D.n_rows <= D.n_cols) ? rows : cols;
    int64_t *idxs_long = (D.n_rows <= D.n_cols) ? cols : rows;
    int64_t nnz;
    if (D.major_axis == 'S') {
        launch_saso<T>(c, seed.key[0], seed.key[1], D.vec_nnz, g.dim_major, g.dim_minor, idxs_short, idxs_long, vals, st);
        nnz = g.full_nnz;
    } else {
        // temporaries: idx, val, counts, offsets, scan storage
        size_t b_idx = align256((size_t)g.full_nnz * 8), b_val = align256((size_t)g.full_nnz * sizeof(T));
        size_t b_cnt = align256((size_t)g.dim_minor * 8);
        size_t scan_bytes = 0;
        RBX_HIP(rocprim::exclusive_scan(nullptr, scan_bytes, (int64_t *)nullptr, (int64_t *)nullptr, (int64_t)0,
                                        (size_t)g.dim_minor, rocprim::plus<int64_t>(), st));
        char *ws = (char *)scratch(SCRATCH_SPARSE_OP, b_idx + b_val + 2 * b_cnt + align256(scan_bytes) + 256, st);
        int64_t *tidx = (int64_t *)ws;
        T *tval = (T *)(ws + b_idx);
        int64_t *counts = (int64_t *)(ws + b_idx + b_val);
        int64_t *offs = (int64_t *)(ws + b_idx + b_val + b_cnt);
        void *tmp = ws + b_idx + b_val + 2 * b_cnt;
        launch_laso<T>(c, seed.key[0], seed.key[1], D.vec_nnz, g.dim_major, g.dim_minor, tidx, tval, counts, st);
        RBX_HIP(rocprim::exclusive_scan(tmp, scan_bytes, counts, offs, (int64_t)0, (size_t)g.dim_minor,
                                        rocprim::plus<int64_t>(), st));
        dim3 grid((unsigned)((g.dim_minor + 255) / 256));
        hipLaunchKernelGGL(laso_compact_kernel<T>, grid, dim3(256), 0, st, tidx, tval, counts, offs, D.vec_nnz,
                           g.dim_minor, idxs_long, idxs_short, vals);
        RBX_HIP(hipGetLastError());
        int64_t last_off = 0, last_cnt = 0;
        RBX_HIP(hipMemcpyAsync(&last_off, offs + g.dim_minor - 1, 8, hipMemcpyDeviceToHost, st));
        RBX_HIP(hipMemcpyAsync(&last_cnt, counts + g.dim_minor - 1, 8, hipMemcpyDeviceToHost, st));
        RBX_HIP(hipStreamSynchronize(st));
        nnz = last_off + last_cnt;
    }
    if (next_state) {
        *next_state = seed;
        ctr_to_words(ctr_add(c, (uint64_t)sparse_incr(D)), next_state->counter);
    }
    return nnz;
}

// ---------------------------------------------------------------- COO -> row CSR
// key = r * m + c for entries inside the d x m window (after the offsets),
// key = d * m (sorts last) otherwise.
// Sort keys r * m + c of the entries inside the window (d * m for the
// others) and their positions; KT / PT are 32-bit when d * m and nnz fit.
template <typename KT, typename PT>
__global__ __launch_bounds__(256) void coo_keys_kernel(const int64_t *rows, const int64_t *cols, int64_t nnz,
                                                       int64_t ro, int64_t co, int64_t d, int64_t m, KT *keys,
                                                       PT *perm) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = rows[e] - ro, c = cols[e] - co;
        const bool in = r >= 0 && r < d && c >= 0 && c < m;
        keys[e] = (KT)(in ? (uint64_t)r * (uint64_t)m + (uint64_t)c : (uint64_t)d * (uint64_t)m);
        perm[e] = (PT)e;
    }
}

template <typename T, typename KT, typename PT>
__global__ __launch_bounds__(256) void csr_from_sorted_kernel(const KT *keys, const PT *perm, const T *vals,
                                                              int64_t nnz, int64_t d, int64_t m, int64_t *row_ptr,
                                                              int32_t *col, T *val) {
    const uint64_t dm = (uint64_t)d * (uint64_t)m;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < nnz; p += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t k = keys[p];
        const int64_t r = k < dm ? (int64_t)(k / (uint64_t)m) : d;
        const int64_t rp = p == 0 ? -1 : (keys[p - 1] < dm ? (int64_t)(keys[p - 1] / (uint64_t)m) : d);
        for (int64_t q = rp + 1; q <= r && q <= d; ++q) row_ptr[q] = p;
        if (p == nnz - 1)
            for (int64_t q = r + 1; q <= d; ++q) row_ptr[q] = nnz;
        if (k < dm) {
            col[p] = (int32_t)(k % (uint64_t)m);
            val[p] = vals[perm[p]];
        }
    }
}

// Row keys only, for entries whose columns already ascend within each row in
// stored order: a stable sort by row then gives the (row, column, position)
// order of the full key in fewer radix passes.
__global__ __launch_bounds__(256) void coo_row_keys_kernel(const int64_t *rows, const int64_t *cols, int64_t nnz,
                                                           int64_t ro, int64_t co, int64_t d, int64_t m,
                                                           uint32_t *keys, int32_t *perm) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = rows[e] - ro, c = cols[e] - co;
        const bool in = r >= 0 && r < d && c >= 0 && c < m;
        keys[e] = (uint32_t)(in ? r : d);
        perm[e] = (int32_t)e;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void csr_from_rowsorted_kernel(const uint32_t *keys, const int32_t *perm,
                                                                 const int64_t *cols, const T *vals, int64_t nnz,
                                                                 int64_t d, int64_t co, int64_t *row_ptr,
                                                                 int32_t *col, T *val) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < nnz; p += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = keys[p];
        const int64_t rp = p == 0 ? -1 : (int64_t)keys[p - 1];
        for (int64_t q = rp + 1; q <= r && q <= d; ++q) row_ptr[q] = p;
        if (p == nnz - 1)
            for (int64_t q = r + 1; q <= d; ++q) row_ptr[q] = nnz;
        if (r < d) {
            const int32_t e = perm[p];
            col[p] = (int32_t)(cols[e] - co);
            val[p] = vals[e];
        }
    }
}

__global__ void zero_rowptr_kernel(int64_t *row_ptr, int64_t d) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q <= d; q += (int64_t)gridDim.x * blockDim.x)
        row_ptr[q] = 0;
}

static int bits_for(uint64_t x) {
    int b = 0;
    while (b < 64 && (x >> b) != 0) ++b;
    return b < 1 ? 1 : b;
}

// ---------------------------------------------------------------- apply
template <typename T>
static void scale_only(T *C, int64_t d, int64_t n, int64_t crs, int64_t ccs, T beta, hipStream_t st);

template <typename T>
__global__ __launch_bounds__(256) void sp_scale_kernel(T *C, int64_t d, int64_t n, int64_t crs, int64_t ccs, T beta) {
    const int64_t total = d * n;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = idx % d, j = idx / d;
        T *c = C + i * crs + j * ccs;
        *c = (beta == (T)0) ? (T)0 : beta * (*c);
    }
}

template <typename T>
static void scale_only(T *C, int64_t d, int64_t n, int64_t crs, int64_t ccs, T beta, hipStream_t st) {
    if (d <= 0 || n <= 0) return;
    int64_t blocks = (d * n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(sp_scale_kernel<T>, dim3((unsigned)blocks), dim3(256), 0, st, C, d, n, crs, ccs, beta);
    RBX_HIP(hipGetLastError());
}

// Sparse rows times a dense row-major G (C = op(A_sparse) * G), vectorised:
// a workgroup owns 32 consecutive rows of C and a chunk of 64 * VEC columns;
// each lane holds VEC consecutive columns, so one 16-byte gather per stored
// entry covers the chunk.  A row's column
// indices and values are read 64 at a time with one coalesced load each and
// broadcast from registers, so the gathers of consecutive entries are
// independent.  Per output entry the entries are applied in stored order in
// one of the reference's three summation forms (below).  Column-major
// outputs go through an LDS tile and are written (and, with beta != 0, read)
// 32 rows = 128 B at a time.
constexpr int CRV_ROWS = 32;
template <typename T>
struct VecOf;
template <>
struct VecOf<float> {
    using type = float4;
    static constexpr int n = 4;
};
template <>
struct VecOf<double> {
    using type = double2;
    static constexpr int n = 2;
};

template <typename T>
__device__ __forceinline__ T vget(const typename VecOf<T>::type &v, int e) {
    return reinterpret_cast<const T *>(&v)[e];
}

// Summation forms, per output entry c with stored entries (v, b) in order:
//   CRV_JIK  c = RN(beta*c) + RN(alpha * sum RN(v*b))      (CSR dot products)
//   CRV_KIB  c = RN(beta*c); c = RN(c + RN(RN(alpha*v)*b))  (rank-1 updates)
//   CRV_JKI  c = RN(beta*c); c = RN(c + RN(v*RN(alpha*b)))  (CSC column kernel)
// (beta = 0 reads nothing and starts from 0; alpha = 0 leaves RN(beta*c).)
enum { CRV_JIK = 0, CRV_KIB = 1, CRV_JKI = 2 };

// NW waves per workgroup (8, or 16 when the rows are few).
template <typename T, typename IDX, bool COLMAJ, int NW>
__global__ __launch_bounds__(64 * NW) void csr_rows_vec_kernel(const int64_t *rowptr, const IDX *colidx, const T *vals,
                                                           int64_t mrows, int64_t ncols_out, T alpha, T beta,
                                                           const T *G, int64_t ldg, T *C, int64_t ldc, int form,
                                                           int cvec, int64_t lrow) {
    using V = typename VecOf<T>::type;
    constexpr int VEC = VecOf<T>::n, CW = 64 * VEC;
    __shared__ T tile[COLMAJ ? CRV_ROWS : 1][COLMAJ ? CW + 1 : 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t r0 = (int64_t)blockIdx.x * CRV_ROWS;
    const int64_t col0 = (int64_t)blockIdx.y * CW;
    const int64_t cl = col0 + (int64_t)lane * VEC;  // this lane's first column
    // a column-major C block is read through LDS (beta * C), 32 rows of a
    // column at a time
    if (COLMAJ && beta != (T)0) {
        // beta * C for the block, read 32 rows of a column at a time
        const int rr = tid & 31;
        for (int cc = tid >> 5; cc < CW; cc += 2 * NW) {
            const int64_t i = r0 + rr, j = col0 + cc;
            tile[rr][cc] = (i < mrows && j < ncols_out) ? beta * C[i + j * ldc] : (T)0;
        }
        __syncthreads();
    }
    for (int rr = wave; rr < CRV_ROWS; rr += NW) {
        const int64_t i = r0 + rr;
        if (i >= mrows) break;
        const int64_t p0 = rowptr[i], p1 = rowptr[i + 1];
        if (p1 - p0 > lrow) continue;  // a long row: summed in segments elsewhere
        T acc[VEC];
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
            T c0 = (T)0;
            if (form != CRV_JIK && beta != (T)0) {
                if (COLMAJ) c0 = tile[rr][lane * VEC + e];
                else if (cl + e < ncols_out) c0 = beta * C[i * ldc + cl + e];
            }
            acc[e] = c0;
        }
        if (alpha != (T)0) {
            for (int64_t pb = p0; pb < p1; pb += 64) {
                const int cnt = (int)(p1 - pb < 64 ? p1 - pb : 64);
                const int64_t myc = lane < cnt ? (int64_t)colidx[pb + lane] : 0;
                const T myv = lane < cnt ? vals[pb + lane] : (T)0;
                auto apply = [&](T v, const V &g) {
                    if (form == CRV_KIB) {
                        const T sc = alpha * v;
#pragma unroll
                        for (int e = 0; e < VEC; ++e) acc[e] = acc[e] + sc * vget<T>(g, e);
                    } else if (form == CRV_JKI) {
#pragma unroll
                        for (int e = 0; e < VEC; ++e) acc[e] = acc[e] + v * (alpha * vget<T>(g, e));
                    } else {
#pragma unroll
                        for (int e = 0; e < VEC; ++e) acc[e] = acc[e] + v * vget<T>(g, e);
                    }
                };
                // groups of 4 gathers in flight, the last group padded (its
                // entries past cnt gather row 0 of G, a valid address, and are
                // not applied; cnt is uniform, so is that test): no
                // one-at-a-time tail paying a full memory latency per entry
                for (int q = 0; q < cnt; q += 4) {
                    V g[4];
                    T v[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int64_t c = __shfl(myc, q + u);  // lanes past cnt hold column 0
                        v[u] = __shfl(myv, q + u);
                        if (cl < ncols_out) g[u] = *(const V *)(G + c * ldg + cl);
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        if (q + u < cnt) apply(v[u], g[u]);
                }
            }
        }
        // finish the JIK form: RN(beta*c) + RN(alpha*sum)
        if (form == CRV_JIK) {
#pragma unroll
            for (int e = 0; e < VEC; ++e) {
                T b0 = (T)0;
                if (beta != (T)0) {
                    if (COLMAJ) b0 = tile[rr][lane * VEC + e];
                    else if (cl + e < ncols_out) b0 = beta * C[i * ldc + cl + e];
                }
                acc[e] = alpha != (T)0 ? b0 + alpha * acc[e] : b0;
            }
        }
        if (COLMAJ) {
#pragma unroll
            for (int e = 0; e < VEC; ++e) tile[rr][lane * VEC + e] = acc[e];
        } else if (cvec && cl + VEC <= ncols_out) {
            V o;
#pragma unroll
            for (int e = 0; e < VEC; ++e) reinterpret_cast<T *>(&o)[e] = acc[e];
            *(V *)(C + i * ldc + cl) = o;
        } else {
#pragma unroll
            for (int e = 0; e < VEC; ++e)
                if (cl + e < ncols_out) C[i * ldc + cl + e] = acc[e];
        }
    }
    if (COLMAJ) {
        // 32 rows of a column per pass: whole 128 B segments
        __syncthreads();
        const int rr = tid & 31;
        const int64_t i = r0 + rr;
        const bool mine = i < mrows && rowptr[i + 1] - rowptr[i] <= lrow;
        for (int cc = tid >> 5; cc < CW; cc += 2 * NW) {
            const int64_t j = col0 + cc;
            if (mine && j < ncols_out) C[i + j * ldc] = tile[rr][cc];
        }
    }
}

// op(B) (m x n, strides brs / bcs) -> row-major copy with leading dimension
// ldo, through 32 x 32 LDS tiles (coalesced on both sides).
template <typename T>
__global__ __launch_bounds__(256) void to_rowmajor_kernel(const T *B, int64_t brs, int64_t bcs, int64_t m, int64_t n,
                                                          T *out, int64_t ldo) {
    __shared__ T t[32][33];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const int64_t c0 = (int64_t)blockIdx.x * 32, j0 = (int64_t)blockIdx.y * 32;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int a = ty + 8 * q;
        // column-major source: lanes walk c; row-major: lanes walk j
        const int64_t c = brs == 1 ? c0 + tx : c0 + a, j = brs == 1 ? j0 + a : j0 + tx;
        const T v = (c < m && j < n) ? B[c * brs + j * bcs] : (T)0;
        if (brs == 1) t[tx][a] = v;
        else t[a][tx] = v;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int a = ty + 8 * q;
        const int64_t c = c0 + a, j = j0 + tx;
        if (c < m && j < ldo) out[c * ldo + j] = j < n ? t[a][tx] : (T)0;
    }
}

// COO entries inside the window [ro, ro + d) x [co, co + m) -> CSR (row
// pointers, 32-bit columns, values) sorted by (row, column), stable for
// duplicates, in library scratch.
template <typename T, typename KT, typename PT>
static void coo_to_csr(const int64_t *rows, const int64_t *cols, const T *vals, int64_t nnz, int64_t ro, int64_t co,
                       int64_t d, int64_t m, int64_t **row_ptr_out, int32_t **col_out, T **val_out, hipStream_t st) {
    size_t sort_bytes = 0;
    const int end_bit = bits_for((uint64_t)d * (uint64_t)m);
    RBX_HIP(rocprim::radix_sort_pairs(nullptr, sort_bytes, (KT *)nullptr, (KT *)nullptr, (PT *)nullptr, (PT *)nullptr,
                                      (size_t)nnz, 0, end_bit, st));
    const size_t bk = align256((size_t)nnz * sizeof(KT)), bp = align256((size_t)nnz * sizeof(PT));
    const size_t brp = align256((size_t)(d + 1) * 8), bcol = align256((size_t)nnz * 4),
                 bval = align256((size_t)nnz * sizeof(T));
    char *ws = (char *)scratch(SCRATCH_SPARSE_CSR, 2 * bk + 2 * bp + brp + bcol + bval + align256(sort_bytes), st);
    KT *keys_in = (KT *)ws, *keys_out = (KT *)(ws + bk);
    PT *perm_in = (PT *)(ws + 2 * bk), *perm_out = (PT *)(ws + 2 * bk + bp);
    char *rest = ws + 2 * bk + 2 * bp;
    int64_t *row_ptr = (int64_t *)rest;
    int32_t *ccol = (int32_t *)(rest + brp);
    T *cval = (T *)(rest + brp + bcol);
    void *sort_tmp = rest + brp + bcol + bval;
    int64_t blocks = (nnz + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    hipLaunchKernelGGL((coo_keys_kernel<KT, PT>), dim3((unsigned)blocks), dim3(256), 0, st, rows, cols, nnz, ro, co, d,
                       m, keys_in, perm_in);
    RBX_HIP(hipGetLastError());
    RBX_HIP(rocprim::radix_sort_pairs(sort_tmp, sort_bytes, keys_in, keys_out, perm_in, perm_out, (size_t)nnz, 0,
                                      end_bit, st));
    hipLaunchKernelGGL(zero_rowptr_kernel, dim3(64), dim3(256), 0, st, row_ptr, d);
    hipLaunchKernelGGL((csr_from_sorted_kernel<T, KT, PT>), dim3((unsigned)blocks), dim3(256), 0, st, keys_out,
                       perm_out, vals, nnz, d, m, row_ptr, ccol, cval);
    RBX_HIP(hipGetLastError());
    *row_ptr_out = row_ptr;
    *col_out = ccol;
    *val_out = cval;
}

// coo_to_csr for entries whose columns ascend within each row in stored order
// (the expansion of a CSR matrix viewed transposed, or of a CSC matrix): the
// same CSR from a stable radix sort on the row alone.
template <typename T>
static void coo_to_csr_by_row(const int64_t *rows, const int64_t *cols, const T *vals, int64_t nnz, int64_t ro,
                              int64_t co, int64_t d, int64_t m, int64_t **row_ptr_out, int32_t **col_out,
                              T **val_out, hipStream_t st) {
    size_t sort_bytes = 0;
    const int end_bit = bits_for((uint64_t)d);
    RBX_HIP(rocprim::radix_sort_pairs(nullptr, sort_bytes, (uint32_t *)nullptr, (uint32_t *)nullptr,
                                      (int32_t *)nullptr, (int32_t *)nullptr, (size_t)nnz, 0, end_bit, st));
    const size_t bk = align256((size_t)nnz * 4), bp = align256((size_t)nnz * 4);
    const size_t brp = align256((size_t)(d + 1) * 8), bcol = align256((size_t)nnz * 4),
                 bval = align256((size_t)nnz * sizeof(T));
    char *ws = (char *)scratch(SCRATCH_SPARSE_CSR, 2 * bk + 2 * bp + brp + bcol + bval + align256(sort_bytes), st);
    uint32_t *keys_in = (uint32_t *)ws, *keys_out = (uint32_t *)(ws + bk);
    int32_t *perm_in = (int32_t *)(ws + 2 * bk), *perm_out = (int32_t *)(ws + 2 * bk + bp);
    char *rest = ws + 2 * bk + 2 * bp;
    int64_t *row_ptr = (int64_t *)rest;
    int32_t *ccol = (int32_t *)(rest + brp);
    T *cval = (T *)(rest + brp + bcol);
    void *sort_tmp = rest + brp + bcol + bval;
    int64_t blocks = (nnz + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    hipLaunchKernelGGL(coo_row_keys_kernel, dim3((unsigned)blocks), dim3(256), 0, st, rows, cols, nnz, ro, co, d, m,
                       keys_in, perm_in);
    RBX_HIP(hipGetLastError());
    RBX_HIP(rocprim::radix_sort_pairs(sort_tmp, sort_bytes, keys_in, keys_out, perm_in, perm_out, (size_t)nnz, 0,
                                      end_bit, st));
    hipLaunchKernelGGL(zero_rowptr_kernel, dim3(64), dim3(256), 0, st, row_ptr, d);
    hipLaunchKernelGGL((csr_from_rowsorted_kernel<T>), dim3((unsigned)blocks), dim3(256), 0, st, keys_out, perm_out,
                       cols, vals, nnz, d, co, row_ptr, ccol, cval);
    RBX_HIP(hipGetLastError());
    *row_ptr_out = row_ptr;
    *col_out = ccol;
    *val_out = cval;
}

// ---------------------------------------------------------------- long rows
// The row kernels give each row to one wave, which applies the row's entries
// in stored order: a row of 10^5 entries (the transposed view of a matrix with
// power-law column degrees) then keeps one wave busy, latency-bound, long
// after the rest of the GPU has finished.  Rows of more than
//   lrow = max(256, 4 x the mean row length)
// entries (so rows of similar lengths, as a sparse operator's, are not split,
// unless they are too few to fill the GPU: long_row_threshold) are cut into
// segments of at most `seg` entries instead: each segment
// is summed by its own wave in FP64 (sum of v * g in stored order), and the
// row's outputs are combined from the segment sums in segment order,
//   c = RN(RN(beta * c) + RN(alpha * sum_s P_s))   (FP64, then rounded to T).
// Deterministic; for such rows the rounding differs from the reference's
// sequential forms by at most the rounding of the sum.  Shorter rows keep the
// row kernels' exact forms.
constexpr int64_t LONG_ROW_MIN = 256, LONG_ROW_MEAN_FACTOR = 4, LONG_SEG = 256;

__global__ void long_rows_clear_kernel(int64_t *count) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *count = 0;
}

// list[0 .. *count) = the rows longer than lrow (in no particular order)
__global__ __launch_bounds__(256) void long_rows_find_kernel(const int64_t *rowptr, int64_t mrows, int64_t lrow,
                                                             int64_t cap, int64_t *count, int64_t *list) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < mrows; i += (int64_t)gridDim.x * blockDim.x)
        if (rowptr[i + 1] - rowptr[i] > lrow) {
            const int64_t pos = (int64_t)atomicAdd((unsigned long long *)count, 1ull);
            if (pos < cap) list[pos] = i;
        }
}

// seg_off[k] = sum over list entries before k of ceil(len / seg); one
// workgroup, 256 entries per pass
__global__ __launch_bounds__(256) void long_rows_plan_kernel(const int64_t *rowptr, const int64_t *count,
                                                             const int64_t *list, int64_t cap, int64_t seg,
                                                             int64_t *seg_off) {
    __shared__ int64_t sh[256];
    const int tid = threadIdx.x;
    const int64_t n = *count < cap ? *count : cap;
    int64_t carry = 0;
    for (int64_t base = 0; base < n; base += 256) {
        const int64_t k = base + tid;
        int64_t v = 0;
        if (k < n) {
            const int64_t row = list[k];
            v = (rowptr[row + 1] - rowptr[row] + seg - 1) / seg;
        }
        sh[tid] = v;
        __syncthreads();
        for (int off = 1; off < 256; off <<= 1) {
            const int64_t add = tid >= off ? sh[tid - off] : 0;
            __syncthreads();
            sh[tid] += add;
            __syncthreads();
        }
        if (k < n) seg_off[k] = carry + sh[tid] - v;
        carry += sh[255];
        __syncthreads();
    }
    if (tid == 0) seg_off[n] = carry;
}

// part[gs * ldp + j] = sum over segment gs's entries of v * G[col, j], in
// FP64.  Waves take segments grid-stride; a lane covers VEC consecutive
// columns (VEC > 1: 16-byte rows of G, ncols_out a multiple of VEC) and
// blockIdx.y picks 64 VEC columns.
template <typename T, typename IDX, int VEC>
__global__ __launch_bounds__(256) void long_rows_partial_kernel(const int64_t *rowptr, const IDX *colidx,
                                                                const T *vals, int64_t ncols_out, const T *G,
                                                                int64_t ldg, const int64_t *count, const int64_t *list,
                                                                const int64_t *seg_off, int64_t cap, int64_t nseg_cap,
                                                                int64_t seg, double *part, int64_t ldp) {
    using V = std::conditional_t<(VEC > 1), typename VecOf<T>::type, T>;
    constexpr int NB = VEC > 1 ? 12 : 16;  // gathers in flight per lane
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t n = *count < cap ? *count : cap;
    if (n == 0) return;
    const int64_t total = seg_off[n] < nseg_cap ? seg_off[n] : nseg_cap;
    const int64_t j = ((int64_t)blockIdx.y * 64 + lane) * VEC;
    const bool jin = j < ncols_out;
    for (int64_t gs = (int64_t)blockIdx.x * 4 + wave; gs < total; gs += (int64_t)gridDim.x * 4) {
        int64_t lo = 0, hi = n - 1;  // last list entry whose segments start at or before gs
        while (lo < hi) {
            const int64_t mid = (lo + hi + 1) >> 1;
            if (seg_off[mid] <= gs) lo = mid;
            else hi = mid - 1;
        }
        const int64_t row = list[lo];
        const int64_t pend = rowptr[row + 1];
        const int64_t p0 = rowptr[row] + (gs - seg_off[lo]) * seg;
        const int64_t p1 = p0 + seg < pend ? p0 + seg : pend;
        double acc[VEC];
#pragma unroll
        for (int e = 0; e < VEC; ++e) acc[e] = 0.0;
        auto load = [&](int64_t c) {
            V g;
            if (jin) {
                g = *(const V *)(G + c * ldg + j);
            } else {
#pragma unroll
                for (int e = 0; e < VEC; ++e) reinterpret_cast<T *>(&g)[e] = (T)0;
            }
            return g;
        };
        auto add = [&](T v, const V &g) {
#pragma unroll
            for (int e = 0; e < VEC; ++e) acc[e] += (double)v * (double)reinterpret_cast<const T *>(&g)[e];
        };
        for (int64_t pb = p0; pb < p1; pb += 64) {
            const int cnt = (int)(p1 - pb < 64 ? p1 - pb : 64);
            const int64_t myc = lane < cnt ? (int64_t)colidx[pb + lane] : 0;
            const T myv = lane < cnt ? vals[pb + lane] : (T)0;
            int q = 0;
            for (; q + NB <= cnt; q += NB) {
                V g[NB];
                T v[NB];
#pragma unroll
                for (int u = 0; u < NB; ++u) {
                    v[u] = __shfl(myv, q + u);
                    g[u] = load(__shfl(myc, q + u));
                }
#pragma unroll
                for (int u = 0; u < NB; ++u) add(v[u], g[u]);
            }
            for (; q < cnt; ++q) add(__shfl(myv, q), load(__shfl(myc, q)));
        }
        if (jin) {
#pragma unroll
            for (int e = 0; e < VEC; ++e) part[gs * ldp + j + e] = acc[e];
        }
    }
}

// C[row, j] = RN(RN(beta * C) + RN(alpha * sum of the row's segment sums))
template <typename T>
__global__ __launch_bounds__(256) void long_rows_combine_kernel(const int64_t *count, const int64_t *list,
                                                                const int64_t *seg_off, int64_t cap, int64_t nseg_cap,
                                                                const double *part, int64_t ldp, int64_t ncols_out,
                                                                T alpha, T beta, T *C, int64_t crs, int64_t ccs) {
    const int64_t n = *count < cap ? *count : cap;
    const int64_t j = (int64_t)blockIdx.y * 256 + threadIdx.x;
    if (j >= ncols_out) return;
    for (int64_t k = blockIdx.x; k < n; k += gridDim.x) {
        const int64_t row = list[k];
        const int64_t s1 = seg_off[k + 1] < nseg_cap ? seg_off[k + 1] : nseg_cap;
        double sum = 0.0;
        int64_t sgi = seg_off[k];
        for (; sgi + 8 <= s1; sgi += 8) {  // eight loads in flight, added in order
            double x[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) x[u] = part[(sgi + u) * ldp + j];
#pragma unroll
            for (int u = 0; u < 8; ++u) sum += x[u];
        }
        for (; sgi < s1; ++sgi) sum += part[sgi * ldp + j];
        T *c = C + row * crs + j * ccs;
        double out = (double)alpha * sum;
        if (beta != (T)0) out = (double)beta * (double)(*c) + out;
        *c = (T)out;
    }
}

// The long-row threshold of a row-kernel call (rows longer than it are split).
// Fewer rows than eight per CU, averaging more than 4096 entries (a sparse
// operator with a small sketch dimension: 1000 rows of 8000 entries), keep
// each of the few waves on a long latency-bound chain even when the rows are
// alike: then every row longer than LONG_ROW_MIN is split.
static int64_t long_row_threshold(int64_t nnz, int64_t mrows) {
    const int64_t mean = (nnz + mrows - 1) / mrows;
    if (mrows < 8 * (int64_t)num_cus() && mean > 4096) return LONG_ROW_MIN;
    const int64_t t = LONG_ROW_MEAN_FACTOR * mean;
    return t > LONG_ROW_MIN ? t : LONG_ROW_MIN;
}

// Sets up and launches the long-row path for a row-kernel call; returns the
// length above which the row kernels must skip a row (INT64_MAX: no split).
// `nnz` bounds the total number of entries the rows hold.  G's rows are
// unit-stride (G[c * ldg + j]); C[i, j] is at C[i * crs + j * ccs].
template <typename T, typename IDX>
static int64_t launch_long_rows(const int64_t *rowptr, const IDX *colidx, const T *vals, int64_t mrows, int64_t nnz,
                                int64_t ncols_out, T alpha, T beta, const T *G, int64_t ldg, T *C, int64_t crs,
                                int64_t ccs, hipStream_t st) {
    if (alpha == (T)0 || mrows == 0 || ncols_out == 0) return INT64_MAX;
    int64_t lrow = long_row_threshold(nnz, mrows);
    constexpr int VN = VecOf<T>::n;
    const bool vec = ncols_out % VN == 0 && ldg % VN == 0 && ((uintptr_t)G % 16) == 0;
    const int64_t cw = vec ? 64 * VN : 64;
    const int64_t ldp = (ncols_out + cw - 1) / cw * cw;
    // keep the segment sums within 512 MiB: longer segments (and a higher
    // threshold) for wide outputs; no split when even that cannot fit
    int64_t seg = LONG_SEG;
    auto sums_bytes = [&]() { return ((double)(nnz / seg) + (double)(nnz / (lrow + 1)) + 2.0) * (double)ldp * 8.0; };
    while (sums_bytes() > (double)(512ll << 20)) {
        if (seg >= nnz) return INT64_MAX;
        seg *= 2;
        if (lrow < seg) lrow = seg;
    }
    if (nnz <= lrow) return INT64_MAX;  // no row can be long
    const int64_t cap = nnz / (lrow + 1) + 1;      // rows longer than lrow
    const int64_t nseg_cap = nnz / seg + cap + 1;  // their segments
    const size_t b_list = align256((size_t)cap * 8), b_off = align256((size_t)(cap + 1) * 8);
    char *ws = (char *)scratch(SCRATCH_LONG_ROWS,
                               256 + b_list + b_off + (size_t)safe_mul(safe_mul(nseg_cap, ldp), 8), st);
    int64_t *count = (int64_t *)ws, *list = (int64_t *)(ws + 256), *seg_off = (int64_t *)(ws + 256 + b_list);
    double *part = (double *)(ws + 256 + b_list + b_off);
    ProfScope prof("long_rows", st);
    hipLaunchKernelGGL(long_rows_clear_kernel, dim3(1), dim3(64), 0, st, count);
    int64_t fb = (mrows + 255) / 256;
    if (fb > 4096) fb = 4096;
    hipLaunchKernelGGL(long_rows_find_kernel, dim3((unsigned)fb), dim3(256), 0, st, rowptr, mrows, lrow, cap, count,
                       list);
    hipLaunchKernelGGL(long_rows_plan_kernel, dim3(1), dim3(256), 0, st, rowptr, (const int64_t *)count,
                       (const int64_t *)list, cap, seg, seg_off);
    const int64_t ny = ldp / cw;
    RBX_REQUIRE(ny <= 65535 && (ncols_out + 255) / 256 <= 65535);
    int64_t px = (nseg_cap + 3) / 4;
    if (px > 4096) px = 4096;
    const dim3 pgrid((unsigned)px, (unsigned)ny);
    if (vec)
        hipLaunchKernelGGL((long_rows_partial_kernel<T, IDX, VN>), pgrid, dim3(256), 0, st, rowptr, colidx, vals,
                           ncols_out, G, ldg, (const int64_t *)count, (const int64_t *)list, (const int64_t *)seg_off,
                           cap, nseg_cap, seg, part, ldp);
    else
        hipLaunchKernelGGL((long_rows_partial_kernel<T, IDX, 1>), pgrid, dim3(256), 0, st, rowptr, colidx, vals,
                           ncols_out, G, ldg, (const int64_t *)count, (const int64_t *)list, (const int64_t *)seg_off,
                           cap, nseg_cap, seg, part, ldp);
    int64_t cx = cap < 1024 ? cap : 1024;
    hipLaunchKernelGGL((long_rows_combine_kernel<T>), dim3((unsigned)cx, (unsigned)((ncols_out + 255) / 256)),
                       dim3(256), 0, st, (const int64_t *)count, (const int64_t *)list, (const int64_t *)seg_off, cap,
                       nseg_cap, (const double *)part, ldp, ncols_out, alpha, beta, C, crs, ccs);
    RBX_HIP(hipGetLastError());
    return lrow;
}

// Launch csr_rows_vec_kernel over mrows x ncols_out: 8 waves per workgroup
// (4 rows per wave: rows of similar length that cluster in one 32-row block
// are shared by more waves than with 4), 16 when 8-wave workgroups would
// leave the GPU short of waves.  Rows longer than the long-row threshold go
// to the segment kernels.
template <typename T, typename IDX>
static void launch_rows_vec(bool colmaj, const int64_t *rowptr, const IDX *colidx, const T *vals, int64_t mrows,
                            int64_t nnz, int64_t ncols_out, T alpha, T beta, const T *G, int64_t ldg, T *C,
                            int64_t ldc, int form, int cvec, hipStream_t st) {
    constexpr int VEC = VecOf<T>::n;
    const int64_t lrow = launch_long_rows<T, IDX>(rowptr, colidx, vals, mrows, nnz, ncols_out, alpha, beta, G, ldg, C,
                                                  colmaj ? 1 : ldc, colmaj ? ldc : 1, st);
    const int64_t nchunk = (ncols_out + 64 * VEC - 1) / (64 * VEC);
    const int64_t nrb = (mrows + CRV_ROWS - 1) / CRV_ROWS;
    RBX_REQUIRE(nchunk <= 65535 && nrb < (1ll << 31));
    const dim3 grid((unsigned)nrb, (unsigned)nchunk);
    const bool wide = nrb * nchunk < 4 * (int64_t)num_cus();
#define RBX_ROWS_VEC(CM, NW)                                                                                  \
    hipLaunchKernelGGL((csr_rows_vec_kernel<T, IDX, CM, NW>), grid, dim3(64 * NW), 0, st, rowptr, colidx, vals, \
                       mrows, ncols_out, alpha, beta, G, ldg, C, ldc, form, cvec, lrow)
    if (colmaj && wide) RBX_ROWS_VEC(true, 16);
    else if (colmaj) RBX_ROWS_VEC(true, 8);
    else if (wide) RBX_ROWS_VEC(false, 16);
    else RBX_ROWS_VEC(false, 8);
#undef RBX_ROWS_VEC
    RBX_HIP(hipGetLastError());
}

// Longest mean row (stored entries per output row) the LDS apply takes: each
// output is one thread's sequential chain, so rows much longer than this are
// better served by the row kernels' segment sums.
constexpr int64_t LDS_APPLY_MAX_MEAN_ROW = 2048;

// A sampled operator's prepared copy (rbx_sparse_skop::prep): its entries as
// a CSR matrix sorted by (row, column, stored position), as left_spmm_coo
// builds it per call, and the LDS apply's entry layout over it.  `vals`,
// `rows`, `cols`, `nnz` and the device identify the arrays it was built from.
struct SparsePrep {
    uint64_t magic;
    int dev;
    size_t val_size;
    int64_t d, m, nnz;
    const void *vals;
    const int64_t *rows, *cols;
    char *mem;  // one allocation: row_ptr, col, val, then the layout
    int64_t *row_ptr;
    int32_t *col;
    void *val;
    char *lay;
    size_t lay_bytes;
    bool unit;  // every value is exactly +1 or -1 (the layout carries the signs)
};
constexpr uint64_t PREP_MAGIC = 0x70657270787862ull;
// the copy may take at most this many times the operator's own bytes
constexpr size_t PREP_MAX_RATIO = 4;

// left_spmm for a COO operator (spmm_dispatch.hh:52-179):
// C = alpha * op(submat(A)) * op(B) + beta * C, A given by COO arrays.  `prep`,
// when given, is the prepared copy of exactly these entries (the whole
// operator, opA = 'N'); the LDS apply then reads it instead of sorting.
template <typename T>
static void left_spmm_coo(char layout, char opA, char opB, int64_t d, int64_t n, int64_t m, T alpha, int64_t A_rows,
                          int64_t A_cols, int64_t nnz, const T *vals, const int64_t *rows, const int64_t *cols,
                          int64_t ro_a, int64_t co_a, const T *B, int64_t ldb, T beta, T *C, int64_t ldc,
                          hipStream_t st, int kib_force = -1, char *fixed_ws = nullptr, size_t fixed_ws_bytes = 0,
                          const SparsePrep *prep = nullptr, bool cols_ascend_in_rows = false, bool unit = false) {
    check_layout(layout);
    check_op(opA);
    check_op(opB);
    RBX_REQUIRE(d >= 0 && n >= 0 && m >= 0);
    if (opA == 'T') {  // transposed view (spmm_dispatch.hh:72-90)
        const int64_t *t = rows; rows = cols; cols = t;
        int64_t x = A_rows; A_rows = A_cols; A_cols = x;
        x = ro_a; ro_a = co_a; co_a = x;
    }
    RBX_REQUIRE(A_rows >= d);
    RBX_REQUIRE(A_cols >= m);
    int64_t rows_B = opB == 'N' ? m : n, cols_B = opB == 'N' ? n : m;
    char layout_opB = opB == 'N' ? layout : (layout == 'C' ? 'R' : 'C');
    if (layout == 'C') {
        RBX_REQUIRE(ldb >= rows_B);
        RBX_REQUIRE(ldc >= d);
    } else {
        RBX_REQUIRE(ldc >= n);
        RBX_REQUIRE(ldb >= cols_B);
    }
    const int64_t crs = layout == 'C' ? 1 : ldc, ccs = layout == 'C' ? ldc : 1;
    if (d == 0 || n == 0) return;
    if (alpha == (T)0 || m == 0 || nnz <= 0) {
        scale_only<T>(C, d, n, crs, ccs, beta, st);
        return;
    }
    RBX_REQUIRE(m < (int64_t)INT32_MAX);
    // op(B)[c, j] strides
    const int64_t brs = layout_opB == 'C' ? 1 : ldb, bcs = layout_opB == 'C' ? ldb : 1;
    // summation form of the reference's kernel (coo_spmm_impl.hh:94-101); a
    // caller whose B is a re-laid-out copy of the reference's B passes the form
    // the reference's own layout selects
    const int kib = kib_force >= 0 ? kib_force : ((layout_opB == 'R' && layout == 'R') ? 1 : 0);
    safe_mul(d + 1, m);
    if (prep != nullptr && layout_opB == 'C') {
        RBX_REQUIRE(opA == 'N' && ro_a == 0 && co_a == 0 && d == prep->d && m == prep->m && nnz == prep->nnz);
        ProfScope prof("csr_lds_apply_kernel", st);
        if (csr_lds_apply_laid<T>(prep->row_ptr, prep->col, (const T *)prep->val, d, n, m, nnz, alpha, beta, B, ldb, C,
                                  crs, ccs, kib ? SPF_KIB : SPF_JKI, prep->lay, prep->lay_bytes, st, prep->unit))
            return;
    }

    // the window's entries as a CSR matrix sorted by (row, column), in scratch
    int64_t *row_ptr;
    int32_t *ccol;
    T *cval;
    if (fixed_ws != nullptr) {
        // caller-provided workspace: the counting-sort build and the LDS apply
        // only (neither allocates)
        if (layout_opB != 'C')
            RBX_FAIL(RBX_UNSUPPORTED, "%s: the allocation-free path needs op(A) with unit stride along the "
                     "contraction (column-major A with opA = 'N', or row-major A with opA = 'T')", __func__);
        if (d > CSRB_MAX_ROWS || nnz > LDS_APPLY_MAX_MEAN_ROW * d)
            RBX_FAIL(RBX_UNSUPPORTED, "%s: the allocation-free path takes at most %lld sketch rows and %lld stored "
                     "entries per row on average (d = %lld, nnz = %lld)", __func__, (long long)CSRB_MAX_ROWS,
                     (long long)LDS_APPLY_MAX_MEAN_ROW, (long long)d, (long long)nnz);
        const size_t need_csr = csr_build_counting_bytes(nnz, d, sizeof(T));
        if (need_csr == 0) RBX_FAIL(RBX_UNSUPPORTED, "%s: nnz = %lld is outside the allocation-free path", __func__,
                                    (long long)nnz);
        const size_t need_apply = csr_lds_apply_ws_bytes(d, m, nnz, sizeof(T));
        const size_t need = need_csr + need_apply;
        if (need > fixed_ws_bytes)
            RBX_FAIL(RBX_INVALID_ARG, "%s: the operator window needs %zu bytes of scratch; the allocation-free "
                     "entry point holds %zu", __func__, need, fixed_ws_bytes);
        if (!csr_build_counting_ws<T>(rows, cols, vals, nnz, ro_a, co_a, d, m, fixed_ws, need_csr, &row_ptr, &ccol,
                                      &cval, st))
            RBX_FAIL(RBX_UNSUPPORTED, "%s: shape outside the allocation-free path", __func__);
        if (!csr_lds_apply<T>(row_ptr, ccol, cval, d, n, m, nnz, alpha, beta, B, ldb, C, crs, ccs,
                              kib ? SPF_KIB : SPF_JKI, fixed_ws + need_csr, need_apply, st))
            RBX_FAIL(RBX_UNSUPPORTED, "%s: shape outside the allocation-free path", __func__);
        return;
    }
    {
        ProfScope prof("coo_to_csr", st);
        // few output rows: stable counting sort by row; otherwise a radix
        // sort on the (row, column) key
        if (csr_build_counting<T>(rows, cols, vals, nnz, ro_a, co_a, d, m, &row_ptr, &ccol, &cval, st)) {
        } else if (cols_ascend_in_rows && d < (int64_t)UINT32_MAX && nnz < (int64_t)INT32_MAX)
            coo_to_csr_by_row<T>(rows, cols, vals, nnz, ro_a, co_a, d, m, &row_ptr, &ccol, &cval, st);
        else if ((uint64_t)d * (uint64_t)m < (1ull << 32) && nnz < (int64_t)INT32_MAX)
            coo_to_csr<T, uint32_t, int32_t>(rows, cols, vals, nnz, ro_a, co_a, d, m, &row_ptr, &ccol, &cval, st);
        else
            coo_to_csr<T, uint64_t, int64_t>(rows, cols, vals, nnz, ro_a, co_a, d, m, &row_ptr, &ccol, &cval, st);
    }
    // op(B) with unit stride along the contraction (a column-major operand of
    // a left sketch), few output rows that are not very long: every output a
    // sequential chain, op(B) streamed once through LDS
    if (layout_opB == 'C' && d <= CSRB_MAX_ROWS && nnz <= LDS_APPLY_MAX_MEAN_ROW * d) {
        ProfScope prof("csr_lds_apply_kernel", st);
        const size_t wsb = csr_lds_apply_ws_bytes(d, m, nnz, sizeof(T));
        char *lws = (char *)scratch(SCRATCH_SPARSE_G, wsb, st);
        if (csr_lds_apply<T>(row_ptr, ccol, cval, d, n, m, nnz, alpha, beta, B, ldb, C, crs, ccs,
                             kib ? SPF_KIB : SPF_JKI, lws, wsb, st, unit))
            return;
    }
    {
        ProfScope prof("csr_apply_kernel", st);
        // op(B) as an aligned row-major matrix: used in place when it is one,
        // else copied (transposed) into scratch
        constexpr int VEC = VecOf<T>::n;
        const T *Gp = B;
        int64_t ldg = ldb;
        if (!(layout_opB == 'R' && ldb % VEC == 0 && ((uintptr_t)B % 16) == 0)) {
            ldg = (n + VEC - 1) / VEC * VEC;
            T *Gs = (T *)scratch(SCRATCH_SPARSE_G, (size_t)safe_mul(m, ldg) * sizeof(T), st);
            const int64_t gx = (m + 31) / 32, gy = (ldg + 31) / 32;
            RBX_REQUIRE(gx < (1ll << 31) && gy <= 65535);
            hipLaunchKernelGGL(to_rowmajor_kernel<T>, dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, st, B, brs, bcs,
                               m, n, Gs, ldg);
            RBX_HIP(hipGetLastError());
            Gp = Gs;
        }
        const int cvec = layout == 'R' && ldc % VEC == 0 && ((uintptr_t)C % 16) == 0;
        launch_rows_vec<T, int32_t>(layout == 'C', row_ptr, ccol, cval, d, nnz, n, alpha, beta, Gp, ldg, C, ldc,
                                    kib ? CRV_KIB : CRV_JKI, cvec, st);
        RBX_HIP(hipGetLastError());
    }
}

// Sampled operator arrays for the duration of a call (skge.hh:483-487).
template <typename T>
struct SkopView {
    int64_t nnz;
    const T *vals;
    const int64_t *rows, *cols;
    bool unit;  // sampled in this call as a SASO: every value is +1 or -1
};

template <typename T>
static SkopView<T> view_or_sample(const rbx_sparse_skop *S, hipStream_t st) {
    RBX_REQUIRE(S != nullptr);
    SparseGeom g = sparse_geom(S->dist);
    if (S->nnz >= 0) {
        RBX_REQUIRE(S->nnz == 0 || (S->vals && S->rows && S->cols));
        RBX_REQUIRE(S->nnz <= g.full_nnz);
        return {S->nnz, (const T *)S->vals, S->rows, S->cols, false};
    }
    const size_t bv = align256((size_t)g.full_nnz * sizeof(T)), bi = align256((size_t)g.full_nnz * 8);
    // The sampled operator lives in the dense-operator slot: LASO sampling
    // itself uses the sparse-operator slot for its temporaries.
    char *ws = (char *)scratch(SCRATCH_DENSE_OP, bv + 2 * bi, st);
    T *vals = (T *)ws;
    int64_t *rows = (int64_t *)(ws + bv), *cols = (int64_t *)(ws + bv + bi);
    int64_t nnz = fill_sparse_impl<T>(S->dist, vals, rows, cols, S->seed_state, nullptr, st);
    return {nnz, vals, rows, cols, S->dist.major_axis == 'S'};
}

// S's prepared copy when it was built from S's current arrays in T on this
// device and the call covers the whole operator with opS = 'N'; else null.
template <typename T>
static const SparsePrep *usable_prep(const rbx_sparse_skop *S, char opS, int64_t d, int64_t m, int64_t ro_s,
                                     int64_t co_s) {
    const SparsePrep *p = (const SparsePrep *)S->prep;
    if (p == nullptr || p->magic != PREP_MAGIC || p->val_size != sizeof(T)) return nullptr;
    if (opS != 'N' || ro_s != 0 || co_s != 0 || d != S->dist.n_rows || m != S->dist.n_cols) return nullptr;
    if (S->nnz != p->nnz || S->vals != p->vals || S->rows != p->rows || S->cols != p->cols) return nullptr;
    if (p->d != d || p->m != m) return nullptr;
    int dev = -1;
    RBX_HIP(hipGetDevice(&dev));
    return dev == p->dev ? p : nullptr;
}

static void prep_free(rbx_sparse_skop *S) {
    SparsePrep *p = (SparsePrep *)S->prep;
    S->prep = nullptr;
    if (p == nullptr || p->magic != PREP_MAGIC) return;
    p->magic = 0;
    // hipFree waits for the work queued on the device, so no queued apply
    // still reads the copy
    if (p->mem != nullptr) (void)hipFree(p->mem);
    delete p;
}

__global__ void prep_flag_zero_kernel(int *f) { *f = 0; }
// *f |= 1 when some value is not exactly +1 or -1 (one atomic per wave that saw one)
template <typename T>
__global__ __launch_bounds__(256) void unit_values_kernel(const T *v, int64_t n, int *f) {
    bool other = false;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        other |= !(v[i] == (T)1 || v[i] == (T)-1);
    if (__any(other) && (threadIdx.x & 63) == 0) atomicOr(f, 1);
}

// rbx_[sd]sparse_skop_prepare: S->prep for S's current arrays, or null when
// the operator is outside the LDS apply or the memory cannot be had.  Finding
// whether every value is +1 or -1 (a SASO's are) reads one flag back, so this
// synchronises the stream.
template <typename T>
static void prepare_skop(rbx_sparse_skop *S, hipStream_t st) {
    prep_free(S);
    if (S->nnz <= 0 || S->vals == nullptr || S->rows == nullptr || S->cols == nullptr) return;
    const int64_t d = S->dist.n_rows, m = S->dist.n_cols, nnz = S->nnz;
    if (d > CSRB_MAX_ROWS || nnz > LDS_APPLY_MAX_MEAN_ROW * d || m >= (int64_t)INT32_MAX) return;
    bool unit = false;
    {
        int *f = (int *)scratch(SCRATCH_SMALL, sizeof(int), st);
        hipLaunchKernelGGL(prep_flag_zero_kernel, dim3(1), dim3(1), 0, st, f);
        int64_t nb = (nnz + 255) / 256;
        if (nb > 4 * (int64_t)num_cus()) nb = 4 * (int64_t)num_cus();
        hipLaunchKernelGGL((unit_values_kernel<T>), dim3((unsigned)nb), dim3(256), 0, st, (const T *)S->vals, nnz, f);
        RBX_HIP(hipGetLastError());
        int h = 1;
        RBX_HIP(hipMemcpyAsync(&h, f, sizeof(h), hipMemcpyDeviceToHost, st));
        RBX_HIP(hipStreamSynchronize(st));
        unit = h == 0;
    }
    const size_t lay = csr_lds_apply_ws_bytes(d, m, nnz, sizeof(T));
    const size_t bR = align256((size_t)(d + 1) * 8), bC = align256((size_t)nnz * 4);
    const size_t bV = align256((size_t)nnz * sizeof(T));
    const size_t total = bR + bC + bV + lay;
    if (total > PREP_MAX_RATIO * (size_t)nnz * (sizeof(T) + 16)) return;
    int64_t *row_ptr;
    int32_t *ccol;
    T *cval;
    if (!csr_build_counting<T>(S->rows, S->cols, (const T *)S->vals, nnz, 0, 0, d, m, &row_ptr, &ccol, &cval, st))
        return;
    int dev = -1;
    RBX_HIP(hipGetDevice(&dev));
    char *mem = nullptr;
    if (hipMalloc((void **)&mem, total) != hipSuccess) {
        (void)hipGetLastError();
        return;  // no copy: every lskges sorts for itself
    }
    SparsePrep *p = new SparsePrep();
    p->magic = PREP_MAGIC;
    p->dev = dev;
    p->val_size = sizeof(T);
    p->d = d;
    p->m = m;
    p->nnz = nnz;
    p->vals = S->vals;
    p->rows = S->rows;
    p->cols = S->cols;
    p->mem = mem;
    p->row_ptr = (int64_t *)mem;
    p->col = (int32_t *)(mem + bR);
    p->val = mem + bR + bC;
    p->lay = mem + bR + bC + bV;
    p->lay_bytes = lay;
    p->unit = unit;
    S->prep = p;  // owned by S from here, so prep_free releases it on any failure below
    RBX_HIP(hipMemcpyAsync(p->row_ptr, row_ptr, (size_t)(d + 1) * 8, hipMemcpyDeviceToDevice, st));
    RBX_HIP(hipMemcpyAsync(p->col, ccol, (size_t)nnz * 4, hipMemcpyDeviceToDevice, st));
    RBX_HIP(hipMemcpyAsync(p->val, cval, (size_t)nnz * sizeof(T), hipMemcpyDeviceToDevice, st));
    if (!csr_lds_layout<T>(p->row_ptr, p->col, (const T *)p->val, d, m, nnz, p->lay, lay, st, unit)) prep_free(S);
}

template <typename T>
static void lskges_impl(char layout, char opS, char opA, int64_t d, int64_t n, int64_t m, T alpha,
                        const rbx_sparse_skop *S, int64_t ro_s, int64_t co_s, const T *A, int64_t lda, T beta, T *B,
                        int64_t ldb, hipStream_t st) {
    SkopView<T> v = view_or_sample<T>(S, st);
    const SparsePrep *prep = S->nnz >= 0 ? usable_prep<T>(S, opS, d, m, ro_s, co_s) : nullptr;
    left_spmm_coo<T>(layout, opS, opA, d, n, m, alpha, S->dist.n_rows, S->dist.n_cols, v.nnz, v.vals, v.rows, v.cols,
                     ro_s, co_s, A, lda, beta, B, ldb, st, -1, nullptr, 0, prep, false, v.unit);
}

// rskges -> right_spmm -> left_spmm with flipped layout and op (spmm_dispatch.hh:181-219)
template <typename T>
static void rskges_impl(char layout, char opA, char opS, int64_t m, int64_t d, int64_t n, T alpha, const T *A,
                        int64_t lda, const rbx_sparse_skop *S, int64_t ro_s, int64_t co_s, T beta, T *B, int64_t ldb,
                        hipStream_t st) {
    check_layout(layout);
    check_op(opS);
    SkopView<T> v = view_or_sample<T>(S, st);
    char trans_opS = opS == 'N' ? 'T' : 'N';
    char trans_layout = layout == 'C' ? 'R' : 'C';
    left_spmm_coo<T>(trans_layout, trans_opS, opA, d, m, n, alpha, S->dist.n_rows, S->dist.n_cols, v.nnz, v.vals,
                     v.rows, v.cols, ro_s, co_s, A, lda, beta, B, ldb, st, -1, nullptr, 0, nullptr, false, v.unit);
}

// ---------------------------------------------------------------- CSR data x dense operator
// left_spmm for CSR, opA = NoTrans (csr_spmm_impl.hh:51-155).  One wave per
// output row; lanes walk output columns.
//   jik: C[i,j] = RN(beta*C) + RN(alpha * sum_l RN(val_l * B[col_l, j]))
//   ikb: C[i,j] = RN(beta*C), then C += RN(RN(alpha*val_l) * B[col_l, j])
template <typename T>
__global__ __launch_bounds__(256) void csr_rows_kernel(const int64_t *rowptr, const int64_t *colidx, const T *vals,
                                                       int64_t mrows, int64_t ncols_out, T alpha, T beta, const T *B,
                                                       int64_t brs, int64_t bcs, T *C, int64_t crs, int64_t ccs,
                                                       int ikb, int64_t lrow) {
    const int lane = threadIdx.x & 63;
    const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= mrows) return;
    const int64_t p0 = rowptr[i], p1 = rowptr[i + 1];
    if (p1 - p0 > lrow) return;  // a long row: summed in segments elsewhere
    for (int64_t j = lane; j < ncols_out; j += 64) {
        T *c = C + i * crs + j * ccs;
        T acc = (beta == (T)0) ? (T)0 : beta * (*c);
        if (alpha != (T)0) {
            if (ikb) {
                for (int64_t p = p0; p < p1; ++p) acc = acc + (alpha * vals[p]) * B[colidx[p] * brs + j * bcs];
            } else {
                T diff = (T)0;
                for (int64_t p = p0; p < p1; ++p) diff = diff + vals[p] * B[colidx[p] * brs + j * bcs];
                acc = acc + alpha * diff;
            }
        }
        *c = acc;
    }
}

// rows[p] = i for p in [rowptr[i], rowptr[i+1]); positions outside [0, nnz)
// are never written
__global__ __launch_bounds__(256) void csr_to_coo_rows_kernel(const int64_t *rowptr, int64_t nrows, int64_t nnz,
                                                              int64_t *rows) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrows) return;
    const int64_t p0 = rowptr[i] < 0 ? 0 : rowptr[i], p1 = rowptr[i + 1] > nnz ? nnz : rowptr[i + 1];
    for (int64_t p = p0; p < p1; ++p) rows[p] = i;
}

// op(submat(S)) of a materialised operator (S.buff in S.layout) gathered into
// a row-major n_p x n_q matrix.
template <typename T>
__global__ __launch_bounds__(256) void gather_op_kernel(const T *src, int64_t rs, int64_t cs, int64_t n_p,
                                                        int64_t n_q, T *dst) {
    const int64_t total = n_p * n_q;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = idx / n_q, q = idx - p * n_q;
        dst[idx] = src[p * rs + q * cs];
    }
}

// G = op(submat(S)) as a row-major n_p x n_q matrix in library scratch: the
// operator is generated from its counters, or gathered from S.buff.
template <typename T>
static T *materialize_op(const rbx_dense_skop *S, char opS, int64_t ro_s, int64_t co_s, int64_t n_p, int64_t n_q,
                         hipStream_t st) {
    const size_t bytes = (size_t)safe_mul(n_p > 0 ? n_p : 1, n_q > 0 ? n_q : 1) * sizeof(T);
    T *G = (T *)scratch(SCRATCH_DENSE_OP, bytes, st);
    if (n_p == 0 || n_q == 0) return G;
    if (S->buff == nullptr) {
        GenMap gm = gen_map(*S, opS, ro_s, co_s, false);
        fill_generated<T>(gm, n_p, n_q, G, n_q, 1, st);
        return G;
    }
    // S.buff holds the full operator in S.layout; G[p, q] = S[ro_s + p, co_s + q]
    // (opS = 'N') or S[ro_s + q, co_s + p] (opS = 'T')
    if (S->layout != 'C' && S->layout != 'R') RBX_FAIL(RBX_INVALID_ARG, "DenseSkOp.layout must be 'C' or 'R'");
    const int64_t frs = S->layout == 'C' ? 1 : S->dist.n_cols, fcs = S->layout == 'C' ? S->dist.n_rows : 1;
    const T *base = (const T *)S->buff + ro_s * frs + co_s * fcs;
    const int64_t rs = opS == 'N' ? frs : fcs, cs = opS == 'N' ? fcs : frs;
    int64_t blocks = (n_p * n_q + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(gather_op_kernel<T>, dim3((unsigned)blocks), dim3(256), 0, st, base, rs, cs, n_p, n_q, G);
    RBX_HIP(hipGetLastError());
    return G;
}

// A sparse data matrix in one of the reference's three formats
// (sparse_data/csr_matrix.hh, csc_matrix.hh, coo_matrix.hh), zero-based.
struct SpData {
    char fmt;  // 'R' CSR, 'C' CSC, 'O' COO
    int64_t n_rows, n_cols, nnz;
    const void *vals;
    const int64_t *ptr;   // CSR rowptr / CSC colptr
    const int64_t *idx;   // CSR colidxs / CSC rowidxs
    const int64_t *rows;  // COO
    const int64_t *cols;  // COO
};

static SpData sp_of(const rbx_csr_matrix *A) {
    RBX_REQUIRE(A != nullptr);
    return SpData{'R', A->n_rows, A->n_cols, A->nnz, A->vals, A->rowptr, A->colidxs, nullptr, nullptr};
}
static SpData sp_of(const rbx_csc_matrix *A) {
    RBX_REQUIRE(A != nullptr);
    return SpData{'C', A->n_rows, A->n_cols, A->nnz, A->vals, A->colptr, A->rowidxs, nullptr, nullptr};
}
static SpData sp_of(const rbx_coo_matrix *A) {
    RBX_REQUIRE(A != nullptr);
    return SpData{'O', A->n_rows, A->n_cols, A->nnz, A->vals, nullptr, nullptr, A->rows, A->cols};
}

// Major index of every stored entry of a compressed matrix (CSR row / CSC
// column), expanded from its pointer array into library scratch.
static const int64_t *expand_ptr(const int64_t *ptr, int64_t nmaj, int64_t nnz, hipStream_t st) {
    int64_t *out = (int64_t *)scratch(SCRATCH_SPARSE_OP, (size_t)(nnz > 0 ? nnz : 1) * 8, st);
    if (nnz > 0 && nmaj > 0) {
        dim3 grid((unsigned)((nmaj + 255) / 256));
        hipLaunchKernelGGL(csr_to_coo_rows_kernel, grid, dim3(256), 0, st, ptr, nmaj, nnz, out);
        RBX_HIP(hipGetLastError());
    }
    return out;
}

// rsksp3 (sparse_data/sksp.hh:277-326) for a CSR, CSC or COO data matrix.
template <typename T>
static void rsksp3_impl(char layout, char opA, char opS, int64_t m, int64_t d, int64_t n, T alpha, const SpData &A,
                        int64_t ro_a, int64_t co_a, const rbx_dense_skop *S, int64_t ro_s, int64_t co_s, T beta, T *B,
                        int64_t ldb, hipStream_t st) {
    RBX_REQUIRE(S != nullptr);
    check_layout(layout);
    check_op(opA);
    check_op(opS);
    RBX_REQUIRE(m >= 0 && d >= 0 && n >= 0);
    RBX_REQUIRE(ro_a >= 0 && co_a >= 0 && A.nnz >= 0);
    DenseGeom geo = dense_geom(S->dist);
    safe_mul(geo.dim_minor, geo.P);
    int64_t rows_S = opS == 'N' ? n : d, cols_S = opS == 'N' ? d : n;
    int64_t rows_A = opA == 'N' ? m : n, cols_A = opA == 'N' ? n : m;
    RBX_REQUIRE(A.n_rows >= rows_A + ro_a);
    RBX_REQUIRE(A.n_cols >= cols_A + co_a);
    RBX_REQUIRE(S->dist.n_rows >= rows_S + ro_s);
    RBX_REQUIRE(S->dist.n_cols >= cols_S + co_s);
    if (layout == 'C') RBX_REQUIRE(ldb >= m);
    else RBX_REQUIRE(ldb >= d);
    // CSR and CSC take the whole matrix; COO accepts a submatrix
    // (spmm_dispatch.hh:95-103)
    if (A.fmt != 'O') {
        RBX_REQUIRE(A.n_rows == rows_A && A.n_cols == cols_A);
        RBX_REQUIRE(ro_a == 0 && co_a == 0);
    }
    if (m == 0 || d == 0) return;
    const int64_t crs = layout == 'C' ? 1 : ldb, ccs = layout == 'C' ? ldb : 1;
    // The reference materialises submat(S) in its natural layout and flips opS
    // when that differs from `layout` (sksp.hh:298-324); the product kernel it
    // then picks depends on that layout, which fixes the summation form.
    // (a materialised operator is used in its own layout, sksp.hh:318-321)
    const char s_layout = S->buff != nullptr ? S->layout : geo.natural;
    const char opS_ref = (s_layout != layout) ? (opS == 'N' ? 'T' : 'N') : opS;
    // row-major op(B) with a row-major C selects the rank-one-update kernels
    // (spmm_dispatch.hh:147-176)
    const int kib = (layout == 'R' && opS_ref == 'N') ? 1 : 0;
    // G = op(submat(S)), n x d, row-major in scratch
    T *G = materialize_op<T>(S, opS, ro_s, co_s, n, d, st);
    // Which kernel family the reference reaches (spmm_dispatch.hh:72-90,
    // 139-176): CSR with op(A) = N, or CSC with op(A) = T (its transpose is the
    // CSR matrix sharing its arrays), runs row by row in stored order; every
    // other case runs column by column, i.e. ascending column of op(A).
    const bool row_kernel = (A.fmt == 'R' && opA == 'N') || (A.fmt == 'C' && opA == 'T');
    if (row_kernel) {
        ProfScope prof("csr_rows_kernel", st);
        constexpr int VEC = VecOf<T>::n;
        // vector path: 16-byte aligned rows of G (d columns wide from a
        // 256-byte base); a row-major B takes vector stores when aligned too
        const int cvec = layout == 'R' && ldb % VEC == 0 && ((uintptr_t)B % 16) == 0;
        if (d % VEC == 0) {
            launch_rows_vec<T, int64_t>(layout == 'C', A.ptr, A.idx, (const T *)A.vals, m, A.nnz, d, alpha, beta,
                                        (const T *)G, d, B, ldb, kib ? CRV_KIB : CRV_JIK, cvec, st);
        } else {
            const int64_t lrow = launch_long_rows<T, int64_t>(A.ptr, A.idx, (const T *)A.vals, m, A.nnz, d, alpha,
                                                              beta, (const T *)G, d, B, crs, ccs, st);
            dim3 grid((unsigned)((m + 3) / 4));
            hipLaunchKernelGGL(csr_rows_kernel<T>, grid, dim3(256), 0, st, A.ptr, A.idx, (const T *)A.vals, m, d,
                               alpha, beta, (const T *)G, (int64_t)d, (int64_t)1, B, crs, ccs, kib, lrow);
        }
        RBX_HIP(hipGetLastError());
        return;
    }
    const int64_t *rows = A.rows, *cols = A.cols;
    if (A.fmt == 'R') {
        rows = expand_ptr(A.ptr, A.n_rows, A.nnz, st);
        cols = A.idx;
    } else if (A.fmt == 'C') {
        rows = A.idx;
        cols = expand_ptr(A.ptr, A.n_cols, A.nnz, st);
    }
    // G is row-major n x d: op(B) = G for a row-major C, G^T of the
    // column-major d x n buffer otherwise
    const char opG = layout == 'C' ? 'T' : 'N';
    // CSR viewed transposed and CSC (the cases that reach here): within each
    // row of the view, the stored order already ascends in the view's column
    left_spmm_coo<T>(layout, opA, opG, m, d, n, alpha, A.n_rows, A.n_cols, A.nnz, (const T *)A.vals, rows, cols, ro_a,
                     co_a, G, d, beta, B, ldb, st, kib, nullptr, 0, nullptr, A.fmt != 'O');
}

// lsksp3 (sparse_data/sksp.hh:132-180): B = alpha * op(submat(S)) * op(A) + beta * B.
// The reference reduces it through right_spmm (spmm_dispatch.hh:182-219) to
// left_spmm on the transposed problem with the opposite layout and op(A)
// flipped; rsksp3's path takes that same left_spmm call when given the
// flipped layout, the flipped op(A) and the flipped op(S) (which makes its
// operator-layout rule match sksp.hh:176-177).
template <typename T>
static void lsksp3_impl(char layout, char opS, char opA, int64_t d, int64_t n, int64_t m, T alpha,
                        const rbx_dense_skop *S, int64_t ro_s, int64_t co_s, const SpData &A, int64_t ro_a,
                        int64_t co_a, T beta, T *B, int64_t ldb, hipStream_t st) {
    check_layout(layout);
    check_op(opS);
    check_op(opA);
    const char tl = layout == 'C' ? 'R' : 'C';
    rsksp3_impl<T>(tl, opA == 'N' ? 'T' : 'N', opS == 'N' ? 'T' : 'N', n, d, m, alpha, A, ro_a, co_a, S, ro_s, co_s,
                   beta, B, ldb, st);
}

// Adjoint of rsksp3 (opA = 'N') with respect to the stored values of A, the
// sampled dense-dense product
//   dvals[k] = alpha * sum_j Bbar[i_k, j] * G[c_k, j],  G = op(submat(S)) (n x d)
// for every stored entry k = (i_k, c_k) of A.  A workgroup owns SD_RB rows of
// A; Bbar is staged through LDS in SD_DC-column chunks (coalesced for either
// layout), each wave walks the stored entries of its rows, and a row of G is
// one coalesced read.  Partial sums over the chunks accumulate in dvals in a
// fixed order.
constexpr int SD_RB = 32;
template <typename T>
struct SddmmChunk {
    static constexpr int DC = sizeof(T) == 8 ? 128 : 256;  // 32 x DC tile: 32 KB
};

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <typename T>
__global__ __launch_bounds__(256) void sddmm_csr_kernel(const int64_t *rowptr, const int64_t *colidx, int64_t m,
                                                        int64_t d, T alpha, const T *Bb, int64_t brs, int64_t bcs,
                                                        const T *G, T *dvals) {
    constexpr int DC = SddmmChunk<T>::DC, NQ = DC / 64;
    __shared__ T tile[SD_RB][DC + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t i0 = (int64_t)blockIdx.x * SD_RB;
    for (int64_t j0 = 0; j0 < d; j0 += DC) {
        const int64_t dc = d - j0 < DC ? d - j0 : DC;
        const bool last = j0 + DC >= d;
        __syncthreads();
        // Bbar[i0 .. i0+SD_RB, j0 .. j0+dc) -> tile, threads walking the
        // unit-stride direction
        for (int idx = tid; idx < SD_RB * DC; idx += 256) {
            int r, c;
            if (brs == 1) { r = idx % SD_RB; c = idx / SD_RB; }
            else { c = idx % DC; r = idx / DC; }
            const int64_t i = i0 + r;
            tile[r][c] = (i < m && c < dc) ? Bb[i * brs + (j0 + c) * bcs] : (T)0;
        }
        __syncthreads();
        for (int r = wave; r < SD_RB; r += 4) {
            const int64_t i = i0 + r;
            if (i >= m) break;
            T gv[NQ];
#pragma unroll
            for (int q = 0; q < NQ; ++q) gv[q] = tile[r][lane + 64 * q];
            const int64_t kb = rowptr[i], ke = rowptr[i + 1];
            for (int64_t k = kb; k < ke; ++k) {
                const T *grow = G + colidx[k] * d + j0;
                T p = (T)0;
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
                    const int j = lane + 64 * q;
                    if (j < dc) p = fma(gv[q], grow[j], p);
                }
                p = wave_sum(p);
                if (lane == 0) {
                    const T tot = j0 == 0 ? p : dvals[k] + p;
                    dvals[k] = last ? alpha * tot : tot;
                }
            }
        }
    }
}

// Vectorised form (d a multiple of the vector width): a group of 16 lanes
// owns one stored entry, so a wave works on four entries of a row at once;
// each lane gathers 16-byte pieces of op(S)'s row and the 16 partial dot
// products are summed with four butterfly steps.
template <typename T>
__global__ __launch_bounds__(256) void sddmm_vec_kernel(const int64_t *rowptr, const int64_t *colidx, int64_t m,
                                                        int64_t d, T alpha, const T *Bb, int64_t brs, int64_t bcs,
                                                        const T *G, T *dvals) {
    using V = typename VecOf<T>::type;
    constexpr int VEC = VecOf<T>::n, DC = SddmmChunk<T>::DC, GS = 16 * VEC, NQ = DC / GS;
    __shared__ T tile[SD_RB][DC + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int grp = lane >> 4, gl = lane & 15;
    const int64_t i0 = (int64_t)blockIdx.x * SD_RB;
    for (int64_t j0 = 0; j0 < d; j0 += DC) {
        const int64_t dc = d - j0 < DC ? d - j0 : DC;
        const bool last = j0 + DC >= d;
        __syncthreads();
        for (int idx = tid; idx < SD_RB * DC; idx += 256) {
            int r, c;
            if (brs == 1) { r = idx % SD_RB; c = idx / SD_RB; }
            else { c = idx % DC; r = idx / DC; }
            const int64_t i = i0 + r;
            tile[r][c] = (i < m && c < dc) ? Bb[i * brs + (j0 + c) * bcs] : (T)0;
        }
        __syncthreads();
        for (int r = wave; r < SD_RB; r += 4) {
            const int64_t i = i0 + r;
            if (i >= m) break;
            const int64_t kb = rowptr[i], ke = rowptr[i + 1];
            for (int64_t pb = kb; pb < ke; pb += 64) {
                // the batch's columns loaded once (lane e: entry pb + e);
                // two steps of four entries (one per 16-lane group) in flight,
                // each lane's NQ pieces of its entry's operator row loaded
                // together
                const int cnt = (int)(ke - pb < 64 ? ke - pb : 64);
                const int64_t myc = lane < cnt ? colidx[pb + lane] : 0;
                for (int q = 0; q < cnt; q += 8) {
                    V gv[2][NQ];
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        const int64_t c = __shfl(myc, q + 4 * h + grp);  // entries past cnt: column 0
                        const T *grow = G + c * d + j0;
#pragma unroll
                        for (int u = 0; u < NQ; ++u) {
                            const int cc = gl * VEC + u * GS;
                            if (cc < dc) gv[h][u] = *(const V *)(grow + cc);
                        }
                    }
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        T p = (T)0;
#pragma unroll
                        for (int u = 0; u < NQ; ++u) {
                            const int cc = gl * VEC + u * GS;
                            if (cc < dc)
#pragma unroll
                                for (int e = 0; e < VEC; ++e) p = fma(tile[r][cc + e], vget<T>(gv[h][u], e), p);
                        }
#pragma unroll
                        for (int o = 8; o >= 1; o >>= 1) p += __shfl_xor(p, o);
                        const int e = q + 4 * h + grp;
                        if (e < cnt && gl == 0) {
                            const int64_t k = pb + e;
                            const T tot = j0 == 0 ? p : dvals[k] + p;
                            dvals[k] = last ? alpha * tot : tot;
                        }
                    }
                }
            }
        }
    }
}

template <typename T>
static void rsksp3_adjoint_impl(char layout, char opS, int64_t m, int64_t d, int64_t n, T alpha,
                                const rbx_csr_matrix *A, const rbx_dense_skop *S, int64_t ro_s, int64_t co_s,
                                const T *Bbar, int64_t ldbb, T *dvals, hipStream_t st) {
    RBX_REQUIRE(A != nullptr && S != nullptr);
    check_layout(layout);
    check_op(opS);
    RBX_REQUIRE(m >= 0 && d >= 0 && n >= 0);
    DenseGeom geo = dense_geom(S->dist);
    safe_mul(geo.dim_minor, geo.P);
    const int64_t rows_S = opS == 'N' ? n : d, cols_S = opS == 'N' ? d : n;
    RBX_REQUIRE(A->n_rows == m && A->n_cols == n);
    RBX_REQUIRE(S->dist.n_rows >= rows_S + ro_s);
    RBX_REQUIRE(S->dist.n_cols >= cols_S + co_s);
    if (layout == 'C') RBX_REQUIRE(ldbb >= m);
    else RBX_REQUIRE(ldbb >= d);
    if (A->nnz == 0 || m == 0) return;
    if (d == 0 || n == 0 || alpha == (T)0) {
        scale_only<T>(dvals, A->nnz, 1, 1, A->nnz, (T)0, st);  // zeros
        return;
    }
    const T *G = materialize_op<T>(S, opS, ro_s, co_s, n, d, st);
    const int64_t brs = layout == 'C' ? 1 : ldbb, bcs = layout == 'C' ? ldbb : 1;
    const int64_t blocks = (m + SD_RB - 1) / SD_RB;
    RBX_REQUIRE(blocks < (1ll << 31));
    ProfScope prof("sddmm_csr_kernel", st);
    if (d % VecOf<T>::n == 0)
        hipLaunchKernelGGL(sddmm_vec_kernel<T>, dim3((unsigned)blocks), dim3(256), 0, st, A->rowptr, A->colidxs, m, d,
                           alpha, Bbar, brs, bcs, (const T *)G, dvals);
    else
        hipLaunchKernelGGL(sddmm_csr_kernel<T>, dim3((unsigned)blocks), dim3(256), 0, st, A->rowptr, A->colidxs, m, d,
                           alpha, Bbar, brs, bcs, (const T *)G, dvals);
    RBX_HIP(hipGetLastError());
}

template <typename T>
static void fill_sparse_skop(rbx_sparse_skop *S, hipStream_t st) {
    RBX_REQUIRE(S != nullptr);
    SparseGeom g = sparse_geom(S->dist);
    if (S->vals == nullptr && S->rows == nullptr && S->cols == nullptr) {
        void *v = nullptr, *r = nullptr, *c = nullptr;
        if (hipMalloc(&v, (size_t)g.full_nnz * sizeof(T)) != hipSuccess ||
            hipMalloc(&r, (size_t)g.full_nnz * 8) != hipSuccess || hipMalloc(&c, (size_t)g.full_nnz * 8) != hipSuccess) {
            if (v) (void)hipFree(v);
            if (r) (void)hipFree(r);
            if (c) (void)hipFree(c);
            RBX_FAIL(RBX_OUT_OF_MEMORY, "fill_sparse: hipMalloc failed");
        }
        S->vals = v;
        S->rows = (int64_t *)r;
        S->cols = (int64_t *)c;
        S->own_memory = 1;
    }
    RBX_REQUIRE(S->vals != nullptr && S->rows != nullptr && S->cols != nullptr);
    prep_free(S);
    S->nnz = fill_sparse_impl<T>(S->dist, (T *)S->vals, S->rows, S->cols, S->seed_state, nullptr, st);
    prepare_skop<T>(S, st);
}

// ---------------------------------------------------------------- allocation-free lskges
// The allocation-free entry points keep their workspace (the operator
// window's CSR) in this file-scope device array: one copy per GPU, resident
// from the library's load on that GPU until the process ends, shared by every
// call on that GPU.  Its host-side name is only a stand-in; the device address
// of each GPU's copy comes from hipGetSymbolAddress, once per device.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
constexpr int NOALLOC_MAX_DEVICES = 64;
static __device__ __align__(256) char noalloc_scratch[NOALLOC_SCRATCH_BYTES];

static char *noalloc_scratch_ptr() {
    static std::mutex mu;
    static char *ptrs[NOALLOC_MAX_DEVICES] = {};
    int dev = 0;
    RBX_HIP(hipGetDevice(&dev));
    RBX_REQUIRE(dev >= 0 && dev < NOALLOC_MAX_DEVICES);
    std::lock_guard<std::mutex> lk(mu);
    if (ptrs[dev] == nullptr) {
        void *p = nullptr;
        RBX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch)));
        RBX_REQUIRE(p != nullptr);
        ptrs[dev] = (char *)p;
    }
    return ptrs[dev];
}

// lskges without any device allocation: S must hold its sampled arrays, op(A)
// must have unit stride along the contraction, and the operator window must
// fit the counting-sort build and the LDS apply; the arithmetic is lskges's.
template <typename T>
static void lskges_noalloc_impl(char layout, char opS, char opA, int64_t d, int64_t n, int64_t m, T alpha,
                                const rbx_sparse_skop *S, int64_t ro_s, int64_t co_s, const T *A, int64_t lda, T beta,
                                T *B, int64_t ldb, hipStream_t st) {
    RBX_REQUIRE(S != nullptr);
    if (S->nnz < 0)
        RBX_FAIL(RBX_UNSUPPORTED, "lskges_noalloc: the SparseSkOp must hold its sampled arrays "
                 "(rbx_[sd]fill_sparse_skop first); sampling would need device memory");
    char *ws = noalloc_scratch_ptr();
    SkopView<T> v = view_or_sample<T>(S, st);
    left_spmm_coo<T>(layout, opS, opA, d, n, m, alpha, S->dist.n_rows, S->dist.n_cols, v.nnz, v.vals, v.rows, v.cols,
                     ro_s, co_s, A, lda, beta, B, ldb, st, -1, ws, NOALLOC_SCRATCH_BYTES,
                     usable_prep<T>(S, opS, d, m, ro_s, co_s));
}

}  // namespace rbx

using namespace rbx;

extern "C" {

rbx_status rbx_sparse_dist_props(const rbx_sparse_dist *dist, int64_t *dim_major, int64_t *dim_minor,
                                 int64_t *full_nnz, double *isometry_scale) {
    RBX_API_BEGIN
    RBX_REQUIRE(dist != nullptr);
    SparseGeom g = sparse_geom(*dist);
    if (dim_major) *dim_major = g.dim_major;
    if (dim_minor) *dim_minor = g.dim_minor;
    if (full_nnz) *full_nnz = g.full_nnz;
    // sparse::isometry_scale (sparse_skops.hh:108-114)
    if (isometry_scale)
        *isometry_scale = dist->major_axis == 'S'
                              ? pow((double)dist->vec_nnz, -0.5)
                              : sqrt((double)g.dim_major / ((double)dist->vec_nnz * (double)g.dim_minor));
    RBX_API_END
}

rbx_status rbx_sparse_skop_init(const rbx_sparse_dist *dist, const rbx_rng_state *seed, rbx_sparse_skop *out) {
    RBX_API_BEGIN
    RBX_REQUIRE(dist != nullptr && seed != nullptr && out != nullptr);
    sparse_geom(*dist);
    memset(out, 0, sizeof(*out));
    out->dist = *dist;
    out->seed_state = *seed;
    out->next_state = *seed;
    ctr_to_words(ctr_add(ctr_from_words(seed->counter), (uint64_t)sparse_incr(*dist)), out->next_state.counter);
    out->nnz = -1;
    RBX_API_END
}

rbx_status rbx_sfill_sparse_unpacked_nosub(const rbx_sparse_dist *D, int64_t *nnz, float *vals, int64_t *rows,
                                           int64_t *cols, const rbx_rng_state *seed, rbx_rng_state *next_state,
                                           void *stream) {
    RBX_API_BEGIN
    RBX_REQUIRE(D != nullptr && seed != nullptr && nnz != nullptr);
    *nnz = fill_sparse_impl<float>(*D, vals, rows, cols, *seed, next_state, (hipStream_t)stream);
    RBX_API_END
}

rbx_status rbx_dfill_sparse_unpacked_nosub(const rbx_sparse_dist *D, int64_t *nnz, double *vals, int64_t *rows,
                                           int64_t *cols, const rbx_rng_state *seed, rbx_rng_state *next_state,
                                           void *stream) {
    RBX_API_BEGIN
    RBX_REQUIRE(D != nullptr && seed != nullptr && nnz != nullptr);
    *nnz = fill_sparse_impl<double>(*D, vals, rows, cols, *seed, next_state, (hipStream_t)stream);
    RBX_API_END
}

rbx_status rbx_sfill_sparse_skop(rbx_sparse_skop *S, void *stream) {
    RBX_API_BEGIN
    fill_sparse_skop<float>(S, (hipStream_t)stream);
    RBX_API_END
}

rbx_status rbx_dfill_sparse_skop(rbx_sparse_skop *S, void *stream) {
    RBX_API_BEGIN
    fill_sparse_skop<double>(S, (hipStream_t)stream);
    RBX_API_END
}

rbx_status rbx_ssparse_skop_prepare(rbx_sparse_skop *S, void *stream) {
    RBX_API_BEGIN
    RBX_REQUIRE(S != nullptr);
    prepare_skop<float>(S, (hipStream_t)stream);
    RBX_API_END
}

rbx_status rbx_dsparse_skop_prepare(rbx_sparse_skop *S, void *stream) {
    RBX_API_BEGIN
    RBX_REQUIRE(S != nullptr);
    prepare_skop<double>(S, (hipStream_t)stream);
    RBX_API_END
}

rbx_status rbx_sparse_skop_free(rbx_sparse_skop *S) {
    RBX_API_BEGIN
    RBX_REQUIRE(S != nullptr);
    prep_free(S);
    if (S->own_memory) {
        if (S->vals) RBX_HIP(hipFree(S->vals));
        if (S->rows) RBX_HIP(hipFree(S->rows));
        if (S->cols) RBX_HIP(hipFree(S->cols));
    }
    S->vals = nullptr;
    S->rows = S->cols = nullptr;
    S->nnz = -1;
    S->own_memory = 0;
    RBX_API_END
}

rbx_status rbx_slskges(char layout, char opS, char opA, int64_t d, int64_t n, int64_t m, float alpha,
                       const rbx_sparse_skop *S, int64_t ro_s, int64_t co_s, const float *A, int64_t lda, float beta,
                       float *B, int64_t ldb, void *stream) {
    RBX_API_BEGIN
    lskges_impl<float>(layout, opS, opA, d, n, m, alpha, S, ro_s, co_s, A, lda, beta, B, ldb, (hipStream_t)stream);
    RBX_API_END
}

rbx_status rbx_dlskges(char layout, char opS, char opA, int64_t d, int64_t n, int64_t m, double alpha,
                       const rbx_sparse_skop *S, int64_t ro_s, int64_t co_s, const double *A, int64_t lda, double beta,
                       double *B, int64_t ldb, void *stream) {
    RBX_API_BEGIN
    lskges_impl<double>(layout, opS, opA, d, n, m, alpha, S, ro_s, co_s, A, lda, beta, B, ldb, (hipStream_t)stream);
    RBX_API_END
}

rbx_status rbx_slskges_noalloc(char layout, char opS, char opA, int64_t d, int64_t n, int64_t m, float alpha,
                               const rbx_sparse_skop *S, int64_t ro_s, int64_t co_s, const float *A, int64_t lda,
                               float beta, float *B, int64_t ldb, void *stream) {
    RBX_API_BEGIN
    lskges_noalloc_impl<float>(layout, opS, opA, d, n, m, alpha, S, ro_s, co_s, A, lda, beta, B, ldb,
                               (hipStream_t)stream);
    RBX_API_END
}

rbx_status rbx_dlskges_noalloc(char layout, char opS, char opA, int64_t d, int64_t n, int64_t m, double alpha,
                               const rbx_sparse_skop *S, int64_t ro_s, int64_t co_s, const double *A, int64_t lda,
                               double beta, double *B, int64_t ldb, void *stream) {
    RBX_API_BEGIN
    lskges_noalloc_impl<double>(layout, opS, opA, d, n, m, alpha, S, ro_s, co_s, A, lda, beta, B, ldb,
                                (hipStream_t)stream);
    RBX_API_END
}

size_t rbx_lskges_noalloc_scratch_bytes(void) { return NOALLOC_SCRATCH_BYTES; }

rbx_status rbx_srskges(char layout, char opA, char opS, int64_t m, int64_t d, int64_t n, float alpha, const float *A,
                       int64_t lda, const rbx_sparse_skop *S, int64_t ro_s, int64_t co_s, float beta, float *B,
                       int64_t ldb, void *stream) {
    RBX_API_BEGIN
    rskges_impl<float>(layout, opA, opS, m, d, n, alpha, A, lda, S, ro_s, co_s, beta, B, ldb, (hipStream_t)stream);
    RBX_API_END
}

rbx_status rbx_drskges(char layout, char opA, char opS, int64_t m, int64_t d, int64_t n, double alpha, const double *A,
                       int64_t lda, const rbx_sparse_skop *S, int64_t ro_s, int64_t co_s, double beta, double *B,
                       int64_t ldb, void *stream) {
    RBX_API_BEGIN
    rskges_impl<double>(layout, opA, opS, m, d, n, alpha, A, lda, S, ro_s, co_s, beta, B, ldb, (hipStream_t)stream);
    RBX_API_END
}

// rsksp3 / lsksp3 entry points for the three data-matrix formats
#define RBX_SKSP3_ENTRY(PFX, T, SFX, MAT)                                                                              \
    rbx_status rbx_##PFX##rsksp3##SFX(char layout, char opA, char opS, int64_t m, int64_t d, int64_t n, T alpha,     \
                                      const MAT *A, int64_t ro_a, int64_t co_a, const rbx_dense_skop *S,            \
                                      int64_t ro_s, int64_t co_s, T beta, T *B, int64_t ldb, void *stream) {        \
        RBX_API_BEGIN                                                                                               \
        rsksp3_impl<T>(layout, opA, opS, m, d, n, alpha, sp_of(A), ro_a, co_a, S, ro_s, co_s, beta, B, ldb,         \
                       (hipStream_t)stream);                                                                        \
        RBX_API_END                                                                                                 \
    }                                                                                                               \
    rbx_status rbx_##PFX##lsksp3##SFX(char layout, char opS, char opA, int64_t d, int64_t n, int64_t m, T alpha,     \
                                      const rbx_dense_skop *S, int64_t ro_s, int64_t co_s, const MAT *A,            \
                                      int64_t ro_a, int64_t co_a, T beta, T *B, int64_t ldb, void *stream) {        \
        RBX_API_BEGIN                                                                                               \
        lsksp3_impl<T>(layout, opS, opA, d, n, m, alpha, S, ro_s, co_s, sp_of(A), ro_a, co_a, beta, B, ldb,         \
                       (hipStream_t)stream);                                                                        \
        RBX_API_END                                                                                                 \
    }

RBX_SKSP3_ENTRY(s, float, , rbx_csr_matrix)
RBX_SKSP3_ENTRY(d, double, , rbx_csr_matrix)
RBX_SKSP3_ENTRY(s, float, _csc, rbx_csc_matrix)
RBX_SKSP3_ENTRY(d, double, _csc, rbx_csc_matrix)
RBX_SKSP3_ENTRY(s, float, _coo, rbx_coo_matrix)
RBX_SKSP3_ENTRY(d, double, _coo, rbx_coo_matrix)

rbx_status rbx_srsksp3_adjoint(char layout, char opS, int64_t m, int64_t d, int64_t n, float alpha,
                               const rbx_csr_matrix *A, const rbx_dense_skop *S, int64_t ro_s, int64_t co_s,
                               const float *Bbar, int64_t ldbb, float *dvals, void *stream) {
    RBX_API_BEGIN
    rsksp3_adjoint_impl<float>(layout, opS, m, d, n, alpha, A, S, ro_s, co_s, Bbar, ldbb, dvals, (hipStream_t)stream);
    RBX_API_END
}

rbx_status rbx_drsksp3_adjoint(char layout, char opS, int64_t m, int64_t d, int64_t n, double alpha,
                               const rbx_csr_matrix *A, const rbx_dense_skop *S, int64_t ro_s, int64_t co_s,
                               const double *Bbar, int64_t ldbb, double *dvals, void *stream) {
    RBX_API_BEGIN
    rsksp3_adjoint_impl<double>(layout, opS, m, d, n, alpha, A, S, ro_s, co_s, Bbar, ldbb, dvals, (hipStream_t)stream);
    RBX_API_END
}


}  // extern "C"
