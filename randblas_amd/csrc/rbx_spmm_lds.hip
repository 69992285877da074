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

// H[b * d + r] becomes the number of row-r entries in blocks before b, and
// tot[r] the row's total.  32 rows per workgroup, each row's blocks cut into
// 8 segments (one thread each): segment sums, their exclusive offsets through
// LDS, then each segment's running offsets written back.
constexpr int CSC_ROWS = 32, CSC_SEGS = 8;
__global__ __launch_bounds__(CSC_ROWS * CSC_SEGS) void csrb_colscan_kernel(int32_t *H, int nb, int d, int64_t *tot) {
    __shared__ int32_t seg[CSC_SEGS][CSC_ROWS];
    const int t = threadIdx.x, lr = t % CSC_ROWS, g = t / CSC_ROWS;
    const int r = blockIdx.x * CSC_ROWS + lr;
    const int nbs = (nb + CSC_SEGS - 1) / CSC_SEGS;
    const int b0 = g * nbs, b1 = b0 + nbs < nb ? b0 + nbs : nb;
    int32_t sum = 0;
    if (r < d) {
#pragma unroll 8
        for (int b = b0; b < b1; ++b) sum += H[(int64_t)b * d + r];
    }
    seg[g][lr] = sum;
    __syncthreads();
    int32_t run = 0;
    for (int q = 0; q < g; ++q) run += seg[q][lr];
    if (r < d) {
#pragma unroll 8
        for (int b = b0; b < b1; ++b) {
            const int64_t i = (int64_t)b * d + r;
            const int32_t x = H[i];
            H[i] = run;
            run += x;
        }
        if (g == CSC_SEGS - 1) tot[r] = run;
    }
}

// One workgroup: row_ptr[0..d) holds the row totals on entry and their
// exclusive scan on exit; row_ptr[d] = the sum.
__global__ __launch_bounds__(1024) void csrb_rowscan_kernel(int d, int64_t *row_ptr) {
    __shared__ int64_t part[1024];
    const int t = threadIdx.x;
    const int per = (d + 1023) / 1024;  // rows per thread (<= 8)
    int64_t v[CSRB_MAX_ROWS / 1024];
    int64_t s = 0;
#pragma unroll
    for (int q = 0; q < CSRB_MAX_ROWS / 1024; ++q) {
        const int r = t * per + q;
        v[q] = (q < per && r < d) ? row_ptr[r] : 0;
        s += v[q];
    }
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int64_t x = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += x;
        __syncthreads();
    }
    int64_t run = t > 0 ? part[t - 1] : 0;
#pragma unroll
    for (int q = 0; q < CSRB_MAX_ROWS / 1024; ++q) {
        const int r = t * per + q;
        if (q < per && r < d) {
            row_ptr[r] = run;
            run += v[q];
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

// Block layout of the counting sort: tiles of CB_TILE entries, at most
// CB_MAX_BLOCKS workgroups, each owning `span` consecutive stored entries.
struct CsrbPlan {
    int nb;
    int64_t span;
    size_t bH, bR, bK, bC, bV;
    size_t bytes() const { return bH + bR + 2 * bK + bC + bV; }
};

static CsrbPlan csrb_plan(int64_t nnz, int64_t d, size_t val_size) {
    CsrbPlan p;
    const int64_t tiles = (nnz + CB_TILE - 1) / CB_TILE;
    const int64_t tpb = (tiles + CB_MAX_BLOCKS - 1) / CB_MAX_BLOCKS;  // tiles per block
    p.nb = (int)((tiles + tpb - 1) / tpb);
    p.span = tpb * CB_TILE;
    p.bH = al256((size_t)p.nb * d * 4);
    p.bR = al256((size_t)(d + 1) * 8);
    p.bK = al256((size_t)nnz * 8);
    p.bC = al256((size_t)nnz * 4);
    p.bV = al256((size_t)nnz * val_size);
    return p;
}

static bool csrb_supported(int64_t nnz, int64_t d, int64_t m) {
    return d > 0 && d <= CSRB_MAX_ROWS && nnz > 0 && nnz < (1ll << 31) && m < (1ll << 31);
}

size_t csr_build_counting_bytes(int64_t nnz, int64_t d, size_t val_size) {
    if (d <= 0 || d > CSRB_MAX_ROWS || nnz <= 0 || nnz >= (1ll << 31)) return 0;
    return csrb_plan(nnz, d, val_size).bytes();
}

template <typename T>
bool csr_build_counting_ws(const int64_t *rows, const int64_t *cols, const T *vals, int64_t nnz, int64_t ro,
                           int64_t co, int64_t d, int64_t m, char *ws, size_t ws_bytes, int64_t **row_ptr_out,
                           int32_t **col_out, T **val_out, hipStream_t st) {
    if (!csrb_supported(nnz, d, m)) return false;
    const CsrbPlan pl = csrb_plan(nnz, d, sizeof(T));
    RBX_REQUIRE(ws != nullptr && ws_bytes >= pl.bytes());
    const int nb = pl.nb;
    const int64_t span = pl.span;
    int32_t *H = (int32_t *)ws;
    int64_t *row_ptr = (int64_t *)(ws + pl.bH);
    uint64_t *keys = (uint64_t *)(ws + pl.bH + pl.bR), *keys2 = (uint64_t *)(ws + pl.bH + pl.bR + pl.bK);
    int32_t *col = (int32_t *)(ws + pl.bH + pl.bR + 2 * pl.bK);
    T *val = (T *)(ws + pl.bH + pl.bR + 2 * pl.bK + pl.bC);
    const int di = (int)d;
    hipLaunchKernelGGL(csrb_hist_kernel, dim3(nb), dim3(CB_THR), 0, st, rows, cols, nnz, ro, co, di, m, span, H);
    RBX_HIP(hipGetLastError());
    hipLaunchKernelGGL(csrb_colscan_kernel, dim3((unsigned)((d + CSC_ROWS - 1) / CSC_ROWS)),
                       dim3(CSC_ROWS * CSC_SEGS), 0, st, H, nb, di, row_ptr);
    RBX_HIP(hipGetLastError());
    hipLaunchKernelGGL(csrb_rowscan_kernel, dim3(1), dim3(1024), 0, st, di, row_ptr);
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

template <typename T>
bool csr_build_counting(const int64_t *rows, const int64_t *cols, const T *vals, int64_t nnz, int64_t ro, int64_t co,
                        int64_t d, int64_t m, int64_t **row_ptr_out, int32_t **col_out, T **val_out, hipStream_t st) {
    if (!csrb_supported(nnz, d, m)) return false;
    const size_t bytes = csrb_plan(nnz, d, sizeof(T)).bytes();
    char *ws = (char *)scratch(SCRATCH_SPARSE_CSR, bytes, st);
    return csr_build_counting_ws<T>(rows, cols, vals, nnz, ro, co, d, m, ws, bytes, row_ptr_out, col_out, val_out,
                                    st);
}

// ---------------------------------------------------------------- apply through LDS
// A workgroup owns NT output rows (one per thread) and J output columns, and
// streams those J columns of the dense operand through LDS in CK-deep chunks
// (double-buffered: the next chunk's loads are in flight while the current one
// is applied).  The stored entries are first laid out by chunk and slot:
// slot q of row r in chunk k is the row's q-th entry among those with a
// column in chunk k (its value at Ev[(k * MAXE + q) * d + r], its column
// within the chunk in 16 bits, two slots per 32-bit word of Ep), and
// startT[k * d + r] where the row's entries of chunk k start in the CSR.  A
// wave's load of a slot word then reads 64 consecutive rows
// (one coalesced request instead of 64 lines), and each thread loads the next
// chunk's slots into a second register bank while it applies the current
// chunk's; slots past the wave's largest count are skipped, and a row with
// more than MAXE entries in a chunk reads the rest from the CSR.  Every
// output is one chain over its row's entries in ascending column.  The row
// groups of one column block run back to back on one XCD, so the block's
// dense data comes from HBM once and from that XCD's L2 after.
// Deepest chunk and slots per row and chunk, by element size: FP32 tiles take
// twice the depth in the same LDS, so rows average up to ten entries per
// chunk there (six in FP64), which evens out the wave's per-chunk maximum.
constexpr int la_ckmax(size_t val_size) { return val_size == 4 ? 2048 : 1024; }
constexpr int la_maxe(size_t val_size) { return val_size == 4 ? 20 : 12; }
// the 16-bit slots of unit layouts: two per word, the column within a chunk below 2^15
static_assert(la_maxe(4) % 2 == 0 && la_maxe(8) % 2 == 0 && la_ckmax(4) <= 0x8000 && la_ckmax(8) <= 0x8000,
              "16-bit slot packing");

// Chunk depth: a power of two in [64, la_ckmax] giving rows about six (FP64)
// or ten (FP32) stored entries per chunk or fewer.
static int la_chunk(int64_t d, int64_t m, int64_t nnz, size_t val_size) {
    const double per_chunk = val_size == 4 ? 10.0 : 6.0;
    int ck = 64;
    while (ck < la_ckmax(val_size) && (double)(2 * ck) * (double)nnz <= per_chunk * (double)d * (double)m) ck *= 2;
    return ck;
}

struct LaPlan {
    int ck;
    int64_t nk;
    size_t bStart, bEv, bEp;
    size_t bytes() const { return bStart + bEv + bEp; }
};
static LaPlan la_plan(int64_t d, int64_t m, int64_t nnz, size_t val_size) {
    LaPlan p;
    p.ck = la_chunk(d, m, nnz, val_size);
    p.nk = (m + p.ck - 1) / p.ck;
    p.bStart = al256((size_t)(p.nk + 1) * (size_t)d * 4);
    p.bEv = al256((size_t)p.nk * la_maxe(val_size) * (size_t)d * val_size);
    p.bEp = al256((size_t)p.nk * la_maxe(val_size) * (size_t)d * 2);
    return p;
}

size_t csr_lds_apply_ws_bytes(int64_t d, int64_t m, int64_t nnz, size_t val_size) {
    return la_plan(d, m, nnz, val_size).bytes();
}

// startT[k * d + r] = position of row r's first entry with column >= k*CK
// (the row's end for k = nk)
__global__ __launch_bounds__(256) void csr_chunk_table_kernel(const int64_t *row_ptr, const int32_t *col, int64_t d,
                                                              int nk, int CK, int32_t *startT) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= d * (int64_t)(nk + 1)) return;
    const int k = (int)(idx / d);
    const int64_t r = idx - (int64_t)k * d;
    int64_t lo = row_ptr[r], hi = row_ptr[r + 1];
    if (k < nk) {
        const int64_t c = (int64_t)k * CK;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if ((int64_t)col[mid] < c) lo = mid + 1;
            else hi = mid;
        }
    } else {
        lo = hi;
    }
    startT[idx] = (int32_t)lo;
}

// One workgroup per row (grid-stride): each entry p of row r goes to slot
// q = p - startT[k * d + r] of its chunk k when q < la_maxe(sizeof(T)): its
// value to Ev[(k * MAXE + q) * d + r], its column within the chunk (below
// 2048) to 16 bits, two slots of a row per 32-bit word (Ep[(k * MAXE / 2 +
// q / 2) * d + r], slot q in half q % 2).  With `unit` (every value is +1 or
// -1) bit 15 holds the value's sign, so the applies that know this read one
// word per two slots and no values.
template <typename T>
__global__ __launch_bounds__(256) void csr_chunk_slots_kernel(const int64_t *row_ptr, const int32_t *col,
                                                              const T *val, int64_t d, int CK,
                                                              const int32_t *startT, T *Ev, uint16_t *Ep,
                                                              int unit) {
    for (int64_t r = blockIdx.x; r < d; r += gridDim.x) {
        const int64_t p1 = row_ptr[r + 1];
        for (int64_t p = row_ptr[r] + threadIdx.x; p < p1; p += blockDim.x) {
            const int32_t c = col[p];
            const int64_t k = c / CK;
            const int64_t q = p - startT[k * d + r];
            constexpr int MAXE = la_maxe(sizeof(T));
            if (q < MAXE) {
                const int64_t o = (k * MAXE + q) * d + r;
                const T v = val[p];
                Ev[o] = v;
                Ep[2 * ((k * (MAXE / 2) + q / 2) * d + r) + (q & 1)] =
                    (uint16_t)((c - k * CK) | ((unit && v < (T)0) ? 0x8000 : 0));
            }
        }
    }
}

// Raw buffer loads: an offset at or past the descriptor's byte count reads 0,
// so tile loads past the operand's end need no branch around them.
constexpr int BUF_DWORD3 = 0x00020000;  // gfx9-family raw buffer descriptor word 3
template <typename T>
__device__ __forceinline__ T buf_load(__amdgpu_buffer_rsrc_t r, uint32_t off);
template <>
__device__ __forceinline__ double buf_load<double>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}
template <>
__device__ __forceinline__ float buf_load<float>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}

// A1: alpha == 1, so the products by alpha (exact then) are left out.
// Slots come from the 16-bit words Ep, two per load (column within the
// chunk).  UNIT: every value is +1 or -1 (a layout built with `unit`), its
// sign in bit 15, and no values are loaded; otherwise the values come from
// Ev and bit 15 is masked off (a unit layout serves both).
template <typename T, int J, int NT, int FORM, bool A1, bool UNIT>
__global__ __launch_bounds__(NT) void csr_lds_apply_kernel(const int32_t *startT, const T *Ev, const uint32_t *Ep,
                                                           int nk, int CK, const int32_t *col, const T *val,
                                                           int64_t d, int64_t n, int64_t m, T alpha, T beta,
                                                           const T *B, int64_t ldb, T *C, int64_t crs, int64_t ccs,
                                                           int RW, int ncb) {
    // Chunk tiles.  Rows of 16 bytes or more (J * sizeof(T) >= 16) are stored
    // row by row, As[buf][c * J + ...], in 16-byte pieces whose order within
    // the row is XOR-swizzled by c, so a stored entry reads its J values with
    // J * sizeof(T) / 16 wide reads and the reads of 64 lanes at unrelated c
    // spread over all bank groups (stores of consecutive c stay conflict-free).
    // Narrower rows are stored column by column, As[buf][jj * LA_CKMAX + c].
    constexpr int LA_CKMAX = la_ckmax(sizeof(T)), LA_MAXE = la_maxe(sizeof(T));
    __shared__ __attribute__((aligned(16))) T As[2][LA_CKMAX * J];
    constexpr bool ROWMAJ = J * (int)sizeof(T) >= 16;
    constexpr int EPP = 16 / (int)sizeof(T);                        // elements per 16-byte piece
    constexpr int NPC = ROWMAJ ? J / EPP : 1;                       // pieces per row
    constexpr int RSH = ROWMAJ ? (J * (int)sizeof(T) == 64 ? 1 : J * (int)sizeof(T) == 32 ? 2 : 3) : 0;
    typedef T piece_t __attribute__((ext_vector_type(16 / sizeof(T))));
    auto slot_of = [](int c, int t) { return t ^ ((c >> RSH) & (NPC - 1)); };
    constexpr int PER = LA_CKMAX / NT;  // contraction indices a thread loads per chunk, at most
    const int tid = threadIdx.x;
    // XCD-aware order over a grid padded to a multiple of 8: tile tt runs on
    // XCD tt / (grid / 8), so a column block's row groups share one L2
    const int W = ncb * RW, L = blockIdx.x, G8 = (int)gridDim.x / 8;
    const int tt = (L % 8) * G8 + L / 8;
    if (tt >= W) return;  // padding: the whole workgroup leaves before any barrier
    const int cb = tt / RW, rg = tt % RW;
    const int64_t j0 = (int64_t)cb * J;
    const int64_t r = (int64_t)rg * NT + tid;
    const bool active = r < d;
    const int64_t rs = active ? r : d - 1;  // a row every load may read (loads are never conditional)

    T acc[J];
#pragma unroll
    for (int jj = 0; jj < J; ++jj) {
        acc[jj] = (T)0;
        if (active && beta != (T)0 && j0 + jj < n) acc[jj] = beta * C[r * crs + (j0 + jj) * ccs];
    }
    auto wave_max = [](int32_t x) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) x = max(x, __shfl_xor(x, o));
        return x;
    };
    auto start_of = [&](int64_t k) {
        const int32_t v = startT[k * d + rs];
        return active ? v : 0;
    };

    // register banks of slots: chunk k in bank k % 2, two 16-bit slots per
    // word of bc
    uint32_t bc0[LA_MAXE / 2], bc1[LA_MAXE / 2];
    T bv0[LA_MAXE], bv1[LA_MAXE];
    // every slot is loaded (a fixed number of loads, so the waits before the
    // tile stores cover only the tile's own, older loads)
    auto load_bank = [&](int64_t k, uint32_t *bc, T *bv) {
#pragma unroll
        for (int q = 0; q < LA_MAXE / 2; ++q) bc[q] = Ep[(k * (LA_MAXE / 2) + q) * d + rs];
        if constexpr (!UNIT) {
#pragma unroll
            for (int q = 0; q < LA_MAXE; ++q) bv[q] = Ev[(k * LA_MAXE + q) * d + rs];
        }
    };
    auto fma_one = [&](int jj, T b, T v) {
        if (FORM == SPF_KIB) acc[jj] = acc[jj] + (A1 ? v : alpha * v) * b;
        else acc[jj] = acc[jj] + v * (A1 ? b : alpha * b);
    };
    constexpr int LOGROWB = J * (int)sizeof(T) == 64 ? 6 : J * (int)sizeof(T) == 32 ? 5 : 4;
    // the entry at chunk-local column c with value v
    auto apply_one = [&](const T *Ab, int32_t c, T v) {
        if constexpr (ROWMAJ) {
            // piece t of row c sits at byte c * ROWB + 16 * (t ^ s(c)), which is
            // (c * ROWB | 16 * s(c)) ^ 16 * t: one XOR per piece
            const uint32_t a0 = ((uint32_t)c << LOGROWB) | ((((uint32_t)c >> RSH) & (NPC - 1)) << 4);
            const char *Ac = (const char *)Ab;
            piece_t pc[NPC];
#pragma unroll
            for (int t = 0; t < NPC; ++t) pc[t] = *(const piece_t *)(Ac + (a0 ^ (uint32_t)(t << 4)));
#pragma unroll
            for (int t = 0; t < NPC; ++t)
#pragma unroll
                for (int e = 0; e < EPP; ++e) fma_one(t * EPP + e, pc[t][e], v);
        } else {
#pragma unroll
            for (int jj = 0; jj < J; ++jj) fma_one(jj, Ab[jj * LA_CKMAX + c], v);
        }
    };
    auto apply = [&](int64_t k, const uint32_t *bc, const T *bv, int32_t s, int32_t e, int32_t wmax) {
        const int32_t cbase = (int32_t)(k * CK);
        const T *Ab = As[k & 1];
        const int32_t cnt = e - s;
#pragma unroll
        for (int q = 0; q < LA_MAXE; ++q)
            if (q < wmax && q < cnt) {
                const uint32_t e = (bc[q >> 1] >> (16 * (q & 1))) & 0xffffu;
                if constexpr (UNIT) apply_one(Ab, (int32_t)(e & 0x7fffu), (e & 0x8000u) ? (T)-1 : (T)1);
                else apply_one(Ab, (int32_t)(e & 0x7fffu), bv[q]);
            }
        for (int32_t p = s + LA_MAXE; p < e; ++p) apply_one(Ab, col[p] - cbase, val[p]);
    };

    // one buffer descriptor per output column: m entries, none past n
    __amdgpu_buffer_rsrc_t rsc[J];
#pragma unroll
    for (int jj = 0; jj < J; ++jj) {
        const bool colok = j0 + jj < n;
        rsc[jj] = __builtin_amdgcn_make_buffer_rsrc((void *)(B + (colok ? (j0 + jj) * ldb : 0)), 0,
                                                    colok ? (int)(m * (int64_t)sizeof(T)) : 0, BUF_DWORD3);
    }
    T ld[PER][J];
    auto load_chunk = [&](int64_t k) {
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const uint32_t off = (uint32_t)((k * CK + u * NT + tid) * (int64_t)sizeof(T));
#pragma unroll
            for (int jj = 0; jj < J; ++jj) ld[u][jj] = buf_load<T>(rsc[jj], off);
        }
    };
    auto store_chunk = [&](int buf) {
#pragma unroll
        for (int u = 0; u < PER; ++u)
            if (u * NT + tid < CK) {
                const int c = u * NT + tid;
                if constexpr (ROWMAJ) {
#pragma unroll
                    for (int t = 0; t < NPC; ++t) {
                        piece_t pc;
#pragma unroll
                        for (int e = 0; e < EPP; ++e) pc[e] = ld[u][t * EPP + e];
                        *(piece_t *)(&As[buf][c * J + slot_of(c, t) * EPP]) = pc;
                    }
                } else {
#pragma unroll
                    for (int jj = 0; jj < J; ++jj) As[buf][jj * LA_CKMAX + c] = ld[u][jj];
                }
            }
    };

    // chunk k's entries are [s_k, e_k) of the CSR, with e_k = s_{k+1}.  The
    // prefetches of chunk k + 1 are issued unconditionally (indices clamped
    // to the last chunk), so every iteration has the same loads in the same
    // order and the waits before each use cover only older loads.
    const int64_t klast = nk - 1;
    int32_t s0 = start_of(0), e0 = start_of(1), s1 = 0, e1 = 0;
    int32_t w0 = wave_max(e0 - s0), w1 = 0;
    int32_t en = start_of(nk >= 2 ? 2 : nk);  // e_{k+1} ahead of its use
    load_chunk(0);
    load_bank(0, bc0, bv0);
    store_chunk(0);
    __syncthreads();
    for (int64_t k = 0; k < nk; k += 2) {
        // chunk k from bank 0; chunk k + 1's tile, then its slots into bank 1
        {
            const int64_t kn = k + 1 < nk ? k + 1 : klast;
            s1 = e0;
            e1 = k + 1 < nk ? en : e0;
            w1 = wave_max(e1 - s1);
            en = start_of(k + 3 <= nk ? k + 3 : nk);  // first: the next wait for it is a short one
            load_chunk(kn);
            load_bank(kn, bc1, bv1);
        }
        apply(k, bc0, bv0, s0, e0, w0);
        store_chunk((int)((k + 1) & 1));
        __syncthreads();
        if (k + 1 >= nk) break;
        // chunk k + 1 from bank 1; chunk k + 2's tile, then its slots into bank 0
        {
            const int64_t kn = k + 2 < nk ? k + 2 : klast;
            s0 = e1;
            e0 = k + 2 < nk ? en : e1;
            w0 = wave_max(e0 - s0);
            en = start_of(k + 4 <= nk ? k + 4 : nk);
            load_chunk(kn);
            load_bank(kn, bc0, bv0);
        }
        apply(k + 1, bc1, bv1, s1, e1, w1);
        store_chunk((int)((k + 2) & 1));
        __syncthreads();
    }
    if (active) {
#pragma unroll
        for (int jj = 0; jj < J; ++jj)
            if (j0 + jj < n) C[r * crs + (j0 + jj) * ccs] = acc[jj];
    }
}

template <typename T, int J, int NT>
static void launch_lds_apply(int form, dim3 grid, const int32_t *startT, const T *Ev, const uint32_t *Ep, int nk,
                             int ck, const int32_t *col, const T *val, int64_t d, int64_t n, int64_t m, T alpha,
                             T beta, const T *B, int64_t ldb, T *C, int64_t crs, int64_t ccs, int RW, int ncb,
                             bool unit, hipStream_t st) {
#define RBX_LA_LAUNCH(F, A1, U)                                                                                   \
    hipLaunchKernelGGL((csr_lds_apply_kernel<T, J, NT, F, A1, U>), grid, dim3(NT), 0, st, startT, Ev, Ep, nk, ck, col, \
                       val, d, n, m, alpha, beta, B, ldb, C, crs, ccs, RW, ncb)
    const bool one = alpha == (T)1;
    // the value-free variant only for the widest column blocks (the others
    // read the values of a unit layout like any other)
    if (unit && J == 8) {
        if (form == SPF_KIB) {
            if (one) RBX_LA_LAUNCH(SPF_KIB, true, true);
            else RBX_LA_LAUNCH(SPF_KIB, false, true);
        } else {
            if (one) RBX_LA_LAUNCH(SPF_JKI, true, true);
            else RBX_LA_LAUNCH(SPF_JKI, false, true);
        }
    } else if (form == SPF_KIB) {
        if (one) RBX_LA_LAUNCH(SPF_KIB, true, false);
        else RBX_LA_LAUNCH(SPF_KIB, false, false);
    } else {
        if (one) RBX_LA_LAUNCH(SPF_JKI, true, false);
        else RBX_LA_LAUNCH(SPF_JKI, false, false);
    }
#undef RBX_LA_LAUNCH
    RBX_HIP(hipGetLastError());
}

template <typename T, int NT>
static void launch_lds_apply_j(int J, int form, dim3 grid, const int32_t *startT, const T *Ev, const uint32_t *Ep,
                               int nk, int ck, const int32_t *col, const T *val, int64_t d, int64_t n, int64_t m,
                               T alpha, T beta, const T *B, int64_t ldb, T *C, int64_t crs, int64_t ccs, int RW,
                               int ncb, bool unit, hipStream_t st) {
#define RBX_LA_ARGS \
    form, grid, startT, Ev, Ep, nk, ck, col, val, d, n, m, alpha, beta, B, ldb, C, crs, ccs, RW, ncb, unit, st
    if (J == 8) launch_lds_apply<T, 8, NT>(RBX_LA_ARGS);
    else if (J == 4) launch_lds_apply<T, 4, NT>(RBX_LA_ARGS);
    else if (J == 2) launch_lds_apply<T, 2, NT>(RBX_LA_ARGS);
    else launch_lds_apply<T, 1, NT>(RBX_LA_ARGS);
#undef RBX_LA_ARGS
}

static bool lds_apply_fits(int64_t d, int64_t m, size_t val_size) {
    // buffer offsets: (m + one chunk) entries must stay below 2^31 bytes
    return d > 0 && m > 0 && (m + la_ckmax(val_size)) * (int64_t)val_size < (1ll << 31);
}

template <typename T>
bool csr_lds_layout(const int64_t *row_ptr, const int32_t *col, const T *val, int64_t d, int64_t m, int64_t nnz,
                    char *ws, size_t ws_bytes, hipStream_t st, bool unit) {
    if (!lds_apply_fits(d, m, sizeof(T)) || ws == nullptr) return false;
    const LaPlan pl = la_plan(d, m, nnz, sizeof(T));
    RBX_REQUIRE(ws_bytes >= pl.bytes());
    if (pl.nk >= (1ll << 30)) return false;
    int32_t *startT = (int32_t *)ws;
    T *Ev = (T *)(ws + pl.bStart);
    uint16_t *Ep = (uint16_t *)(ws + pl.bStart + pl.bEv);
    const int64_t tcells = d * (pl.nk + 1);
    hipLaunchKernelGGL(csr_chunk_table_kernel, dim3((unsigned)((tcells + 255) / 256)), dim3(256), 0, st, row_ptr, col,
                       d, (int)pl.nk, pl.ck, startT);
    RBX_HIP(hipGetLastError());
    const int64_t cus = num_cus();
    const int64_t sg = d < 4 * cus ? d : 4 * cus;
    hipLaunchKernelGGL((csr_chunk_slots_kernel<T>), dim3((unsigned)sg), dim3(256), 0, st, row_ptr, col, val, d, pl.ck,
                       (const int32_t *)startT, Ev, Ep, (int)unit);
    RBX_HIP(hipGetLastError());
    return true;
}

template <typename T>
bool csr_lds_apply_laid(const int64_t *row_ptr, const int32_t *col, const T *val, int64_t d, int64_t n, int64_t m,
                        int64_t nnz, T alpha, T beta, const T *B, int64_t ldb, T *C, int64_t crs, int64_t ccs,
                        int form, const char *ws, size_t ws_bytes, hipStream_t st, bool unit) {
    (void)row_ptr;
    if (n <= 0 || !lds_apply_fits(d, m, sizeof(T)) || ws == nullptr) return false;
    const LaPlan pl = la_plan(d, m, nnz, sizeof(T));
    RBX_REQUIRE(ws_bytes >= pl.bytes());
    // 512 rows per workgroup (256 for few rows), the widest column block
    // (8, 4, 2, 1) that still gives about a workgroup per CU
    const int NT = d > 256 ? 512 : 256;
    const int64_t RW = (d + NT - 1) / NT;
    const int64_t cus = num_cus();
    int J = 8;
    while (J > 1 && ((n + J - 1) / J) * RW < cus * 3 / 4) J /= 2;
    const int64_t ncb = (n + J - 1) / J;
    if (ncb * RW >= (1ll << 30) || pl.nk >= (1ll << 30)) return false;
    const int32_t *startT = (const int32_t *)ws;
    const T *Ev = (const T *)(ws + pl.bStart);
    const uint32_t *Ep = (const uint32_t *)(ws + pl.bStart + pl.bEv);
    const int nki = (int)pl.nk;
    const dim3 grid((unsigned)((ncb * RW + 7) / 8 * 8));
    const int rw = (int)RW, nc = (int)ncb;
    if (NT == 512)
        launch_lds_apply_j<T, 512>(J, form, grid, startT, Ev, Ep, nki, pl.ck, col, val, d, n, m, alpha, beta, B, ldb, C,
                                   crs, ccs, rw, nc, unit, st);
    else
        launch_lds_apply_j<T, 256>(J, form, grid, startT, Ev, Ep, nki, pl.ck, col, val, d, n, m, alpha, beta, B, ldb, C,
                                   crs, ccs, rw, nc, unit, st);
    return true;
}

template <typename T>
bool csr_lds_apply(const int64_t *row_ptr, const int32_t *col, const T *val, int64_t d, int64_t n, int64_t m,
                   int64_t nnz, T alpha, T beta, const T *B, int64_t ldb, T *C, int64_t crs, int64_t ccs, int form,
                   char *ws, size_t ws_bytes, hipStream_t st, bool unit) {
    if (d <= 0 || n <= 0 || m <= 0 || ws == nullptr) return false;
    // every check of the apply before anything is launched
    const int NT = d > 256 ? 512 : 256;
    const int64_t RW = (d + NT - 1) / NT;
    if (n * RW >= (1ll << 30)) return false;
    if (!csr_lds_layout<T>(row_ptr, col, val, d, m, nnz, ws, ws_bytes, st, unit)) return false;
    return csr_lds_apply_laid<T>(row_ptr, col, val, d, n, m, nnz, alpha, beta, B, ldb, C, crs, ccs, form, ws, ws_bytes,
                                 st, unit);
}

template bool csr_build_counting_ws<float>(const int64_t *, const int64_t *, const float *, int64_t, int64_t,
                                           int64_t, int64_t, int64_t, char *, size_t, int64_t **, int32_t **,
                                           float **, hipStream_t);
template bool csr_build_counting_ws<double>(const int64_t *, const int64_t *, const double *, int64_t, int64_t,
                                            int64_t, int64_t, int64_t, char *, size_t, int64_t **, int32_t **,
                                            double **, hipStream_t);
template bool csr_build_counting<float>(const int64_t *, const int64_t *, const float *, int64_t, int64_t, int64_t,
                                        int64_t, int64_t, int64_t **, int32_t **, float **, hipStream_t);
template bool csr_build_counting<double>(const int64_t *, const int64_t *, const double *, int64_t, int64_t, int64_t,
                                         int64_t, int64_t, int64_t **, int32_t **, double **, hipStream_t);
template bool csr_lds_apply<float>(const int64_t *, const int32_t *, const float *, int64_t, int64_t, int64_t,
                                   int64_t, float, float, const float *, int64_t, float *, int64_t, int64_t, int,
                                   char *, size_t, hipStream_t, bool);
template bool csr_lds_layout<float>(const int64_t *, const int32_t *, const float *, int64_t, int64_t, int64_t,
                                    char *, size_t, hipStream_t, bool);
template bool csr_lds_layout<double>(const int64_t *, const int32_t *, const double *, int64_t, int64_t, int64_t,
                                     char *, size_t, hipStream_t, bool);
template bool csr_lds_apply_laid<float>(const int64_t *, const int32_t *, const float *, int64_t, int64_t, int64_t,
                                        int64_t, float, float, const float *, int64_t, float *, int64_t, int64_t, int,
                                        const char *, size_t, hipStream_t, bool);
template bool csr_lds_apply_laid<double>(const int64_t *, const int32_t *, const double *, int64_t, int64_t, int64_t,
                                         int64_t, double, double, const double *, int64_t, double *, int64_t, int64_t,
                                         int, const char *, size_t, hipStream_t, bool);
template bool csr_lds_apply<double>(const int64_t *, const int32_t *, const double *, int64_t, int64_t, int64_t,
                                    int64_t, double, double, const double *, int64_t, double *, int64_t, int64_t, int,
                                    char *, size_t, hipStream_t, bool);

}  // namespace rbx
