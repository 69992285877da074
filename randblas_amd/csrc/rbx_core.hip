// rbx_core.hip -- error slot, RNG-state helpers, raw Philox streams and the
// materialising fill_dense kernel.
#include "rbx_internal.hh"
#include "rbx_fastgen.hh"

#include <math.h>
#include <string.h>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

namespace rbx {

static thread_local char g_err[1024];

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}
void clear_error() { g_err[0] = 0; }

int num_cus() {
    static const int cached = [] {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0)
            return cus;
        return 256;
    }();
    return cached;
}

// ---------------------------------------------------------------------------
// Per-stream workspaces.
//
// Every (device, stream) pair owns its own scratch slots and rocBLAS handle,
// so calls on different streams never share memory.  A host thread that
// touches a stream's workspace holds that workspace's (recursive) mutex until
// its outermost API call returns (ApiScope), so two host threads that enqueue
// work on one stream take turns and never see each other's pointers change.
// Growing a slot is stream-ordered: the old buffer is released with
// hipFreeAsync on the same stream (after the work already enqueued there that
// may still read it) and the new one comes from hipMallocAsync; no device-wide
// synchronisation, no hipFree while other streams run.
struct StreamWs {
    std::recursive_mutex mu;
    int dev = 0;
    hipStream_t stream = nullptr;
    void *ptr[SCRATCH_NSLOTS] = {};
    size_t size[SCRATCH_NSLOTS] = {};
    void *blas = nullptr;
    void (*blas_destroy)(void *) = nullptr;
    hipStream_t aux = nullptr;                  // helper stream for work that overlaps this stream's
    hipEvent_t ev[WS_EVENTS] = {};              // fork / join events between the two
};

static std::mutex g_ws_mutex;  // guards the map, not the workspaces
static std::map<std::pair<int, hipStream_t>, std::unique_ptr<StreamWs>> g_ws;
static thread_local int t_api_depth = 0;
static thread_local std::vector<StreamWs *> t_held;

ApiScope::ApiScope() { ++t_api_depth; }
ApiScope::~ApiScope() {
    if (--t_api_depth == 0) {
        for (StreamWs *w : t_held) w->mu.unlock();
        t_held.clear();
    }
}

static bool pools_supported(int dev) {
    static std::mutex mu;
    static std::map<int, bool> known;
    std::lock_guard<std::mutex> lk(mu);
    auto it = known.find(dev);
    if (it != known.end()) return it->second;
    int v = 0;
    bool ok = hipDeviceGetAttribute(&v, hipDeviceAttributeMemoryPoolsSupported, dev) == hipSuccess && v != 0;
    known[dev] = ok;
    return ok;
}

// Stream-ordered release / allocation (a stream synchronisation where the
// device has no memory pools).
static void ws_free(StreamWs &w, void *p) {
    if (p == nullptr) return;
    if (pools_supported(w.dev)) {
        RBX_HIP(hipFreeAsync(p, w.stream));
    } else {
        RBX_HIP(hipStreamSynchronize(w.stream));
        RBX_HIP(hipFree(p));
    }
}
static void *ws_alloc(StreamWs &w, size_t bytes) {
    void *p = nullptr;
    hipError_t e = pools_supported(w.dev) ? hipMallocAsync(&p, bytes, w.stream) : hipMalloc(&p, bytes);
    if (e != hipSuccess || p == nullptr) {
        (void)hipGetLastError();  // the failed allocation is reported here, not by the next call
        RBX_FAIL(RBX_OUT_OF_MEMORY, "scratch: allocation of %zu bytes failed", bytes);
    }
    return p;
}

StreamWs &stream_ws(hipStream_t stream) {
    if (t_api_depth == 0) RBX_FAIL(RBX_INVALID_ARG, "stream_ws: used outside an API call");
    int dev = 0;
    RBX_HIP(hipGetDevice(&dev));
    StreamWs *w;
    {
        std::lock_guard<std::mutex> lk(g_ws_mutex);
        auto &slot = g_ws[{dev, stream}];
        if (!slot) {
            slot.reset(new StreamWs);
            slot->dev = dev;
            slot->stream = stream;
        }
        w = slot.get();
    }
    for (StreamWs *h : t_held)
        if (h == w) return *w;
    w->mu.lock();
    t_held.push_back(w);
    return *w;
}

void *scratch(int slot, size_t bytes, hipStream_t stream) {
    StreamWs &w = stream_ws(stream);
    if (bytes == 0) bytes = 256;
    if (w.size[slot] < bytes) {
        void *old = w.ptr[slot];
        w.ptr[slot] = nullptr;
        w.size[slot] = 0;
        ws_free(w, old);
        w.ptr[slot] = ws_alloc(w, bytes);
        w.size[slot] = bytes;
    }
    return w.ptr[slot];
}

void *&stream_blas_handle(hipStream_t stream, void (**destroy)(void *)) {
    StreamWs &w = stream_ws(stream);
    *destroy = w.blas_destroy;
    return w.blas;
}
void set_stream_blas_destroy(hipStream_t stream, void (*destroy)(void *)) { stream_ws(stream).blas_destroy = destroy; }

hipStream_t ws_aux_stream(hipStream_t stream) {
    StreamWs &w = stream_ws(stream);
    if (w.aux == nullptr) RBX_HIP(hipStreamCreateWithFlags(&w.aux, hipStreamNonBlocking));
    return w.aux;
}

hipEvent_t ws_event(hipStream_t stream, int i) {
    StreamWs &w = stream_ws(stream);
    if (i < 0 || i >= WS_EVENTS) RBX_FAIL(RBX_INVALID_ARG, "ws_event: index %d out of range", i);
    if (w.ev[i] == nullptr) RBX_HIP(hipEventCreateWithFlags(&w.ev[i], hipEventDisableTiming));
    return w.ev[i];
}

static void release_ws(StreamWs &w) {
    for (int s = 0; s < SCRATCH_NSLOTS; ++s) {
        void *p = w.ptr[s];
        w.ptr[s] = nullptr;
        w.size[s] = 0;
        ws_free(w, p);
    }
    if (w.blas != nullptr && w.blas_destroy != nullptr) w.blas_destroy(w.blas);
    w.blas = nullptr;
    // the helper stream's work is always joined into `stream` before the call
    // that issued it returns, so it is finished once `stream` reaches here
    if (w.aux != nullptr) {
        RBX_HIP(hipStreamSynchronize(w.aux));
        RBX_HIP(hipStreamDestroy(w.aux));
        w.aux = nullptr;
    }
    for (int i = 0; i < WS_EVENTS; ++i)
        if (w.ev[i] != nullptr) {
            (void)hipEventDestroy(w.ev[i]);
            w.ev[i] = nullptr;
        }
}

// ---------------------------------------------------------------------------
// Kernel timing records (rbx_profile_*)
struct ProfRec {
    const char *name;
    hipEvent_t a, b;
};
static std::mutex g_prof_mutex;
static bool g_prof_on = false;
static std::vector<ProfRec> g_prof;

ProfScope::ProfScope(const char *n, hipStream_t s) : name(n), stream(s) {
    if (!g_prof_on) return;
    if (hipEventCreate(&start) != hipSuccess) { start = nullptr; return; }
    if (hipEventRecord(start, stream) != hipSuccess) { (void)hipEventDestroy(start); start = nullptr; }
}

ProfScope::~ProfScope() {
    if (start == nullptr) return;
    hipEvent_t stop;
    if (hipEventCreate(&stop) != hipSuccess) { (void)hipEventDestroy(start); return; }
    if (hipEventRecord(stop, stream) != hipSuccess) {
        (void)hipEventDestroy(start);
        (void)hipEventDestroy(stop);
        return;
    }
    std::lock_guard<std::mutex> lk(g_prof_mutex);
    g_prof.push_back({name, start, stop});
}

// ---------------------------------------------------------------------------
// Raw Philox words: one thread per counter, out[4*i + w].
__global__ __launch_bounds__(256) void philox_stream_kernel(Ctr128 base, uint32_t k0, uint32_t k1,
                                                            int64_t n, uint4 *out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) out[i] = philox4x32_10(ctr_add(base, (uint64_t)i), k0, k1);
}

// ---------------------------------------------------------------------------
// fill_dense_unpacked (dense_skops.hh:563-606) on the device.
//
// The submatrix is n_rows x n_cols at (ro_s, co_s); the generated matrix
// G[p,q] = S[ro_s+p, co_s+q] is written to buff[p*rs + q*cs].  One thread
// owns one Philox counter of the natural layout and writes the (up to) four
// entries it produces, so every entry is computed exactly once.
template <typename T>
__global__ __launch_bounds__(256) void fill_dense_kernel(GenMap g, int64_t n_p, int64_t n_q, T *buff,
                                                         int64_t rs, int64_t cs, int64_t cell_first,
                                                         int64_t n_cells_per_vec, int64_t n_vecs) {
    const int64_t total = n_cells_per_vec * n_vecs;
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t bstart = g.b_off;
    const int64_t n_b = g.b_along_rows ? n_p : n_q;
    for (; idx < total; idx += stride) {
        int64_t v = idx / n_cells_per_vec;          // offset along the minor (a) direction
        int64_t cell = cell_first + idx % n_cells_per_vec;
        float4 w = gen4(g.gauss != 0, ctr_add(g.ctr, (uint64_t)((g.a_off + v) * g.P + cell)), g.k0, g.k1);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            int64_t bb = 4 * cell + t - bstart;      // offset along the major (b) direction
            if (bb < 0 || bb >= n_b) continue;
            int64_t p = g.b_along_rows ? bb : v;
            int64_t q = g.b_along_rows ? v : bb;
            buff[p * rs + q * cs] = entry_value<T>(f4_get(w, t), g.gauss != 0);
        }
    }
}

// The same entries when the major (b) direction is unit-stride, every cell
// is whole (b_off and n_b multiples of 4) and rows start 16-byte aligned:
// a 2-D grid (cells x minor indices, no index division) and one 16-byte
// store per counter (FP32; two for FP64).
template <typename T>
__global__ __launch_bounds__(256) void fill_dense_vec_kernel(GenMap g, T *buff, int64_t stride_a, int64_t cells,
                                                             int64_t n_a) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cells) return;
    const bool gauss = g.gauss != 0;
    const int64_t cell = g.b_off / 4 + c;
    for (int64_t v = blockIdx.y; v < n_a; v += gridDim.y) {
        const Ctr128 ctr = ctr_add(g.ctr, (uint64_t)((g.a_off + v) * g.P + cell));
        T *dst = buff + v * stride_a + 4 * c;
        if constexpr (sizeof(T) == 2) {
            // bfloat16 entries: the fast generator (exact by construction)
            *(uint2 *)dst = gauss ? gen4_bf16<true>(ctr, g.k0, g.k1) : gen4_bf16<false>(ctr, g.k0, g.k1);
            continue;
        }
        const float4 w = gen4(gauss, ctr, g.k0, g.k1);
        if constexpr (sizeof(T) == 4) {
            *(float4 *)dst = make_float4(entry_value<float>(w.x, gauss), entry_value<float>(w.y, gauss),
                                         entry_value<float>(w.z, gauss), entry_value<float>(w.w, gauss));
        } else {
            *(double2 *)dst = make_double2(entry_value<double>(w.x, gauss), entry_value<double>(w.y, gauss));
            *(double2 *)(dst + 2) = make_double2(entry_value<double>(w.z, gauss), entry_value<double>(w.w, gauss));
        }
    }
}

// The same entries for an output whose unit stride runs along the minor
// (a) direction, where one thread per counter would write 4 strided values:
// a workgroup generates a tile of 64 minor indices x 16 counters (64 major
// indices) into LDS and writes it 64 minor-contiguous values at a time.
constexpr int FT_A = 64, FT_C = 16, FT_B = 4 * FT_C;
template <typename T>
__global__ __launch_bounds__(256) void fill_dense_tr_kernel(GenMap g, int64_t n_p, int64_t n_q, T *buff, int64_t rs,
                                                            int64_t cs, int64_t cell_first, int64_t n_cells,
                                                            int64_t n_a) {
    __shared__ T tile[FT_B][FT_A + 1];
    const int tid = threadIdx.x;
    const int64_t v0 = (int64_t)blockIdx.x * FT_A;            // first minor index of the tile
    const int64_t c0 = cell_first + (int64_t)blockIdx.y * FT_C;  // first counter cell of the tile
    const int64_t n_b = g.b_along_rows ? n_p : n_q;
    const int64_t bbase = 4 * c0 - g.b_off;                    // major index of the tile's first entry
#pragma unroll
    for (int it = 0; it < FT_A * FT_C / 256; ++it) {
        const int e = tid + 256 * it;
        const int va = e / FT_C, cl = e % FT_C;  // neighbouring threads: neighbouring counters
        const int64_t v = v0 + va, cell = c0 + cl;
        if (v < n_a && cell < cell_first + n_cells) {
            const Ctr128 ctr = ctr_add(g.ctr, (uint64_t)((g.a_off + v) * g.P + cell));
            if constexpr (sizeof(T) == 2) {
                // bfloat16 entries: the fast generator (exact by construction)
                const uint2 pk = g.gauss != 0 ? gen4_bf16<true>(ctr, g.k0, g.k1) : gen4_bf16<false>(ctr, g.k0, g.k1);
                tile[4 * cl][va] = (T)(pk.x & 0xffffu);
                tile[4 * cl + 1][va] = (T)(pk.x >> 16);
                tile[4 * cl + 2][va] = (T)(pk.y & 0xffffu);
                tile[4 * cl + 3][va] = (T)(pk.y >> 16);
            } else {
                const float4 w = gen4(g.gauss != 0, ctr, g.k0, g.k1);
#pragma unroll
                for (int t = 0; t < 4; ++t) tile[4 * cl + t][va] = entry_value<T>(f4_get(w, t), g.gauss != 0);
            }
        }
    }
    __syncthreads();
    for (int e = tid; e < FT_B * FT_A; e += 256) {
        const int bl = e / FT_A, va = e % FT_A;  // lanes walk the minor (unit-stride) direction
        const int64_t bb = bbase + bl, v = v0 + va;
        if (bb < 0 || bb >= n_b || v >= n_a) continue;
        const int64_t p = g.b_along_rows ? bb : v, q = g.b_along_rows ? v : bb;
        buff[p * rs + q * cs] = tile[bl][va];
    }
}

// Write the generated matrix G (n_p x n_q) to buff[p*rs + q*cs].
template <typename T>
void fill_generated(const GenMap &g, int64_t n_p, int64_t n_q, T *buff, int64_t rs, int64_t cs, hipStream_t stream) {
    if (n_p <= 0 || n_q <= 0) return;
    int64_t n_b = g.b_along_rows ? n_p : n_q;
    int64_t n_a = g.b_along_rows ? n_q : n_p;
    int64_t cell_first = g.b_off / 4;
    int64_t cells = (g.b_off + n_b - 1) / 4 - cell_first + 1;
    const int64_t stride_a = g.b_along_rows ? cs : rs, stride_b = g.b_along_rows ? rs : cs;
    ProfScope prof("fill_dense_kernel", stream);
    if (stride_b == 1 && g.b_off % 4 == 0 && n_b % 4 == 0 && stride_a % 4 == 0 && ((uintptr_t)buff & 15) == 0 &&
        (n_a == 1 || stride_a >= n_b)) {
        const int64_t gx = (cells + 255) / 256;
        const int64_t gy = n_a < 4096 ? n_a : 4096;
        RBX_REQUIRE(gx < (1ll << 31));
        hipLaunchKernelGGL(fill_dense_vec_kernel<T>, dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, stream, g, buff,
                           stride_a, cells, n_a);
    } else if (stride_a == 1 && stride_b != 1 && n_a > 1) {
        const int64_t gx = (n_a + FT_A - 1) / FT_A, gy = (cells + FT_C - 1) / FT_C;
        RBX_REQUIRE(gx < (1ll << 31) && gy <= 65535);
        hipLaunchKernelGGL(fill_dense_tr_kernel<T>, dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, stream, g, n_p,
                           n_q, buff, rs, cs, cell_first, cells, n_a);
    } else {
        int64_t total = safe_mul(cells, n_a);
        int64_t blocks = (total + 255) / 256;
        int64_t cap = (int64_t)num_cus() * 64;
        if (blocks > cap) blocks = cap;
        hipLaunchKernelGGL(fill_dense_kernel<T>, dim3((unsigned)blocks), dim3(256), 0, stream, g, n_p, n_q, buff, rs,
                           cs, cell_first, cells, n_a);
    }
    RBX_HIP(hipGetLastError());
}
template void fill_generated<float>(const GenMap &, int64_t, int64_t, float *, int64_t, int64_t, hipStream_t);
template void fill_generated<double>(const GenMap &, int64_t, int64_t, double *, int64_t, int64_t, hipStream_t);
template void fill_generated<uint16_t>(const GenMap &, int64_t, int64_t, uint16_t *, int64_t, int64_t, hipStream_t);

template <typename T>
static void fill_dense_impl(char layout, const rbx_dense_dist *D, int64_t n_rows, int64_t n_cols,
                            int64_t ro_s, int64_t co_s, T *buff, const rbx_rng_state *seed,
                            rbx_rng_state *next_state, hipStream_t stream) {
    RBX_REQUIRE(D != nullptr && seed != nullptr);
    check_layout(layout);
    DenseGeom geo = dense_geom(*D);
    RBX_REQUIRE(n_rows > 0 && n_cols > 0);
    RBX_REQUIRE(ro_s >= 0 && co_s >= 0);
    RBX_REQUIRE(D->n_rows >= n_rows + ro_s);
    RBX_REQUIRE(D->n_cols >= n_cols + co_s);
    RBX_REQUIRE(buff != nullptr);
    safe_mul(geo.dim_minor, geo.P);
    safe_mul(n_rows, n_cols);
    rbx_dense_skop S;
    S.dist = *D;
    S.seed_state = *seed;
    GenMap g = gen_map(S, 'N', ro_s, co_s, false);
    // natural-orientation extents
    int64_t n_b = g.b_along_rows ? n_rows : n_cols;
    int64_t n_a = g.b_along_rows ? n_cols : n_rows;
    int64_t cell_first = g.b_off / 4;
    int64_t cell_last = (g.b_off + n_b - 1) / 4;
    int64_t cells = cell_last - cell_first + 1;
    int64_t rs = layout == 'C' ? 1 : n_cols, cs = layout == 'C' ? n_rows : 1;
    fill_generated<T>(g, n_rows, n_cols, buff, rs, cs, stream);
    if (next_state) {
        // fill_dense_submat_impl returns seed + ctr_mat_start + n_srows*stride
        // (dense_skops.hh:109-169): ctr_mat_start = a_off*P + b_off/4.
        Ctr128 c = ctr_add(g.ctr, (uint64_t)(g.a_off * g.P + cell_first));
        c = ctr_add(c, (uint64_t)safe_mul(n_a, g.P));
        next_state->key[0] = seed->key[0];
        next_state->key[1] = seed->key[1];
        ctr_to_words(c, next_state->counter);
    }
}

template <typename T>
static void fill_dense_skop_impl(rbx_dense_skop *S, hipStream_t stream) {
    RBX_REQUIRE(S != nullptr);
    DenseGeom geo = dense_geom(S->dist);
    if (S->buff == nullptr) {
        size_t bytes = (size_t)safe_mul(safe_mul(S->dist.n_rows, S->dist.n_cols), sizeof(T));
        void *p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) RBX_FAIL(RBX_OUT_OF_MEMORY, "fill_dense: hipMalloc of %zu bytes failed", bytes);
        S->buff = p;
        S->own_memory = 1;
    }
    if (S->layout != 'C' && S->layout != 'R') S->layout = geo.natural;
    fill_dense_impl<T>(S->layout, &S->dist, S->dist.n_rows, S->dist.n_cols, 0, 0, (T *)S->buff,
                       &S->seed_state, nullptr, stream);
}

}  // namespace rbx

using namespace rbx;

extern "C" {

const char *rbx_last_error_message(void) { return g_err; }

const char *rbx_version(void) { return "randblas_amd 0.1 (gfx950)"; }

rbx_status rbx_release_workspace(void *stream) {
    RBX_API_BEGIN
    int dev = 0;
    RBX_HIP(hipGetDevice(&dev));
    bool known;
    {
        std::lock_guard<std::mutex> lk(g_ws_mutex);
        known = g_ws.count({dev, (hipStream_t)stream}) != 0;
    }
    if (known) release_ws(stream_ws((hipStream_t)stream));
    RBX_API_END
}

rbx_status rbx_release_all_workspaces(void) {
    RBX_API_BEGIN
    int cur = 0;
    RBX_HIP(hipGetDevice(&cur));
    std::vector<StreamWs *> all;
    {
        std::lock_guard<std::mutex> lk(g_ws_mutex);
        for (auto &kv : g_ws) all.push_back(kv.second.get());
    }
    for (StreamWs *w : all) {
        std::lock_guard<std::recursive_mutex> lk(w->mu);
        RBX_HIP(hipSetDevice(w->dev));
        RBX_HIP(hipDeviceSynchronize());
        for (int s = 0; s < SCRATCH_NSLOTS; ++s) {
            if (w->ptr[s] != nullptr) RBX_HIP(hipFree(w->ptr[s]));
            w->ptr[s] = nullptr;
            w->size[s] = 0;
        }
        if (w->blas != nullptr && w->blas_destroy != nullptr) w->blas_destroy(w->blas);
        w->blas = nullptr;
        if (w->aux != nullptr) {
            RBX_HIP(hipStreamDestroy(w->aux));
            w->aux = nullptr;
        }
        for (int i = 0; i < WS_EVENTS; ++i)
            if (w->ev[i] != nullptr) {
                (void)hipEventDestroy(w->ev[i]);
                w->ev[i] = nullptr;
            }
    }
    RBX_HIP(hipSetDevice(cur));
    RBX_API_END
}

rbx_status rbx_profile_enable(int on) {
    RBX_API_BEGIN
    g_prof_on = on != 0;
    RBX_API_END
}

int rbx_profile_collect(char *names, int name_len, int *counts, double *total_ms, int max_entries) {
    clear_error();
    std::vector<ProfRec> recs;
    {
        std::lock_guard<std::mutex> lk(g_prof_mutex);
        recs.swap(g_prof);
    }
    std::vector<std::string> nm;
    std::vector<int> cnt;
    std::vector<double> tot;
    int rc = 0;
    for (auto &r : recs) {
        float ms = 0.f;
        if (hipEventSynchronize(r.b) != hipSuccess || hipEventElapsedTime(&ms, r.a, r.b) != hipSuccess) rc = -1;
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
        size_t i = 0;
        while (i < nm.size() && nm[i] != r.name) ++i;
        if (i == nm.size()) { nm.push_back(r.name); cnt.push_back(0); tot.push_back(0.0); }
        cnt[i] += 1;
        tot[i] += ms;
    }
    if (rc < 0) { set_error("rbx_profile_collect: event query failed"); return -1; }
    int n = 0;
    for (size_t i = 0; i < nm.size() && n < max_entries; ++i, ++n) {
        if (names && name_len > 0) snprintf(names + (size_t)n * name_len, name_len, "%s", nm[i].c_str());
        if (counts) counts[n] = cnt[i];
        if (total_ms) total_ms[n] = tot[i];
    }
    return n;
}

rbx_status rbx_rng_state_from_key(uint64_t key, rbx_rng_state *out) {
    RBX_API_BEGIN
    RBX_REQUIRE(out != nullptr);
    memset(out, 0, sizeof(*out));
    out->key[0] = (uint32_t)key;
    out->key[1] = (uint32_t)(key >> 32);
    RBX_API_END
}

rbx_status rbx_rng_state_incr(rbx_rng_state *state, uint64_t n) {
    RBX_API_BEGIN
    RBX_REQUIRE(state != nullptr);
    ctr_to_words(ctr_add(ctr_from_words(state->counter), n), state->counter);
    RBX_API_END
}

rbx_status rbx_philox4x32_10(const rbx_rng_state *state, int64_t n_counters, uint32_t *out, void *stream) {
    RBX_API_BEGIN
    RBX_REQUIRE(state != nullptr);
    RBX_REQUIRE(n_counters >= 0);
    if (n_counters == 0) return RBX_OK;
    RBX_REQUIRE(out != nullptr);
    int64_t blocks = (n_counters + 255) / 256;
    int64_t cap = (int64_t)num_cus() * 64;
    if (blocks > cap) blocks = cap;
    ProfScope prof("philox_stream_kernel", (hipStream_t)stream);
    hipLaunchKernelGGL(philox_stream_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                       ctr_from_words(state->counter), state->key[0], state->key[1], n_counters,
                       (uint4 *)out);
    RBX_HIP(hipGetLastError());
    RBX_API_END
}

rbx_status rbx_dense_dist_props(const rbx_dense_dist *dist, int64_t *dim_major, int64_t *dim_minor,
                                char *natural_layout, double *isometry_scale) {
    RBX_API_BEGIN
    RBX_REQUIRE(dist != nullptr);
    DenseGeom g = dense_geom(*dist);
    if (dim_major) *dim_major = g.dim_major;
    if (dim_minor) *dim_minor = g.dim_minor;
    if (natural_layout) *natural_layout = g.natural;
    if (isometry_scale) *isometry_scale = pow((double)g.dim_minor, -0.5);
    RBX_API_END
}

rbx_status rbx_dense_skop_init(const rbx_dense_dist *dist, const rbx_rng_state *seed, rbx_dense_skop *out) {
    RBX_API_BEGIN
    RBX_REQUIRE(dist != nullptr && seed != nullptr && out != nullptr);
    DenseGeom g = dense_geom(*dist);
    memset(out, 0, sizeof(*out));
    out->dist = *dist;
    out->seed_state = *seed;
    out->next_state = *seed;
    // dense::compute_next_state (dense_skops.hh:172-185)
    ctr_to_words(ctr_add(ctr_from_words(seed->counter), (uint64_t)safe_mul(g.P, g.dim_minor)),
                 out->next_state.counter);
    out->buff = nullptr;
    out->layout = g.natural;
    out->own_memory = 0;
    RBX_API_END
}

rbx_status rbx_sfill_dense_unpacked(char layout, const rbx_dense_dist *D, int64_t n_rows, int64_t n_cols,
                                    int64_t ro_s, int64_t co_s, float *buff, const rbx_rng_state *seed,
                                    rbx_rng_state *next_state, void *stream) {
    RBX_API_BEGIN
    fill_dense_impl<float>(layout, D, n_rows, n_cols, ro_s, co_s, buff, seed, next_state, (hipStream_t)stream);
    RBX_API_END
}

rbx_status rbx_dfill_dense_unpacked(char layout, const rbx_dense_dist *D, int64_t n_rows, int64_t n_cols,
                                    int64_t ro_s, int64_t co_s, double *buff, const rbx_rng_state *seed,
                                    rbx_rng_state *next_state, void *stream) {
    RBX_API_BEGIN
    fill_dense_impl<double>(layout, D, n_rows, n_cols, ro_s, co_s, buff, seed, next_state, (hipStream_t)stream);
    RBX_API_END
}

rbx_status rbx_sfill_dense_skop(rbx_dense_skop *S, void *stream) {
    RBX_API_BEGIN
    fill_dense_skop_impl<float>(S, (hipStream_t)stream);
    RBX_API_END
}

rbx_status rbx_dfill_dense_skop(rbx_dense_skop *S, void *stream) {
    RBX_API_BEGIN
    fill_dense_skop_impl<double>(S, (hipStream_t)stream);
    RBX_API_END
}

rbx_status rbx_dense_skop_free(rbx_dense_skop *S) {
    RBX_API_BEGIN
    RBX_REQUIRE(S != nullptr);
    if (S->own_memory && S->buff != nullptr) {
        RBX_HIP(hipFree(S->buff));
    }
    S->buff = nullptr;
    S->own_memory = 0;
    RBX_API_END
}

}  // extern "C"
