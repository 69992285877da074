// rbx_internal.hh -- host-side plumbing shared by the library's translation
// units: error reporting, argument validation, operator geometry.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdarg.h>
#include <string>

#include "../../include/randblas_amd.h"
#include "rb_philox.hh"

namespace rbx {

// Thread-local error slot behind rbx_last_error_message().
void set_error(const char *fmt, ...);
void clear_error();

struct Fail {
    rbx_status st;
};

// randblas_require equivalent (exceptions.hh:152-162): throws inside the
// library, converted to a status at the C ABI boundary by RBX_API_BEGIN/END.
#define RBX_REQUIRE(cond)                                                              \
    do {                                                                               \
        if (!(cond)) {                                                                 \
            ::rbx::set_error("%s: requirement failed: %s", __func__, #cond);           \
            throw ::rbx::Fail{RBX_INVALID_ARG};                                        \
        }                                                                              \
    } while (0)

#define RBX_FAIL(status, ...)                                                          \
    do {                                                                               \
        ::rbx::set_error(__VA_ARGS__);                                                 \
        throw ::rbx::Fail{status};                                                     \
    } while (0)

#define RBX_HIP(call)                                                                  \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            ::rbx::set_error("%s: %s failed: %s", __func__, #call, hipGetErrorString(e_)); \
            throw ::rbx::Fail{RBX_HIP_ERROR};                                          \
        }                                                                              \
    } while (0)

// Marks one API call on this thread; when the outermost call returns, the
// per-stream workspaces it used are released to other host threads.
struct ApiScope {
    ApiScope();
    ~ApiScope();
    ApiScope(const ApiScope &) = delete;
    ApiScope &operator=(const ApiScope &) = delete;
};

#define RBX_API_BEGIN                                                                  \
    ::rbx::ApiScope rbx_api_scope_;                                                    \
    ::rbx::clear_error();                                                              \
    try {
#define RBX_API_END                                                                    \
    }                                                                                  \
    catch (const ::rbx::Fail &f) {                                                     \
        return f.st;                                                                   \
    }                                                                                  \
    catch (...) {                                                                      \
        ::rbx::set_error("%s: unexpected C++ exception", __func__);                    \
        return RBX_INVALID_ARG;                                                        \
    }                                                                                  \
    return RBX_OK;

// safe_int_product (base.hh:260-274)
inline int64_t safe_mul(int64_t a, int64_t b) {
    if (a == 0 || b == 0) return 0;
    int64_t c;
    if (__builtin_mul_overflow(a, b, &c)) {
        RBX_FAIL(RBX_OVERFLOW, "Overflow when multiplying a (=%lld) and b(=%lld)", (long long)a,
                 (long long)b);
    }
    return c;
}

inline void check_layout(char layout) {
    if (layout != 'C' && layout != 'R') RBX_FAIL(RBX_INVALID_ARG, "invalid layout '%c'", layout);
}
inline void check_op(char op) {
    if (op != 'N' && op != 'T') RBX_FAIL(RBX_INVALID_ARG, "invalid op '%c'", op);
}

// DenseDist geometry (dense_skops.hh:313-329, 187-199)
struct DenseGeom {
    int64_t dim_major, dim_minor, P;  // P = counters per major vector
    char natural;                     // 'R' or 'C'
};
inline DenseGeom dense_geom(const rbx_dense_dist &D) {
    if (D.n_rows <= 0 || D.n_cols <= 0) RBX_FAIL(RBX_INVALID_ARG, "DenseDist: n_rows and n_cols must be positive");
    if (D.family != 'G' && D.family != 'U') RBX_FAIL(RBX_INVALID_ARG, "DenseDist: family must be 'G' or 'U'");
    if (D.major_axis != 'L' && D.major_axis != 'S') RBX_FAIL(RBX_INVALID_ARG, "DenseDist: major_axis must be 'L' or 'S'");
    DenseGeom g;
    bool lng = D.major_axis == 'L';
    int64_t mx = D.n_rows > D.n_cols ? D.n_rows : D.n_cols;
    int64_t mn = D.n_rows < D.n_cols ? D.n_rows : D.n_cols;
    g.dim_major = lng ? mx : mn;
    g.dim_minor = lng ? mn : mx;
    g.P = (g.dim_major + 3) / 4;
    bool wide = D.n_rows < D.n_cols;
    if (wide && lng) g.natural = 'R';
    else if (wide) g.natural = 'C';
    else if (lng) g.natural = 'C';
    else g.natural = 'R';
    return g;
}

// Mapping of a generated matrix G (p = row, q = column of G) onto the
// natural-layout coordinates (a = minor index, b = major index) of the
// operator: entry (a, b) is word b%4 of Philox(counter + a*P + b/4).
struct GenMap {
    Ctr128 ctr;
    uint32_t k0, k1;
    int64_t P;
    int64_t a_off, b_off;  // natural coordinates of G(0,0)
    int b_along_rows;      // 1: b = b_off + p, a = a_off + q; 0: b = b_off + q, a = a_off + p
    int gauss;
};

// G = op(submat(S)) where submat starts at (ro_s, co_s); `transposed` applies
// an extra transpose (G = op(submat(S))^T).
inline GenMap gen_map(const rbx_dense_skop &S, char opS, int64_t ro_s, int64_t co_s, bool transposed) {
    DenseGeom g = dense_geom(S.dist);
    GenMap m;
    m.ctr = ctr_from_words(S.seed_state.counter);
    m.k0 = S.seed_state.key[0];
    m.k1 = S.seed_state.key[1];
    m.P = g.P;
    m.gauss = S.dist.family == 'G';
    // op(submat(S))[p,q] = S[ro_s+p, co_s+q] (N) or S[ro_s+q, co_s+p] (T)
    bool swap_rc = (opS == 'T') != transposed;  // G[p,q] = S[ro+q, co+p]
    // S[r, c] -> natural (a, b) = (r, c) if RowMajor natural else (c, r)
    if (g.natural == 'R') {
        m.a_off = ro_s;
        m.b_off = co_s;
        m.b_along_rows = swap_rc ? 1 : 0;
    } else {
        m.a_off = co_s;
        m.b_off = ro_s;
        m.b_along_rows = swap_rc ? 0 : 1;
    }
    return m;
}

int num_cus();

// Library-owned device scratch: grow-only buffers per (device, stream, slot).
// Calls on different streams never share a buffer; host threads that enqueue
// on the same stream take turns (the workspace stays locked until the
// thread's outermost API call returns).  A slot's contents live until the
// next call on the same stream that uses it; one call uses each slot for one
// purpose (SCRATCH_DENSE_OP: an operator gathered or generated for the
// sparse-data products; SCRATCH_SPARSE_G: the row-major copy of the dense
// operand inside left_spmm_coo; SCRATCH_SPLITK: the K-range partial products
// of the fused dense sketch; SCRATCH_SOLVE / SCRATCH_SOLVE2: the sketched
// system of a sketch-and-solve call and its adjoint's vectors;
// SCRATCH_LONG_ROWS: the segment sums of the sparse row kernels' long rows).
// Growth is stream-ordered (hipFreeAsync / hipMallocAsync on `stream`).
enum ScratchSlot { SCRATCH_BF16 = 0, SCRATCH_SPARSE_OP = 1, SCRATCH_SPARSE_CSR = 2, SCRATCH_DENSE_OP = 3,
                   SCRATCH_SMALL = 4, SCRATCH_SPARSE_G = 5, SCRATCH_SPLITK = 6, SCRATCH_SOLVE = 7,
                   SCRATCH_SOLVE2 = 8, SCRATCH_LONG_ROWS = 9, SCRATCH_NSLOTS = 10 };
void *scratch(int slot, size_t bytes, hipStream_t stream);
// A helper stream owned by `stream`'s workspace (non-blocking, default
// priority) and up to WS_EVENTS events, for calls that overlap part of their
// own work with the rest; such a call joins all helper work back into
// `stream` before it returns.  Released with the workspace.
constexpr int WS_EVENTS = 8;
hipStream_t ws_aux_stream(hipStream_t stream);
hipEvent_t ws_event(hipStream_t stream, int i);
// The rocBLAS handle slot of the stream's workspace (created by its user)
// and its destroy function, called by rbx_release_workspace.
void *&stream_blas_handle(hipStream_t stream, void (**destroy)(void *));
void set_stream_blas_destroy(hipStream_t stream, void (*destroy)(void *));

// Write the generated matrix G (n_p x n_q) of a GenMap to buff[p*rs + q*cs].
template <typename T>
void fill_generated(const GenMap &g, int64_t n_p, int64_t n_q, T *buff, int64_t rs, int64_t cs, hipStream_t stream);

// Event-bracketed launch record used by rbx_profile_* (a no-op unless enabled).
struct ProfScope {
    ProfScope(const char *name, hipStream_t stream);
    ~ProfScope();
    const char *name;
    hipStream_t stream;
    hipEvent_t start = nullptr;
};

}  // namespace rbx
