// RandBLAS.hh -- header-only C++ layer of randblas_amd with RandBLAS's own
// type and function names, forwarding to the C ABI in randblas_amd.h.
//
// Include this header instead of RandBLAS/RandBLAS.hh (RandBLAS.hh:33-41) and
// link librandblas_amd.so.  Differences a caller sees:
//   * every matrix / array pointer is a DEVICE pointer (hipMalloc'd memory);
//   * calls are asynchronous on the stream set by RandBLAS::gpu::set_stream
//     (default: the null stream), except where a host value is returned
//     (fill_sparse's nnz, require_symmetric);
//   * operators whose buff is nullptr are never materialised by the sketch
//     functions: their entries are regenerated inside the product kernels;
//   * fill_dense(S) / fill_sparse(S) allocate device memory (hipMalloc) that
//     the operator's destructor frees;
//   * SparseSkOp and CSRMatrix use int64_t indices (the only index type the
//     reference's sketch_general accepts for sparse operators, skge.hh:784).
// Errors are thrown as RandBLAS::Error (std::overflow_error for overflow),
// as the reference does (exceptions.hh:57-95, base.hh:260-274).
#ifndef RANDBLAS_AMD_CPP_HH
#define RANDBLAS_AMD_CPP_HH

#include <cmath>
#include <cstdint>
#include <cstring>
#include <exception>
#include <limits>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>

#include "randblas_amd.h"

// BLAS++ (blas.hh) defines these enums with the same values; they are
// declared here only when BLAS++ has not been included first.
#ifndef BLAS_HH
namespace blas {
enum class Layout : char { ColMajor = 'C', RowMajor = 'R' };
enum class Op : char { NoTrans = 'N', Trans = 'T', ConjTrans = 'C' };
enum class Side : char { Left = 'L', Right = 'R' };
}  // namespace blas
#endif

namespace RandBLAS {

// ---------------------------------------------------------------- errors
class Error : public std::exception {
  public:
    explicit Error(std::string msg) : msg_(std::move(msg)) {}
    const char *what() const noexcept override { return msg_.c_str(); }

  private:
    std::string msg_;
};

class Unsupported : public Error {
  public:
    using Error::Error;
};

namespace gpu {
// Stream used by every call of this layer on the calling thread (a
// hipStream_t passed as void*; nullptr = the null stream).
inline void *&current_stream() {
    static thread_local void *s = nullptr;
    return s;
}
inline void set_stream(void *stream) { current_stream() = stream; }
inline void *stream() { return current_stream(); }
}  // namespace gpu

namespace detail {
inline void check(rbx_status s, const char *func) {
    if (s == RBX_OK) return;
    std::string msg = std::string(func) + ": " + rbx_last_error_message();
    if (s == RBX_OVERFLOW) throw std::overflow_error(msg);
    if (s == RBX_UNSUPPORTED) throw Unsupported(msg);
    throw Error(msg);
}
inline char ch(blas::Layout l) { return static_cast<char>(l); }
inline char ch(blas::Op o) { return o == blas::Op::NoTrans ? 'N' : 'T'; }
template <typename T>
constexpr bool is_real() {
    return std::is_same<T, float>::value || std::is_same<T, double>::value;
}
}  // namespace detail

// ---------------------------------------------------------------- RNG
// r123array-like fixed array of N words of type U with incr(n): adds n to
// the little-endian multi-word integer (Random123 array.h semantics, pinned by
// test_r123.cc:706-797).
template <int N, typename U>
struct r123array {
    using value_type = U;
    static constexpr int static_size = N;
    U v[N];
    U &operator[](int i) { return v[i]; }
    const U &operator[](int i) const { return v[i]; }
    r123array &incr(uint64_t n = 1) {
        static_assert(sizeof(U) == 4, "32-bit words");
        uint64_t carry = n;
        for (int i = 0; i < N && carry != 0; ++i) {
            const uint64_t s = (uint64_t)v[i] + (carry & 0xFFFFFFFFull);
            v[i] = (U)(s & 0xFFFFFFFFull);
            carry = (carry >> 32) + (s >> 32);
        }
        return *this;
    }
    bool operator==(const r123array &o) const { return std::memcmp(v, o.v, sizeof(v)) == 0; }
    bool operator!=(const r123array &o) const { return !(*this == o); }
};

// Philox4x32-10 (Random123), the reference's DefaultRNG (base.hh:53).
struct Philox4x32 {
    using ctr_type = r123array<4, uint32_t>;
    using key_type = r123array<2, uint32_t>;
};
using DefaultRNG = Philox4x32;

// RNGState (base.hh:64-164).
template <typename RNG = DefaultRNG>
struct RNGState {
    static_assert(std::is_same<RNG, Philox4x32>::value, "randblas_amd implements Philox4x32-10 only");
    using generator = RNG;
    using ctr_type = typename RNG::ctr_type;
    using key_type = typename RNG::key_type;
    using ctr_uint = typename ctr_type::value_type;
    using key_uint = typename key_type::value_type;
    static constexpr int len_c = ctr_type::static_size;
    static constexpr int len_k = key_type::static_size;

    ctr_type counter;
    key_type key;

    RNGState() : counter{}, key{} {}
    RNGState(uint64_t k) : counter{}, key{} { key.incr(k); }
    RNGState(const key_type &k) : counter{}, key(k) {}
    RNGState(const ctr_type &c, const key_type &k) : counter(c), key(k) {}
    bool operator==(const RNGState &s) const { return counter == s.counter && key == s.key; }
    bool operator!=(const RNGState &s) const { return !(*this == s); }

    rbx_rng_state c() const {
        rbx_rng_state s;
        for (int i = 0; i < 4; ++i) s.counter[i] = counter.v[i];
        for (int i = 0; i < 2; ++i) s.key[i] = key.v[i];
        return s;
    }
    static RNGState from_c(const rbx_rng_state &s) {
        RNGState r;
        for (int i = 0; i < 4; ++i) r.counter.v[i] = s.counter[i];
        for (int i = 0; i < 2; ++i) r.key.v[i] = s.key[i];
        return r;
    }
};

// ---------------------------------------------------------------- dense operators
enum class ScalarDist : char { Gaussian = 'G', Uniform = 'U' };
enum class Axis : char { Short = 'S', Long = 'L' };

// DenseDist (dense_skops.hh:231-350).
struct DenseDist {
    const int64_t n_rows;
    const int64_t n_cols;
    const Axis major_axis;
    const int64_t dim_major;
    const int64_t dim_minor;
    const double isometry_scale;
    const ScalarDist family;
    const blas::Layout natural_layout;

    DenseDist(int64_t n_rows, int64_t n_cols, ScalarDist family = ScalarDist::Gaussian,
              Axis major_axis = Axis::Long)
        : DenseDist(n_rows, n_cols, family, major_axis, props(n_rows, n_cols, family, major_axis)) {}

    rbx_dense_dist c() const {
        return rbx_dense_dist{n_rows, n_cols, static_cast<char>(family), static_cast<char>(major_axis)};
    }

  private:
    struct Props {
        int64_t dmaj, dmin;
        char lay;
        double iso;
    };
    static Props props(int64_t r, int64_t c, ScalarDist f, Axis a) {
        rbx_dense_dist d{r, c, static_cast<char>(f), static_cast<char>(a)};
        Props p{};
        detail::check(rbx_dense_dist_props(&d, &p.dmaj, &p.dmin, &p.lay, &p.iso), "DenseDist");
        return p;
    }
    DenseDist(int64_t r, int64_t c, ScalarDist f, Axis a, Props p)
        : n_rows(r), n_cols(c), major_axis(a), dim_major(p.dmaj), dim_minor(p.dmin), isometry_scale(p.iso),
          family(f), natural_layout(p.lay == 'C' ? blas::Layout::ColMajor : blas::Layout::RowMajor) {}
};

namespace dense {
// dense::compute_next_state (dense_skops.hh:172-185).
template <typename RNG>
RNGState<RNG> compute_next_state(const DenseDist &D, const RNGState<RNG> &state) {
    rbx_dense_dist d = D.c();
    rbx_rng_state s = state.c();
    rbx_dense_skop tmp;
    detail::check(rbx_dense_skop_init(&d, &s, &tmp), "compute_next_state");
    return RNGState<RNG>::from_c(tmp.next_state);
}
}  // namespace dense

// DenseSkOp (dense_skops.hh:357-478).  buff is a device pointer.
template <typename T, typename RNG = DefaultRNG>
struct DenseSkOp {
    static_assert(detail::is_real<T>(), "T must be float or double");
    using distribution_t = DenseDist;
    using state_t = RNGState<RNG>;
    using scalar_t = T;

    const DenseDist dist;
    const state_t seed_state;
    const state_t next_state;
    const int64_t n_rows;
    const int64_t n_cols;
    bool own_memory;
    T *buff = nullptr;
    const blas::Layout layout;

    DenseSkOp(DenseDist dist, const state_t &seed_state)
        : dist(dist), seed_state(seed_state), next_state(dense::compute_next_state(dist, seed_state)),
          n_rows(dist.n_rows), n_cols(dist.n_cols), own_memory(true), buff(nullptr),
          layout(dist.natural_layout) {}

    DenseSkOp(DenseSkOp &&S)
        : dist(S.dist), seed_state(S.seed_state), next_state(S.next_state), n_rows(S.n_rows), n_cols(S.n_cols),
          own_memory(S.own_memory), buff(S.buff), layout(S.layout) {
        S.buff = nullptr;
    }
    DenseSkOp(const DenseSkOp &) = delete;
    DenseSkOp &operator=(const DenseSkOp &) = delete;

    ~DenseSkOp() {
        if (own_memory && buff != nullptr) {
            rbx_dense_skop c = this->c();
            c.own_memory = 1;
            (void)rbx_dense_skop_free(&c);
        }
    }

    rbx_dense_skop c() const {
        rbx_dense_skop s;
        s.dist = dist.c();
        s.seed_state = seed_state.c();
        s.next_state = next_state.c();
        s.buff = (void *)buff;
        s.layout = detail::ch(layout);
        s.own_memory = 0;
        return s;
    }
};

// fill_dense_unpacked (dense_skops.hh:563-606).
template <typename T, typename RNG>
RNGState<RNG> fill_dense_unpacked(blas::Layout layout, const DenseDist &D, int64_t n_rows, int64_t n_cols,
                                  int64_t ro_s, int64_t co_s, T *buff, const RNGState<RNG> &seed) {
    static_assert(detail::is_real<T>(), "T must be float or double");
    rbx_dense_dist d = D.c();
    rbx_rng_state s = seed.c(), nxt;
    rbx_status st;
    if constexpr (std::is_same<T, float>::value)
        st = rbx_sfill_dense_unpacked(detail::ch(layout), &d, n_rows, n_cols, ro_s, co_s, buff, &s, &nxt,
                                      gpu::stream());
    else
        st = rbx_dfill_dense_unpacked(detail::ch(layout), &d, n_rows, n_cols, ro_s, co_s, buff, &s, &nxt,
                                      gpu::stream());
    detail::check(st, "fill_dense_unpacked");
    return RNGState<RNG>::from_c(nxt);
}

// fill_dense(D, buff, seed) (dense_skops.hh:623-626): the full operator in its
// natural layout.
template <typename T, typename RNG>
RNGState<RNG> fill_dense(const DenseDist &D, T *buff, const RNGState<RNG> &seed) {
    return fill_dense_unpacked(D.natural_layout, D, D.n_rows, D.n_cols, 0, 0, buff, seed);
}

// fill_dense(S) (dense_skops.hh:649-658): allocates S.buff on the device when
// it is null (the operator then owns it) and fills it.
template <typename T, typename RNG>
void fill_dense(DenseSkOp<T, RNG> &S) {
    rbx_dense_skop c = S.c();
    if (S.buff == nullptr && !S.own_memory)
        throw Error("fill_dense: an operator that does not own its memory needs a buffer");
    rbx_status st;
    if constexpr (std::is_same<T, float>::value)
        st = rbx_sfill_dense_skop(&c, gpu::stream());
    else
        st = rbx_dfill_dense_skop(&c, gpu::stream());
    detail::check(st, "fill_dense");
    S.buff = (T *)c.buff;
}

// ---------------------------------------------------------------- sparse operators
// SparseDist (sparse_skops.hh:131-246).
struct SparseDist {
    const int64_t n_rows;
    const int64_t n_cols;
    const Axis major_axis;
    const int64_t dim_major;
    const int64_t dim_minor;
    const double isometry_scale;
    const int64_t vec_nnz;
    const int64_t full_nnz;

    SparseDist(int64_t n_rows, int64_t n_cols, int64_t vec_nnz = 4, Axis major_axis = Axis::Short)
        : SparseDist(n_rows, n_cols, vec_nnz, major_axis, props(n_rows, n_cols, vec_nnz, major_axis)) {}

    rbx_sparse_dist c() const { return rbx_sparse_dist{n_rows, n_cols, vec_nnz, static_cast<char>(major_axis)}; }

  private:
    struct Props {
        int64_t dmaj, dmin, full;
        double iso;
    };
    static Props props(int64_t r, int64_t c, int64_t k, Axis a) {
        rbx_sparse_dist d{r, c, k, static_cast<char>(a)};
        Props p{};
        detail::check(rbx_sparse_dist_props(&d, &p.dmaj, &p.dmin, &p.full, &p.iso), "SparseDist");
        return p;
    }
    SparseDist(int64_t r, int64_t c, int64_t k, Axis a, Props p)
        : n_rows(r), n_cols(c), major_axis(a), dim_major(p.dmaj), dim_minor(p.dmin), isometry_scale(p.iso),
          vec_nnz(k), full_nnz(p.full) {}
};

// compute_next_state(SparseDist, state) (sparse_skops.hh:266-283).
template <typename RNG>
RNGState<RNG> compute_next_state(const SparseDist &D, const RNGState<RNG> &state) {
    rbx_sparse_dist d = D.c();
    rbx_rng_state s = state.c();
    rbx_sparse_skop tmp;
    detail::check(rbx_sparse_skop_init(&d, &s, &tmp), "compute_next_state");
    return RNGState<RNG>::from_c(tmp.next_state);
}

// SparseSkOp (sparse_skops.hh:289-450).  vals / rows / cols are device arrays.
template <typename T, typename RNG = DefaultRNG, typename sint_t = int64_t>
struct SparseSkOp {
    static_assert(detail::is_real<T>(), "T must be float or double");
    static_assert(std::is_same<sint_t, int64_t>::value, "randblas_amd sparse operators use int64_t indices");
    using distribution_t = SparseDist;
    using state_t = RNGState<RNG>;
    using scalar_t = T;
    using index_t = sint_t;

    const SparseDist dist;
    const state_t seed_state;
    const state_t next_state;
    const int64_t n_rows;
    const int64_t n_cols;
    bool own_memory;
    int64_t nnz;
    T *vals;
    sint_t *rows;
    sint_t *cols;
    // the library's prepared copy of the sampled entries (rbx_sparse_skop::prep)
    void *prep = nullptr;

    SparseSkOp(SparseDist dist, const state_t &seed_state)
        : dist(dist), seed_state(seed_state), next_state(compute_next_state(dist, seed_state)),
          n_rows(dist.n_rows), n_cols(dist.n_cols), own_memory(true), nnz(-1), vals(nullptr), rows(nullptr),
          cols(nullptr) {}

    SparseSkOp(SparseDist dist, const state_t &seed_state, const state_t &next_state, int64_t nnz, T *vals,
               sint_t *rows, sint_t *cols)
        : dist(dist), seed_state(seed_state), next_state(next_state), n_rows(dist.n_rows), n_cols(dist.n_cols),
          own_memory(false), nnz(nnz), vals(vals), rows(rows), cols(cols) {}

    SparseSkOp(SparseSkOp &&S)
        : dist(S.dist), seed_state(S.seed_state), next_state(S.next_state), n_rows(S.n_rows), n_cols(S.n_cols),
          own_memory(S.own_memory), nnz(S.nnz), vals(S.vals), rows(S.rows), cols(S.cols), prep(S.prep) {
        S.vals = nullptr;
        S.rows = nullptr;
        S.cols = nullptr;
        S.nnz = -1;
        S.prep = nullptr;
    }
    SparseSkOp(const SparseSkOp &) = delete;
    SparseSkOp &operator=(const SparseSkOp &) = delete;

    ~SparseSkOp() {
        const bool arrays = own_memory && (vals != nullptr || rows != nullptr || cols != nullptr);
        if (arrays || prep != nullptr) {
            rbx_sparse_skop c = this->c();
            c.own_memory = arrays ? 1 : 0;
            (void)rbx_sparse_skop_free(&c);
        }
    }

    rbx_sparse_skop c() const {
        rbx_sparse_skop s;
        s.dist = dist.c();
        s.seed_state = seed_state.c();
        s.next_state = next_state.c();
        s.nnz = nnz;
        s.vals = (void *)vals;
        s.rows = rows;
        s.cols = cols;
        s.own_memory = 0;
        s.prep = prep;
        return s;
    }
};

// fill_sparse_unpacked_nosub (sparse_skops.hh:515-565).  Synchronises the
// stream (nnz is a host value).
template <typename T, typename sint_t, typename RNG>
RNGState<RNG> fill_sparse_unpacked_nosub(const SparseDist &D, int64_t &nnz, T *vals, sint_t *rows, sint_t *cols,
                                         const RNGState<RNG> &seed) {
    static_assert(std::is_same<sint_t, int64_t>::value, "int64_t indices");
    rbx_sparse_dist d = D.c();
    rbx_rng_state s = seed.c(), nxt;
    rbx_status st;
    if constexpr (std::is_same<T, float>::value)
        st = rbx_sfill_sparse_unpacked_nosub(&d, &nnz, vals, rows, cols, &s, &nxt, gpu::stream());
    else
        st = rbx_dfill_sparse_unpacked_nosub(&d, &nnz, vals, rows, cols, &s, &nxt, gpu::stream());
    detail::check(st, "fill_sparse_unpacked_nosub");
    return RNGState<RNG>::from_c(nxt);
}

// fill_sparse(S) (sparse_skops.hh:587-603).
template <typename T, typename RNG, typename sint_t>
void fill_sparse(SparseSkOp<T, RNG, sint_t> &S) {
    if (S.vals == nullptr && !S.own_memory)
        throw Error("fill_sparse: an operator that does not own its memory needs arrays");
    rbx_sparse_skop c = S.c();
    rbx_status st;
    if constexpr (std::is_same<T, float>::value)
        st = rbx_sfill_sparse_skop(&c, gpu::stream());
    else
        st = rbx_dfill_sparse_skop(&c, gpu::stream());
    S.prep = c.prep;  // the fill replaced the prepared copy, also when it failed
    detail::check(st, "fill_sparse");
    S.vals = (T *)c.vals;
    S.rows = c.rows;
    S.cols = c.cols;
    S.nnz = c.nnz;
}

// ---------------------------------------------------------------- dense sketches
namespace dense {
// lskge3 (skge.hh:154-202).
template <typename T, typename RNG>
void lskge3(blas::Layout layout, blas::Op opS, blas::Op opA, int64_t d, int64_t n, int64_t m, T alpha,
            const DenseSkOp<T, RNG> &S, int64_t ro_s, int64_t co_s, const T *A, int64_t lda, T beta, T *B,
            int64_t ldb) {
    rbx_dense_skop c = S.c();
    rbx_status st;
    if constexpr (std::is_same<T, float>::value)
        st = rbx_slskge3(detail::ch(layout), detail::ch(opS), detail::ch(opA), d, n, m, alpha, &c, ro_s, co_s, A,
                         lda, beta, B, ldb, gpu::stream());
    else
        st = rbx_dlskge3(detail::ch(layout), detail::ch(opS), detail::ch(opA), d, n, m, alpha, &c, ro_s, co_s, A,
                         lda, beta, B, ldb, gpu::stream());
    detail::check(st, "lskge3");
}

// rskge3 (skge.hh:307-355).
template <typename T, typename RNG>
void rskge3(blas::Layout layout, blas::Op opA, blas::Op opS, int64_t m, int64_t d, int64_t n, T alpha, const T *A,
            int64_t lda, const DenseSkOp<T, RNG> &S, int64_t ro_s, int64_t co_s, T beta, T *B, int64_t ldb) {
    rbx_dense_skop c = S.c();
    rbx_status st;
    if constexpr (std::is_same<T, float>::value)
        st = rbx_srskge3(detail::ch(layout), detail::ch(opA), detail::ch(opS), m, d, n, alpha, A, lda, &c, ro_s,
                         co_s, beta, B, ldb, gpu::stream());
    else
        st = rbx_drskge3(detail::ch(layout), detail::ch(opA), detail::ch(opS), m, d, n, alpha, A, lda, &c, ro_s,
                         co_s, beta, B, ldb, gpu::stream());
    detail::check(st, "rskge3");
}
}  // namespace dense

// ---------------------------------------------------------------- sparse sketches
namespace sparse {
// lskges (skge.hh:465-492).
template <typename T, typename RNG, typename sint_t>
void lskges(blas::Layout layout, blas::Op opS, blas::Op opA, int64_t d, int64_t n, int64_t m, T alpha,
            const SparseSkOp<T, RNG, sint_t> &S, int64_t ro_s, int64_t co_s, const T *A, int64_t lda, T beta, T *B,
            int64_t ldb) {
    rbx_sparse_skop c = S.c();
    rbx_status st;
    if constexpr (std::is_same<T, float>::value)
        st = rbx_slskges(detail::ch(layout), detail::ch(opS), detail::ch(opA), d, n, m, alpha, &c, ro_s, co_s, A,
                         lda, beta, B, ldb, gpu::stream());
    else
        st = rbx_dlskges(detail::ch(layout), detail::ch(opS), detail::ch(opA), d, n, m, alpha, &c, ro_s, co_s, A,
                         lda, beta, B, ldb, gpu::stream());
    detail::check(st, "lskges");
}

// rskges (skge.hh:598-626).
template <typename T, typename RNG, typename sint_t>
void rskges(blas::Layout layout, blas::Op opA, blas::Op opS, int64_t m, int64_t d, int64_t n, T alpha, const T *A,
            int64_t lda, const SparseSkOp<T, RNG, sint_t> &S, int64_t ro_s, int64_t co_s, T beta, T *B,
            int64_t ldb) {
    rbx_sparse_skop c = S.c();
    rbx_status st;
    if constexpr (std::is_same<T, float>::value)
        st = rbx_srskges(detail::ch(layout), detail::ch(opA), detail::ch(opS), m, d, n, alpha, A, lda, &c, ro_s,
                         co_s, beta, B, ldb, gpu::stream());
    else
        st = rbx_drskges(detail::ch(layout), detail::ch(opA), detail::ch(opS), m, d, n, alpha, A, lda, &c, ro_s,
                         co_s, beta, B, ldb, gpu::stream());
    detail::check(st, "rskges");
}
}  // namespace sparse

// ---------------------------------------------------------------- sketch_general
// (skge.hh:756-1199): submatrix forms ...
template <typename T, typename RNG>
void sketch_general(blas::Layout layout, blas::Op opS, blas::Op opA, int64_t d, int64_t n, int64_t m, T alpha,
                    const DenseSkOp<T, RNG> &S, int64_t ro_s, int64_t co_s, const T *A, int64_t lda, T beta, T *B,
                    int64_t ldb) {
    dense::lskge3(layout, opS, opA, d, n, m, alpha, S, ro_s, co_s, A, lda, beta, B, ldb);
}
template <typename T, typename RNG, typename sint_t>
void sketch_general(blas::Layout layout, blas::Op opS, blas::Op opA, int64_t d, int64_t n, int64_t m, T alpha,
                    const SparseSkOp<T, RNG, sint_t> &S, int64_t ro_s, int64_t co_s, const T *A, int64_t lda,
                    T beta, T *B, int64_t ldb) {
    sparse::lskges(layout, opS, opA, d, n, m, alpha, S, ro_s, co_s, A, lda, beta, B, ldb);
}
template <typename T, typename RNG>
void sketch_general(blas::Layout layout, blas::Op opA, blas::Op opS, int64_t m, int64_t d, int64_t n, T alpha,
                    const T *A, int64_t lda, const DenseSkOp<T, RNG> &S, int64_t ro_s, int64_t co_s, T beta, T *B,
                    int64_t ldb) {
    dense::rskge3(layout, opA, opS, m, d, n, alpha, A, lda, S, ro_s, co_s, beta, B, ldb);
}
template <typename T, typename RNG, typename sint_t>
void sketch_general(blas::Layout layout, blas::Op opA, blas::Op opS, int64_t m, int64_t d, int64_t n, T alpha,
                    const T *A, int64_t lda, const SparseSkOp<T, RNG, sint_t> &S, int64_t ro_s, int64_t co_s,
                    T beta, T *B, int64_t ldb) {
    sparse::rskges(layout, opA, opS, m, d, n, alpha, A, lda, S, ro_s, co_s, beta, B, ldb);
}

namespace detail {
// the full-operator overloads require op(S) to be exactly r x c
// (skge.hh:1089-1095, 1191-1197)
template <typename SKOP>
void require_full(const SKOP &S, blas::Op opS, int64_t r, int64_t c) {
    const bool ok = opS == blas::Op::NoTrans ? (S.n_rows == r && S.n_cols == c) : (S.n_rows == c && S.n_cols == r);
    if (!ok) throw Error("sketch_general: operator dimensions do not match the full-operator call");
}
}  // namespace detail

// ... and full-operator forms.
template <typename T, typename SKOP>
auto sketch_general(blas::Layout layout, blas::Op opS, blas::Op opA, int64_t d, int64_t n, int64_t m, T alpha,
                    const SKOP &S, const T *A, int64_t lda, T beta, T *B, int64_t ldb)
    -> decltype((void)S.seed_state, void()) {
    detail::require_full(S, opS, d, m);
    sketch_general(layout, opS, opA, d, n, m, alpha, S, 0, 0, A, lda, beta, B, ldb);
}
template <typename T, typename SKOP>
auto sketch_general(blas::Layout layout, blas::Op opA, blas::Op opS, int64_t m, int64_t d, int64_t n, T alpha,
                    const T *A, int64_t lda, const SKOP &S, T beta, T *B, int64_t ldb)
    -> decltype((void)S.seed_state, void()) {
    detail::require_full(S, opS, n, d);
    sketch_general(layout, opA, opS, m, d, n, alpha, A, lda, S, 0, 0, beta, B, ldb);
}

// ---------------------------------------------------------------- sketch_vector
// (skve.hh:140-246): y = alpha * op(submat(S)) * x + beta * y, with (d, m) the
// dimensions of submat(S) before op; computed as a RowMajor sketch_general
// with one column and the increments as leading dimensions.
template <typename T, typename SKOP>
void sketch_vector(blas::Op opS, int64_t d, int64_t m, T alpha, const SKOP &S, int64_t ro_s, int64_t co_s,
                   const T *x, int64_t incx, T beta, T *y, int64_t incy) {
    const int64_t _d = opS == blas::Op::Trans ? m : d;
    const int64_t _m = opS == blas::Op::Trans ? d : m;
    sketch_general(blas::Layout::RowMajor, opS, blas::Op::NoTrans, _d, (int64_t)1, _m, alpha, S, ro_s, co_s, x,
                   incx, beta, y, incy);
}
template <typename T, typename SKOP>
void sketch_vector(blas::Op opS, T alpha, const SKOP &S, const T *x, int64_t incx, T beta, T *y, int64_t incy) {
    sketch_vector(opS, S.n_rows, S.n_cols, alpha, S, 0, 0, x, incx, beta, y, incy);
}

// ---------------------------------------------------------------- sketch_symmetric
namespace util {
// util::require_symmetric (util.hh:127-148), on the device; synchronises.
template <typename T>
void require_symmetric(blas::Layout layout, const T *A, int64_t n, int64_t lda, T tol) {
    rbx_status st;
    if constexpr (std::is_same<T, float>::value)
        st = rbx_srequire_symmetric(detail::ch(layout), A, n, lda, tol, gpu::stream());
    else
        st = rbx_drequire_symmetric(detail::ch(layout), A, n, lda, tol, gpu::stream());
    detail::check(st, "require_symmetric");
}
}  // namespace util

// ---------------------------------------------------------------- index sampling
// weights_to_cdf (util.hh:459-472), in place on the device; synchronises.
template <typename T>
void weights_to_cdf(int64_t n, T *w, T error_if_below = -std::sqrt(std::numeric_limits<T>::epsilon())) {
    rbx_status st;
    if constexpr (std::is_same<T, float>::value) st = rbx_sweights_to_cdf(n, w, error_if_below, gpu::stream());
    else st = rbx_dweights_to_cdf(n, w, error_if_below, gpu::stream());
    detail::check(st, "weights_to_cdf");
}

// sample_indices_iid (util.hh:493-514): k iid draws from a device CDF.
template <typename T, typename sint_t, typename state_t = RNGState<>>
state_t sample_indices_iid(int64_t n, const T *cdf, int64_t k, sint_t *samples, const state_t &state) {
    static_assert(std::is_same<sint_t, int64_t>::value, "int64_t indices");
    rbx_rng_state s = state.c(), next;
    rbx_status st;
    if constexpr (std::is_same<T, float>::value)
        st = rbx_ssample_indices_iid(n, cdf, k, samples, &s, &next, gpu::stream());
    else
        st = rbx_dsample_indices_iid(n, cdf, k, samples, &s, &next, gpu::stream());
    detail::check(st, "sample_indices_iid");
    return state_t::from_c(next);
}

// sample_indices_iid_uniform with Rademacher signs (util.hh:516-550).
template <typename T, typename sint_t, typename state_t = RNGState<>>
state_t sample_indices_iid_uniform(int64_t n, int64_t k, sint_t *samples, T *rademachers, const state_t &state) {
    static_assert(std::is_same<sint_t, int64_t>::value, "int64_t indices");
    rbx_rng_state s = state.c(), next;
    rbx_status st;
    if constexpr (std::is_same<T, float>::value)
        st = rbx_ssample_indices_iid_uniform_signed(n, k, samples, rademachers, &s, &next, gpu::stream());
    else
        st = rbx_dsample_indices_iid_uniform_signed(n, k, samples, rademachers, &s, &next, gpu::stream());
    detail::check(st, "sample_indices_iid_uniform");
    return state_t::from_c(next);
}

// sample_indices_iid_uniform (util.hh:552-560).
template <typename sint_t = int64_t, typename state_t = RNGState<>>
state_t sample_indices_iid_uniform(int64_t n, int64_t k, sint_t *samples, const state_t &state) {
    static_assert(std::is_same<sint_t, int64_t>::value, "int64_t indices");
    rbx_rng_state s = state.c(), next;
    detail::check(rbx_sample_indices_iid_uniform(n, k, samples, &s, &next, gpu::stream()),
                  "sample_indices_iid_uniform");
    return state_t::from_c(next);
}

// sketch_symmetric (sksy.hh:158-530): A is n x n, symmetric, stored as a
// general matrix; checked first with require_symmetric(sym_check_tol).
template <typename T, typename SKOP>
void sketch_symmetric(blas::Layout layout, int64_t d, int64_t n, T alpha, const SKOP &S, int64_t ro_s, int64_t co_s,
                      const T *A, int64_t lda, T beta, T *B, int64_t ldb, T sym_check_tol = 0) {
    util::require_symmetric(layout, A, n, lda, sym_check_tol);
    sketch_general(layout, blas::Op::NoTrans, blas::Op::NoTrans, d, n, n, alpha, S, ro_s, co_s, A, lda, beta, B,
                   ldb);
}
template <typename T, typename SKOP>
void sketch_symmetric(blas::Layout layout, int64_t n, int64_t d, T alpha, const T *A, int64_t lda, const SKOP &S,
                      int64_t ro_s, int64_t co_s, T beta, T *B, int64_t ldb, T sym_check_tol = 0) {
    util::require_symmetric(layout, A, n, lda, sym_check_tol);
    sketch_general(layout, blas::Op::NoTrans, blas::Op::NoTrans, n, d, n, alpha, A, lda, S, ro_s, co_s, beta, B,
                   ldb);
}
template <typename T, typename SKOP>
auto sketch_symmetric(blas::Layout layout, T alpha, const SKOP &S, const T *A, int64_t lda, T beta, T *B,
                      int64_t ldb, T sym_check_tol = 0) -> decltype((void)S.seed_state, void()) {
    sketch_symmetric(layout, S.n_rows, S.n_cols, alpha, S, 0, 0, A, lda, beta, B, ldb, sym_check_tol);
}
template <typename T, typename SKOP>
auto sketch_symmetric(blas::Layout layout, T alpha, const T *A, int64_t lda, const SKOP &S, T beta, T *B,
                      int64_t ldb, T sym_check_tol = 0) -> decltype((void)S.seed_state, void()) {
    sketch_symmetric(layout, S.n_rows, S.n_cols, alpha, A, lda, S, 0, 0, beta, B, ldb, sym_check_tol);
}

// ---------------------------------------------------------------- sparse data
namespace sparse_data {
enum class IndexBase : int { Zero = 0, One = 1 };

// CSRMatrix (sparse_data/csr_matrix.hh:73-245), zero-based, device arrays,
// borrowed (the expert constructor; this layer never allocates CSR data).
template <typename T, typename sint_t = int64_t>
struct CSRMatrix {
    static_assert(std::is_same<sint_t, int64_t>::value, "int64_t indices");
    using scalar_t = T;
    using index_t = sint_t;
    const int64_t n_rows;
    const int64_t n_cols;
    const bool own_memory;
    int64_t nnz;
    IndexBase index_base;
    T *vals;
    sint_t *rowptr;
    sint_t *colidxs;

    CSRMatrix(int64_t n_rows, int64_t n_cols, int64_t nnz, T *vals, sint_t *rowptr, sint_t *colidxs,
              IndexBase index_base = IndexBase::Zero)
        : n_rows(n_rows), n_cols(n_cols), own_memory(false), nnz(nnz), index_base(index_base), vals(vals),
          rowptr(rowptr), colidxs(colidxs) {
        if (index_base != IndexBase::Zero) throw Unsupported("CSRMatrix: only zero-based indices are supported");
    }
    rbx_csr_matrix c() const { return rbx_csr_matrix{n_rows, n_cols, nnz, (void *)vals, rowptr, colidxs}; }
};

// CSCMatrix (sparse_data/csc_matrix.hh:74-244), zero-based, device arrays, borrowed.
template <typename T, typename sint_t = int64_t>
struct CSCMatrix {
    static_assert(std::is_same<sint_t, int64_t>::value, "int64_t indices");
    using scalar_t = T;
    using index_t = sint_t;
    const int64_t n_rows;
    const int64_t n_cols;
    const bool own_memory;
    int64_t nnz;
    IndexBase index_base;
    T *vals;
    sint_t *rowidxs;
    sint_t *colptr;

    CSCMatrix(int64_t n_rows, int64_t n_cols, int64_t nnz, T *vals, sint_t *rowidxs, sint_t *colptr,
              IndexBase index_base = IndexBase::Zero)
        : n_rows(n_rows), n_cols(n_cols), own_memory(false), nnz(nnz), index_base(index_base), vals(vals),
          rowidxs(rowidxs), colptr(colptr) {
        if (index_base != IndexBase::Zero) throw Unsupported("CSCMatrix: only zero-based indices are supported");
    }
    rbx_csc_matrix c() const { return rbx_csc_matrix{n_rows, n_cols, nnz, (void *)vals, rowidxs, colptr}; }
};

// NonzeroSort (sparse_data/base.hh:103-107).
enum class NonzeroSort : char { CSC = 'C', CSR = 'R', None = 'N' };

// COOMatrix (sparse_data/coo_matrix.hh:83-308), zero-based, device arrays,
// borrowed; the entries may be in any order (the library sorts a copy).
template <typename T, typename sint_t = int64_t>
struct COOMatrix {
    static_assert(std::is_same<sint_t, int64_t>::value, "int64_t indices");
    using scalar_t = T;
    using index_t = sint_t;
    const int64_t n_rows;
    const int64_t n_cols;
    const bool own_memory;
    int64_t nnz;
    IndexBase index_base;
    T *vals;
    sint_t *rows;
    sint_t *cols;

    COOMatrix(int64_t n_rows, int64_t n_cols, int64_t nnz, T *vals, sint_t *rows, sint_t *cols,
              IndexBase index_base = IndexBase::Zero)
        : n_rows(n_rows), n_cols(n_cols), own_memory(false), nnz(nnz), index_base(index_base), vals(vals),
          rows(rows), cols(cols) {
        if (index_base != IndexBase::Zero) throw Unsupported("COOMatrix: only zero-based indices are supported");
    }
    rbx_coo_matrix c() const { return rbx_coo_matrix{n_rows, n_cols, nnz, (void *)vals, rows, cols}; }

    // sort_arrays (coo_matrix.hh:222-226): stable, in place on the device.
    void sort_arrays(NonzeroSort s) {
        if (s == NonzeroSort::None) return;
        rbx_coo_matrix a = c();
        rbx_status st;
        if constexpr (std::is_same<T, float>::value) st = rbx_scoo_sort((char)s, &a, gpu::stream());
        else st = rbx_dcoo_sort((char)s, &a, gpu::stream());
        if (st != RBX_OK) throw Error(std::string("sort_arrays: ") + rbx_last_error_message());
    }
    // coo_arrays_determine_sort (sparse_data/base.hh:142-167); synchronises.
    NonzeroSort determine_sort() const {
        rbx_coo_matrix a = c();
        char o = 'N';
        if (rbx_coo_determine_sort(&a, &o, gpu::stream()) != RBX_OK)
            throw Error(std::string("determine_sort: ") + rbx_last_error_message());
        return static_cast<NonzeroSort>(o);
    }
};

// Format conversions (sparse_data/conversions.hh:49-126) into caller-allocated
// device arrays of the destination (nnz entries; n_rows + 1 / n_cols + 1
// pointers).  COO -> compressed sorts stably.
namespace formats {
#define RBX_CONV_OVERLOAD(T, PFX)                                                        \
    inline rbx_status conv(const rbx_csr_matrix *a, rbx_coo_matrix *b, T *) {            \
        return rbx_##PFX##csr_to_coo(a, b, gpu::stream());                               \
    }                                                                                    \
    inline rbx_status conv(const rbx_csc_matrix *a, rbx_coo_matrix *b, T *) {            \
        return rbx_##PFX##csc_to_coo(a, b, gpu::stream());                               \
    }                                                                                    \
    inline rbx_status conv(const rbx_coo_matrix *a, rbx_csr_matrix *b, T *) {            \
        return rbx_##PFX##coo_to_csr(a, b, gpu::stream());                               \
    }                                                                                    \
    inline rbx_status conv(const rbx_coo_matrix *a, rbx_csc_matrix *b, T *) {            \
        return rbx_##PFX##coo_to_csc(a, b, gpu::stream());                               \
    }
RBX_CONV_OVERLOAD(float, s)
RBX_CONV_OVERLOAD(double, d)
#undef RBX_CONV_OVERLOAD
template <typename From, typename To>
void convert(const From &src, To &dst, const char *name) {
    static_assert(std::is_same<typename From::scalar_t, typename To::scalar_t>::value, "same scalar type");
    auto a = src.c();
    auto b = dst.c();
    if (conv(&a, &b, (typename From::scalar_t *)nullptr) != RBX_OK)
        throw Error(std::string(name) + ": " + rbx_last_error_message());
    dst.nnz = b.nnz;
}
}  // namespace formats

template <typename T>
void csr_to_coo(const CSRMatrix<T, int64_t> &csr, COOMatrix<T, int64_t> &coo) {
    formats::convert(csr, coo, "csr_to_coo");
}
template <typename T>
void csc_to_coo(const CSCMatrix<T, int64_t> &csc, COOMatrix<T, int64_t> &coo) {
    formats::convert(csc, coo, "csc_to_coo");
}
template <typename T>
void coo_to_csr(const COOMatrix<T, int64_t> &coo, CSRMatrix<T, int64_t> &csr) {
    formats::convert(coo, csr, "coo_to_csr");
}
template <typename T>
void coo_to_csc(const COOMatrix<T, int64_t> &coo, CSCMatrix<T, int64_t> &csc) {
    formats::convert(coo, csc, "coo_to_csc");
}

namespace formats {
// One C entry point per (precision, format) pair.
#define RBX_SKSP3_OVERLOAD(T, PFX, MAT, SFX)                                                                           \
    inline rbx_status rsksp3_c(char l, char oa, char os, int64_t m, int64_t d, int64_t n, T alpha, const MAT *a,      \
                               int64_t ro_a, int64_t co_a, const rbx_dense_skop *s, int64_t ro_s, int64_t co_s,        \
                               T beta, T *B, int64_t ldb, void *st) {                                                 \
        return rbx_##PFX##rsksp3##SFX(l, oa, os, m, d, n, alpha, a, ro_a, co_a, s, ro_s, co_s, beta, B, ldb, st);     \
    }                                                                                                                 \
    inline rbx_status lsksp3_c(char l, char os, char oa, int64_t d, int64_t n, int64_t m, T alpha,                   \
                               const rbx_dense_skop *s, int64_t ro_s, int64_t co_s, const MAT *a, int64_t ro_a,        \
                               int64_t co_a, T beta, T *B, int64_t ldb, void *st) {                                   \
        return rbx_##PFX##lsksp3##SFX(l, os, oa, d, n, m, alpha, s, ro_s, co_s, a, ro_a, co_a, beta, B, ldb, st);     \
    }
RBX_SKSP3_OVERLOAD(float, s, rbx_csr_matrix, )
RBX_SKSP3_OVERLOAD(double, d, rbx_csr_matrix, )
RBX_SKSP3_OVERLOAD(float, s, rbx_csc_matrix, _csc)
RBX_SKSP3_OVERLOAD(double, d, rbx_csc_matrix, _csc)
RBX_SKSP3_OVERLOAD(float, s, rbx_coo_matrix, _coo)
RBX_SKSP3_OVERLOAD(double, d, rbx_coo_matrix, _coo)
#undef RBX_SKSP3_OVERLOAD
}  // namespace formats

// rsksp3 (sparse_data/sksp.hh:277-326): B = alpha * op(submat(A)) * op(submat(S)) + beta * B,
// A a CSRMatrix, CSCMatrix or COOMatrix.
template <typename T, typename SpMat, typename RNG>
void rsksp3(blas::Layout layout, blas::Op opA, blas::Op opS, int64_t m, int64_t d, int64_t n, T alpha,
            const SpMat &A, int64_t ro_a, int64_t co_a, const DenseSkOp<T, RNG> &S, int64_t ro_s, int64_t co_s,
            T beta, T *B, int64_t ldb) {
    static_assert(std::is_same<typename SpMat::scalar_t, T>::value, "A and B must have the same scalar type");
    auto a = A.c();
    rbx_dense_skop s = S.c();
    detail::check(formats::rsksp3_c(detail::ch(layout), detail::ch(opA), detail::ch(opS), m, d, n, alpha, &a, ro_a,
                                   co_a, &s, ro_s, co_s, beta, B, ldb, gpu::stream()),
                  "rsksp3");
}

// lsksp3 (sparse_data/sksp.hh:132-180): B = alpha * op(submat(S)) * op(submat(A)) + beta * B.
template <typename T, typename RNG, typename SpMat>
void lsksp3(blas::Layout layout, blas::Op opS, blas::Op opA, int64_t d, int64_t n, int64_t m, T alpha,
            const DenseSkOp<T, RNG> &S, int64_t ro_s, int64_t co_s, const SpMat &A, int64_t ro_a, int64_t co_a,
            T beta, T *B, int64_t ldb) {
    static_assert(std::is_same<typename SpMat::scalar_t, T>::value, "A and B must have the same scalar type");
    auto a = A.c();
    rbx_dense_skop s = S.c();
    detail::check(formats::lsksp3_c(detail::ch(layout), detail::ch(opS), detail::ch(opA), d, n, m, alpha, &s, ro_s,
                                   co_s, &a, ro_a, co_a, beta, B, ldb, gpu::stream()),
                  "lsksp3");
}
}  // namespace sparse_data

// sketch_sparse, left-sketch form (sparse_data/sksp.hh:418-439).
template <typename T, typename RNG, typename SpMat>
auto sketch_sparse(blas::Layout layout, blas::Op opS, blas::Op opA, int64_t d, int64_t n, int64_t m, T alpha,
                   const DenseSkOp<T, RNG> &S, int64_t ro_s, int64_t co_s, const SpMat &A, T beta, T *B,
                   int64_t ldb) -> decltype((void)A.c(), void()) {
    sparse_data::lsksp3(layout, opS, opA, d, n, m, alpha, S, ro_s, co_s, A, 0, 0, beta, B, ldb);
}

// sketch_sparse, right-sketch form (sparse_data/sksp.hh:520-539).
template <typename T, typename SpMat, typename RNG>
auto sketch_sparse(blas::Layout layout, blas::Op opA, blas::Op opS, int64_t m, int64_t d, int64_t n, T alpha,
                   const SpMat &A, const DenseSkOp<T, RNG> &S, int64_t ro_s, int64_t co_s, T beta, T *B,
                   int64_t ldb) -> decltype((void)A.c(), void()) {
    sparse_data::rsksp3(layout, opA, opS, m, d, n, alpha, A, 0, 0, S, ro_s, co_s, beta, B, ldb);
}

// ---------------------------------------------------------------- extensions
// (no reference counterpart)
namespace amd {
// Adjoint of lskge3 with respect to A: Abar = alpha * op(submat(S))^T * Bbar + beta * Abar.
template <typename T, typename RNG>
void lskge3_adjoint(blas::Layout layout, blas::Op opS, int64_t d, int64_t n, int64_t m, T alpha,
                    const DenseSkOp<T, RNG> &S, int64_t ro_s, int64_t co_s, const T *Bbar, int64_t ldbb, T beta,
                    T *Abar, int64_t ldab) {
    rbx_dense_skop c = S.c();
    rbx_status st;
    if constexpr (std::is_same<T, float>::value)
        st = rbx_slskge3_adjoint(detail::ch(layout), detail::ch(opS), d, n, m, alpha, &c, ro_s, co_s, Bbar, ldbb,
                                 beta, Abar, ldab, gpu::stream());
    else
        st = rbx_dlskge3_adjoint(detail::ch(layout), detail::ch(opS), d, n, m, alpha, &c, ro_s, co_s, Bbar, ldbb,
                                 beta, Abar, ldab, gpu::stream());
    detail::check(st, "lskge3_adjoint");
}

// Adjoint of rskge3 with respect to A: Abar = alpha * Bbar * op(submat(S))^T + beta * Abar.
template <typename T, typename RNG>
void rskge3_adjoint(blas::Layout layout, blas::Op opS, int64_t m, int64_t d, int64_t n, T alpha, const T *Bbar,
                    int64_t ldbb, const DenseSkOp<T, RNG> &S, int64_t ro_s, int64_t co_s, T beta, T *Abar,
                    int64_t ldab) {
    rbx_dense_skop c = S.c();
    rbx_status st;
    if constexpr (std::is_same<T, float>::value)
        st = rbx_srskge3_adjoint(detail::ch(layout), detail::ch(opS), m, d, n, alpha, Bbar, ldbb, &c, ro_s, co_s,
                                 beta, Abar, ldab, gpu::stream());
    else
        st = rbx_drskge3_adjoint(detail::ch(layout), detail::ch(opS), m, d, n, alpha, Bbar, ldbb, &c, ro_s, co_s,
                                 beta, Abar, ldab, gpu::stream());
    detail::check(st, "rskge3_adjoint");
}

// Gradient of B = alpha * A * op(submat(S)) (A CSR, opA = NoTrans) with
// respect to A's stored values: dvals[k] = alpha * sum_j Bbar[i_k, j] * op(submat(S))[c_k, j].
template <typename T, typename RNG>
void rsksp3_adjoint(blas::Layout layout, blas::Op opS, int64_t m, int64_t d, int64_t n, T alpha,
                    const sparse_data::CSRMatrix<T, int64_t> &A, const DenseSkOp<T, RNG> &S, int64_t ro_s,
                    int64_t co_s, const T *Bbar, int64_t ldbb, T *dvals) {
    rbx_csr_matrix a = A.c();
    rbx_dense_skop s = S.c();
    rbx_status st;
    if constexpr (std::is_same<T, float>::value)
        st = rbx_srsksp3_adjoint(detail::ch(layout), detail::ch(opS), m, d, n, alpha, &a, &s, ro_s, co_s, Bbar, ldbb,
                                 dvals, gpu::stream());
    else
        st = rbx_drsksp3_adjoint(detail::ch(layout), detail::ch(opS), m, d, n, alpha, &a, &s, ro_s, co_s, Bbar, ldbb,
                                 dvals, gpu::stream());
    detail::check(st, "rsksp3_adjoint");
}

// BF16-split left sketch of an FP32 matrix (n_terms bfloat16 terms of A).
template <typename RNG>
void lskge3_bf16split(blas::Layout layout, blas::Op opS, blas::Op opA, int64_t d, int64_t n, int64_t m, float alpha,
                      const DenseSkOp<float, RNG> &S, int64_t ro_s, int64_t co_s, const float *A, int64_t lda,
                      float beta, float *B, int64_t ldb, int n_terms = 3) {
    rbx_dense_skop c = S.c();
    detail::check(rbx_slskge3_bf16split(detail::ch(layout), detail::ch(opS), detail::ch(opA), d, n, m, alpha, &c,
                                        ro_s, co_s, A, lda, beta, B, ldb, n_terms, gpu::stream()),
                  "lskge3_bf16split");
}
}  // namespace amd

}  // namespace RandBLAS

#endif  // RANDBLAS_AMD_CPP_HH
