// rbx_gemm_common.hh -- the column-major product description shared by the
// dense sketch kernels, and the two tile producers: operator tiles generated
// from Philox counters, and strided dense tiles loaded from memory.
#pragma once
#include "rbx_internal.hh"

namespace rbx {

struct GemmProblem {
    int64_t M, N, K;
    double alpha, beta;
    const void *D;        // dense operand
    int64_t drs, dcs;     // element (r,c) of D at D[r*drs + c*dcs]
    GenMap g;             // generated operator (used when sbuf == nullptr)
    const void *sbuf;     // materialised operator, element (p,q) at sbuf[p*srs + q*scs]
    int64_t soff, srs, scs;  // element offset of G(0,0) in sbuf
    int gen_is_x;         // 1: X = operator, Y = D; 0: X = D, Y = operator
    void *C;
    int64_t crs, ccs;
    // split K: nsplit > 1 runs K ranges of kw (the last one shorter) as
    // blockIdx.z, each writing its raw product to part + z * M * N
    // (column-major M x N); a reduction then forms alpha * sum + beta * C
    int nsplit = 1;
    int64_t kw = 0;
    void *part = nullptr;
    // the last column of D and of C held apart from the rest (a right-hand
    // side beside a matrix): D's column N - 1 at xD[k * xdrs], C's at
    // xC[i * xcrs]; only the wide kernels' tail takes these
    const void *xD = nullptr;
    int64_t xdrs = 0;
    void *xC = nullptr;
    int64_t xcrs = 0;
};

// The sub-problem of K range z: the dense operand and the operator are
// advanced along K, the output is the split's partial (alpha 1, beta 0).
template <typename T>
__device__ __forceinline__ GemmProblem split_view(const GemmProblem &in, int z) {
    if (in.nsplit <= 1) return in;
    GemmProblem q = in;
    const int64_t k0 = (int64_t)z * in.kw;
    q.K = in.K - k0 < in.kw ? in.K - k0 : in.kw;
    if (in.gen_is_x) {
        // X = operator (M x K, K = its column index q), Y = D (K x N, K = row)
        q.D = (const T *)in.D + k0 * in.drs;
        if (in.sbuf) q.soff = in.soff + k0 * in.scs;
        else if (in.g.b_along_rows) q.g.a_off = in.g.a_off + k0;
        else q.g.b_off = in.g.b_off + k0;
    } else {
        // X = D (M x K, K = column), Y = operator (K x N, K = its row index p)
        q.D = (const T *)in.D + k0 * in.dcs;
        if (in.sbuf) q.soff = in.soff + k0 * in.srs;
        else if (in.g.b_along_rows) q.g.b_off = in.g.b_off + k0;
        else q.g.a_off = in.g.a_off + k0;
    }
    q.C = (T *)in.part + (int64_t)z * in.M * in.N;
    q.crs = 1;
    q.ccs = in.M;
    q.alpha = 1.0;
    q.beta = 0.0;
    return q;
}

// Point the problem at S->buff when the operator is materialised
// (lskge3/rskge3 with S.buff != nullptr: offset_and_ldim, skge.hh:195-198).
inline void set_mem_operand(GemmProblem &pb, const rbx_dense_skop &S, char opS, int64_t ro_s, int64_t co_s,
                            bool transposed) {
    pb.sbuf = nullptr;
    pb.soff = 0;
    if (S.buff == nullptr) return;
    char lay = S.layout;
    if (lay != 'C' && lay != 'R') RBX_FAIL(RBX_INVALID_ARG, "DenseSkOp.layout must be 'C' or 'R' when buff is set");
    const int64_t frs = lay == 'C' ? 1 : S.dist.n_cols;
    const int64_t fcs = lay == 'C' ? S.dist.n_rows : 1;
    const bool swap_rc = (opS == 'T') != transposed;
    pb.sbuf = S.buff;
    pb.soff = ro_s * frs + co_s * fcs;
    pb.srs = swap_rc ? fcs : frs;
    pb.scs = swap_rc ? frs : fcs;
}

// Generate the PR x QC tile G[p0:p0+PR, q0:q0+QC) of the operator into LDS at
// lds[(p-p0)*ldp + (q-q0)*ldq].  Entries outside [0,p_lim) x [0,q_lim) are
// written as zero.  Each thread produces whole Philox counters (four
// consecutive entries along the major axis).
template <typename T, int PR, int QC, int NT>
__device__ __forceinline__ void gen_tile(const GenMap &g, int64_t p0, int64_t q0, int64_t p_lim, int64_t q_lim,
                                         T *lds, int ldp, int ldq, int tid) {
    const bool gauss = g.gauss != 0;
    if (!g.b_along_rows) {
        // major axis along q: one "vector" per tile row p
        const int64_t bstart = g.b_off + q0;
        const int64_t c_first = bstart >> 2;
        const int ncell = (int)(((bstart + QC - 1) >> 2) - c_first + 1);
        for (int idx = tid; idx < PR * ncell; idx += NT) {
            const int pp = idx / ncell, cc = idx - pp * ncell;
            const int64_t p = p0 + pp;
            const int64_t cell = c_first + cc;
            float4 w = make_float4(0.f, 0.f, 0.f, 0.f);
            const bool pv = p < p_lim;
            if (pv) w = gen4(gauss, ctr_add(g.ctr, (uint64_t)((g.a_off + p) * g.P + cell)), g.k0, g.k1);
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int qq = (int)(4 * cell + t - bstart);
                if (qq >= 0 && qq < QC) {
                    const bool ok = pv && (q0 + qq) < q_lim;
                    lds[pp * ldp + qq * ldq] = ok ? entry_value<T>(f4_get(w, t), gauss) : (T)0;
                }
            }
        }
    } else {
        // major axis along p: one "vector" per tile column q
        const int64_t bstart = g.b_off + p0;
        const int64_t c_first = bstart >> 2;
        const int ncell = (int)(((bstart + PR - 1) >> 2) - c_first + 1);
        for (int idx = tid; idx < QC * ncell; idx += NT) {
            const int qq = idx / ncell, cc = idx - qq * ncell;
            const int64_t q = q0 + qq;
            const int64_t cell = c_first + cc;
            float4 w = make_float4(0.f, 0.f, 0.f, 0.f);
            const bool qv = q < q_lim;
            if (qv) w = gen4(gauss, ctr_add(g.ctr, (uint64_t)((g.a_off + q) * g.P + cell)), g.k0, g.k1);
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int pp = (int)(4 * cell + t - bstart);
                if (pp >= 0 && pp < PR) {
                    const bool ok = qv && (p0 + pp) < p_lim;
                    lds[pp * ldp + qq * ldq] = ok ? entry_value<T>(f4_get(w, t), gauss) : (T)0;
                }
            }
        }
    }
}

// Load the PR x QC tile of a strided matrix (element (p,q) at src[p*rs + q*cs])
// into LDS at lds[(p-p0)*ldp + (q-q0)*ldq], zero-filling outside the limits.
// Consecutive threads walk the unit-stride direction when there is one.
template <typename T, int PR, int QC, int NT>
__device__ __forceinline__ void load_tile(const T *src, int64_t rs, int64_t cs, int64_t p0, int64_t q0, int64_t p_lim,
                                          int64_t q_lim, T *lds, int ldp, int ldq, int tid) {
    if (rs == 1) {
        for (int idx = tid; idx < PR * QC; idx += NT) {
            const int pp = idx % PR, qq = idx / PR;
            const int64_t p = p0 + pp, q = q0 + qq;
            lds[pp * ldp + qq * ldq] = (p < p_lim && q < q_lim) ? src[p + q * cs] : (T)0;
        }
    } else {
        for (int idx = tid; idx < PR * QC; idx += NT) {
            const int qq = idx % QC, pp = idx / QC;
            const int64_t p = p0 + pp, q = q0 + qq;
            lds[pp * ldp + qq * ldq] = (p < p_lim && q < q_lim) ? src[p * rs + q * cs] : (T)0;
        }
    }
}

template <typename T>
void scale_output(T *C, int64_t M, int64_t N, int64_t crs, int64_t ccs, double beta, hipStream_t stream);

template <typename T>
GemmProblem plan_left(char layout, char opS, char opA, int64_t d, int64_t n, int64_t m, double alpha,
                      const rbx_dense_skop *S, int64_t ro_s, int64_t co_s, const T *A, int64_t lda, double beta,
                      T *B, int64_t ldb);

// FP32 problems with a generated operator on the X side and large interior
// shapes: runs the covered part on the wide-tile kernel and returns the
// remaining sub-problems in rest[] (their count), or -1 when it does not
// apply (rbx_bf16.hip).
int f32_wide_run(const GemmProblem &pb, GemmProblem rest[2], hipStream_t stream);
// The same for FP64 problems (tiles of up to 64 x 512).
int f64_wide_run(const GemmProblem &pb, GemmProblem rest[2], hipStream_t stream);

// [B b] = op(S) * [A a] for column-major A (m x n, lda) and B (d x n, ldb)
// with one more column a (m) -> b (d) held apart, alpha 1, beta 0, in one
// pass over the generated operator (the wide kernels' tail column).  Returns
// false, having launched nothing, when the wide kernels do not take the
// shape (the caller then sketches A and a separately).
template <typename T>
bool left_sketch_with_column(char opS, int64_t d, int64_t n, int64_t m, const rbx_dense_skop *S, int64_t ro_s,
                             int64_t co_s, const T *A, int64_t lda, const T *a, T *B, int64_t ldb, T *b,
                             hipStream_t stream);

}  // namespace rbx
