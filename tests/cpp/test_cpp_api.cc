// test_cpp_api.cc -- drives the C++ layer (randblas_amd/cpp/RandBLAS.hh) the
// way a RandBLAS user would.
//
//   test_cpp_api host          host-side semantics only (no GPU needed)
//   test_cpp_api gpu OUTDIR    sketches on the GPU; raw results are written to
//                              OUTDIR/<name>.bin for the Python test to check
//                              against the CPU oracle
//
// Exit status 0 when every in-program check passed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "RandBLAS.hh"

using blas::Layout;
using blas::Op;
using RandBLAS::Axis;
using RandBLAS::DenseDist;
using RandBLAS::DenseSkOp;
using RandBLAS::RNGState;
using RandBLAS::ScalarDist;
using RandBLAS::SparseDist;
using RandBLAS::SparseSkOp;

static int n_checks = 0, n_fail = 0;
#define EXPECT(cond)                                                               \
    do {                                                                           \
        ++n_checks;                                                                \
        if (!(cond)) {                                                             \
            ++n_fail;                                                              \
            std::fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);   \
        }                                                                          \
    } while (0)

#define HIPCHK(x)                                                                  \
    do {                                                                           \
        hipError_t e_ = (x);                                                       \
        if (e_ != hipSuccess) {                                                    \
            std::fprintf(stderr, "HIP error %s at %s:%d\n", hipGetErrorString(e_), \
                         __FILE__, __LINE__);                                      \
            std::exit(2);                                                          \
        }                                                                          \
    } while (0)

template <typename F>
static bool throws_error(F f, const char *needle = nullptr) {
    try {
        f();
    } catch (const RandBLAS::Error &e) {
        return needle == nullptr || std::strstr(e.what(), needle) != nullptr;
    } catch (...) {
        return false;
    }
    return false;
}

static void host_checks() {
    // RNGState(uint64 k): counter 0, key {k mod 2^32, k >> 32} (base.hh:119)
    RNGState<> s(0x100000002ull);
    EXPECT(s.key[0] == 2 && s.key[1] == 1);
    EXPECT(s.counter[0] == 0 && s.counter[1] == 0 && s.counter[2] == 0 && s.counter[3] == 0);
    RNGState<> t;
    EXPECT(t.key[0] == 0 && t.key[1] == 0);
    // multi-word counter increments carry across words (test_r123.cc:706-797)
    RNGState<> u;
    u.counter.v[0] = 0xFFFFFFFFu;
    u.counter.v[1] = 0xFFFFFFFFu;
    u.counter.incr(1);
    EXPECT(u.counter[0] == 0 && u.counter[1] == 0 && u.counter[2] == 1 && u.counter[3] == 0);
    RNGState<> v;
    v.counter.v[0] = 0xFFFFFFFEu;
    v.counter.incr(0x100000005ull);  // 0xFFFFFFFE + 0x1_00000005 = 0x2_00000003
    EXPECT(v.counter[0] == 3 && v.counter[1] == 2 && v.counter[2] == 0);
    // same arithmetic as the library's own incr
    for (uint64_t n : {1ull, 7ull, 0xFFFFFFFFull, 0x123456789ABCull, ~0ull}) {
        RNGState<> a;
        a.counter.v[0] = 0xFFFFFFF0u;
        a.counter.v[1] = 0x80000000u;
        rbx_rng_state c = a.c();
        EXPECT(rbx_rng_state_incr(&c, n) == RBX_OK);
        a.counter.incr(n);
        EXPECT(RNGState<>::from_c(c) == a);
    }

    // DenseDist (dense_skops.hh:313-329) and natural layouts (:187-199)
    DenseDist D(400, 2000);
    EXPECT(D.dim_major == 2000 && D.dim_minor == 400);
    EXPECT(D.natural_layout == Layout::RowMajor);
    EXPECT(std::fabs(D.isometry_scale - std::pow(400.0, -0.5)) < 1e-15);
    EXPECT(DenseDist(2000, 400).natural_layout == Layout::ColMajor);
    DenseDist Ds(400, 2000, ScalarDist::Uniform, Axis::Short);
    EXPECT(Ds.dim_major == 400 && Ds.natural_layout == Layout::ColMajor);
    // next_state = seed + dim_minor * ceil(dim_major / 4) (dense_skops.hh:172-185)
    DenseSkOp<double> S(D, RNGState<>(42));
    EXPECT(S.next_state.counter[0] == 400u * 500u && S.next_state.key[0] == 42);
    EXPECT(S.layout == Layout::RowMajor && S.buff == nullptr && S.own_memory);
    EXPECT(throws_error([] { DenseDist bad(0, 5); }));

    // SparseDist (sparse_skops.hh:208-245)
    SparseDist SD(2000, 100000, 8);
    EXPECT(SD.dim_major == 2000 && SD.dim_minor == 100000 && SD.full_nnz == 800000);
    EXPECT(std::fabs(SD.isometry_scale - std::pow(8.0, -0.5)) < 1e-15);
    SparseSkOp<float> SS(SD, RNGState<>(7));
    EXPECT(SS.next_state.counter[0] == 800000u && SS.nnz == -1 && SS.own_memory);
    EXPECT(throws_error([] { SparseDist bad(10, 20, 11); }, "vec_nnz"));
    SparseDist SL(30, 500, 7, Axis::Long);
    EXPECT(SL.dim_major == 500 && SL.dim_minor == 30);
    EXPECT(RandBLAS::compute_next_state(SL, RNGState<>(1)).counter[0] == 30u * 4u);
}

template <typename T>
static T *dev_alloc(size_t n) {
    T *p = nullptr;
    HIPCHK(hipMalloc((void **)&p, n * sizeof(T) + 16));
    return p;
}

template <typename T>
static void dump(const std::string &dir, const char *name, const T *dev, size_t n) {
    std::vector<T> h(n);
    HIPCHK(hipMemcpy(h.data(), dev, n * sizeof(T), hipMemcpyDeviceToHost));
    std::string path = dir + "/" + name + ".bin";
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) {
        std::perror(path.c_str());
        std::exit(2);
    }
    std::fwrite(h.data(), sizeof(T), n, f);
    std::fclose(f);
}

static void gpu_checks(const std::string &out) {
    hipStream_t stream;
    HIPCHK(hipStreamCreate(&stream));
    RandBLAS::gpu::set_stream(stream);

    // Dense left sketch: S 400 x 2000 Gaussian (key 42), A 2000 x 200 from
    // DenseDist(2000, 200) key 1 (natural ColMajor), B = S * A.
    const int64_t d = 400, m = 2000, n = 200;
    double *A = dev_alloc<double>(m * n), *B = dev_alloc<double>(d * n), *Bt = dev_alloc<double>(n * d);
    DenseDist DA(m, n);
    RNGState<> nxt = RandBLAS::fill_dense(DA, A, RNGState<>(1));
    EXPECT(nxt.counter[0] == 200u * 500u);
    DenseSkOp<double> S(DenseDist(d, m), RNGState<>(42));
    RandBLAS::sketch_general(Layout::ColMajor, Op::NoTrans, Op::NoTrans, d, n, m, 1.0, S, A, m, 0.0, B, d);
    // right form: Bt (n x d) = A^T * S^T
    RandBLAS::sketch_general(Layout::ColMajor, Op::Trans, Op::Trans, n, d, m, 1.0, A, m, S, 0.0, Bt, n);
    HIPCHK(hipStreamSynchronize(stream));
    dump(out, "dense_A", A, m * n);
    dump(out, "dense_B", B, d * n);
    dump(out, "dense_Bt", Bt, n * d);

    // sketch_vector: y = S * A[:, 0]
    double *y = dev_alloc<double>(d);
    RandBLAS::sketch_vector(Op::NoTrans, 1.0, S, A, (int64_t)1, 0.0, y, (int64_t)1);
    HIPCHK(hipStreamSynchronize(stream));
    dump(out, "vector_y", y, d);

    // adjoint: Abar = S^T * B (m x n)
    double *Abar = dev_alloc<double>(m * n);
    RandBLAS::amd::lskge3_adjoint(Layout::ColMajor, Op::NoTrans, d, n, m, 1.0, S, 0, 0, B, d, 0.0, Abar, m);
    HIPCHK(hipStreamSynchronize(stream));
    dump(out, "adjoint_Abar", Abar, m * n);

    // argument checks surface as RandBLAS::Error
    EXPECT(throws_error([&] {
        RandBLAS::sketch_general(Layout::ColMajor, Op::NoTrans, Op::NoTrans, d, n, m, 1.0, S, 0, 0, A, m - 1, 0.0,
                                 B, d);
    }));
    EXPECT(throws_error([&] {
        RandBLAS::sketch_general(Layout::ColMajor, Op::NoTrans, Op::NoTrans, d - 1, n, m, 1.0, S, A, m, 0.0, B, d);
    }));

    // materialised operator: fill_dense(S) allocates and owns a device buffer
    {
        DenseSkOp<float> Sf(DenseDist(64, 100), RNGState<>(5));
        RandBLAS::fill_dense(Sf);
        EXPECT(Sf.buff != nullptr && Sf.own_memory);
        HIPCHK(hipStreamSynchronize(stream));
        dump(out, "dense_fill_f32", Sf.buff, 64 * 100);
    }  // destructor frees Sf.buff

    // sparse left sketch: SASO 200 x 5000, vec_nnz 8, key 7; A2 5000 x 50 Uniform (key 3)
    const int64_t sd = 200, sm = 5000, sn = 50;
    double *A2 = dev_alloc<double>(sm * sn), *B2 = dev_alloc<double>(sd * sn);
    RandBLAS::fill_dense(DenseDist(sm, sn, ScalarDist::Uniform), A2, RNGState<>(3));
    SparseSkOp<double> SS(SparseDist(sd, sm, 8), RNGState<>(7));
    RandBLAS::sketch_general(Layout::ColMajor, Op::NoTrans, Op::NoTrans, sd, sn, sm, 1.0, SS, A2, sm, 0.0, B2, sd);
    HIPCHK(hipStreamSynchronize(stream));
    dump(out, "sparse_A", A2, sm * sn);
    dump(out, "sparse_B", B2, sd * sn);
    RandBLAS::fill_sparse(SS);
    EXPECT(SS.nnz == 8 * sm && SS.vals != nullptr);
    dump(out, "sparse_vals", SS.vals, SS.nnz);
    dump(out, "sparse_rows", SS.rows, SS.nnz);
    dump(out, "sparse_cols", SS.cols, SS.nnz);

    // sketch_symmetric on an asymmetric matrix (test_sketch_symmetric.cc:390-420)
    {
        const double h[9] = {1.0, 0.0, 0.0, 5.0, 2.0, 0.0, 0.0, 0.0, 3.0};
        double *A3 = dev_alloc<double>(9), *B3 = dev_alloc<double>(6);
        HIPCHK(hipMemcpy(A3, h, sizeof(h), hipMemcpyHostToDevice));
        DenseSkOp<double> S3(DenseDist(2, 3, ScalarDist::Uniform, Axis::Short), RNGState<>(42));
        EXPECT(throws_error(
            [&] { RandBLAS::sketch_symmetric(Layout::ColMajor, (int64_t)2, (int64_t)3, 1.0, S3, 0, 0, A3, 3, 0.0, B3, 2); },
            "Symmetry check failed. |A(0,1) - A(1,0)| was 5.000000e+00"));
        HIPCHK(hipFree(A3));
        HIPCHK(hipFree(B3));
    }

    // rsksp3: CSR A (50 x 300, 5 nonzeros per row) times S (300 x 16 Gaussian, key 42)
    {
        const int64_t cm = 50, cn = 300, cd = 16, per = 5;
        std::vector<int64_t> rowptr(cm + 1), colidx;
        std::vector<double> vals;
        for (int64_t i = 0; i < cm; ++i) {
            rowptr[i] = (int64_t)colidx.size();
            for (int64_t k = 0; k < per; ++k) {
                colidx.push_back((i * 7 + k * 61) % cn);
                vals.push_back(std::sin((double)(i * per + k)));
            }
            std::sort(colidx.begin() + rowptr[i], colidx.end());
        }
        rowptr[cm] = (int64_t)colidx.size();
        const int64_t nnz = rowptr[cm];
        double *v = dev_alloc<double>(nnz), *B4 = dev_alloc<double>(cm * cd);
        int64_t *rp = dev_alloc<int64_t>(cm + 1), *ci = dev_alloc<int64_t>(nnz);
        HIPCHK(hipMemcpy(v, vals.data(), nnz * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(rp, rowptr.data(), (cm + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ci, colidx.data(), nnz * sizeof(int64_t), hipMemcpyHostToDevice));
        RandBLAS::sparse_data::CSRMatrix<double> Acsr(cm, cn, nnz, v, rp, ci);
        DenseSkOp<double> S4(DenseDist(cn, cd), RNGState<>(42));
        RandBLAS::sketch_sparse(Layout::ColMajor, Op::NoTrans, Op::NoTrans, cm, cd, cn, 1.0, Acsr, S4, 0, 0, 0.0, B4,
                                cm);
        HIPCHK(hipStreamSynchronize(stream));
        dump(out, "csr_vals", v, nnz);
        dump(out, "csr_rowptr", rp, cm + 1);
        dump(out, "csr_colidx", ci, nnz);
        dump(out, "csr_B", B4, cm * cd);

        // The same matrix as CSC and as COO (entries reversed): with alpha = 1,
        // beta = 0 and ascending columns every format sums the same products in
        // the same order, so the results agree bit for bit.
        std::vector<int64_t> colptr(cn + 1, 0), rowidx(nnz), crow, ccol;
        std::vector<double> cvals(nnz), ovals;
        for (int64_t p = 0; p < nnz; ++p) colptr[colidx[p] + 1]++;
        for (int64_t c = 0; c < cn; ++c) colptr[c + 1] += colptr[c];
        std::vector<int64_t> fill(colptr.begin(), colptr.end() - 1);
        for (int64_t i = 0; i < cm; ++i)
            for (int64_t p = rowptr[i]; p < rowptr[i + 1]; ++p) {
                rowidx[fill[colidx[p]]] = i;
                cvals[fill[colidx[p]]++] = vals[p];
            }
        for (int64_t i = cm - 1; i >= 0; --i)
            for (int64_t p = rowptr[i + 1] - 1; p >= rowptr[i]; --p) {
                crow.push_back(i);
                ccol.push_back(colidx[p]);
                ovals.push_back(vals[p]);
            }
        double *cv = dev_alloc<double>(nnz), *ov = dev_alloc<double>(nnz), *B5 = dev_alloc<double>(cm * cd),
               *B6 = dev_alloc<double>(cm * cd);
        int64_t *cp = dev_alloc<int64_t>(cn + 1), *cr = dev_alloc<int64_t>(nnz), *orow = dev_alloc<int64_t>(nnz),
                *ocol = dev_alloc<int64_t>(nnz);
        HIPCHK(hipMemcpy(cv, cvals.data(), nnz * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ov, ovals.data(), nnz * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(cp, colptr.data(), (cn + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(cr, rowidx.data(), nnz * sizeof(int64_t), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(orow, crow.data(), nnz * sizeof(int64_t), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ocol, ccol.data(), nnz * sizeof(int64_t), hipMemcpyHostToDevice));
        RandBLAS::sparse_data::CSCMatrix<double> Acsc(cm, cn, nnz, cv, cr, cp);
        RandBLAS::sparse_data::COOMatrix<double> Acoo(cm, cn, nnz, ov, orow, ocol);
        RandBLAS::sketch_sparse(Layout::ColMajor, Op::NoTrans, Op::NoTrans, cm, cd, cn, 1.0, Acsc, S4, 0, 0, 0.0, B5,
                                cm);
        RandBLAS::sparse_data::rsksp3(Layout::ColMajor, Op::NoTrans, Op::NoTrans, cm, cd, cn, 1.0, Acoo, 0, 0, S4, 0,
                                      0, 0.0, B6, cm);
        HIPCHK(hipStreamSynchronize(stream));
        std::vector<double> h4(cm * cd), h5(cm * cd), h6(cm * cd);
        HIPCHK(hipMemcpy(h4.data(), B4, cm * cd * sizeof(double), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(h5.data(), B5, cm * cd * sizeof(double), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(h6.data(), B6, cm * cd * sizeof(double), hipMemcpyDeviceToHost));
        EXPECT(std::memcmp(h4.data(), h5.data(), h4.size() * sizeof(double)) == 0);
        EXPECT(std::memcmp(h4.data(), h6.data(), h4.size() * sizeof(double)) == 0);

        // conversions: the reversed COO sorted back to CSR gives the CSR arrays;
        // CSC -> COO is CSC-ordered
        EXPECT(Acoo.determine_sort() == RandBLAS::sparse_data::NonzeroSort::None);
        int64_t *rp2 = dev_alloc<int64_t>(cm + 1), *ci2 = dev_alloc<int64_t>(nnz);
        double *v2 = dev_alloc<double>(nnz);
        RandBLAS::sparse_data::CSRMatrix<double> Acsr2(cm, cn, 0, v2, rp2, ci2);
        RandBLAS::sparse_data::coo_to_csr(Acoo, Acsr2);
        std::vector<int64_t> hrp(cm + 1), hci(nnz);
        std::vector<double> hv(nnz);
        HIPCHK(hipMemcpy(hrp.data(), rp2, (cm + 1) * sizeof(int64_t), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(hci.data(), ci2, nnz * sizeof(int64_t), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(hv.data(), v2, nnz * sizeof(double), hipMemcpyDeviceToHost));
        EXPECT(Acsr2.nnz == nnz && hrp == rowptr && hci == colidx && hv == vals);
        RandBLAS::sparse_data::COOMatrix<double> Bcoo(cm, cn, 0, ov, orow, ocol);
        RandBLAS::sparse_data::csc_to_coo(Acsc, Bcoo);
        EXPECT(Bcoo.nnz == nnz && Bcoo.determine_sort() == RandBLAS::sparse_data::NonzeroSort::CSC);
        Bcoo.sort_arrays(RandBLAS::sparse_data::NonzeroSort::CSR);
        EXPECT(Bcoo.determine_sort() == RandBLAS::sparse_data::NonzeroSort::CSR);
        for (void *p : {(void *)rp2, (void *)ci2, (void *)v2}) HIPCHK(hipFree(p));
        for (void *p : {(void *)cv, (void *)ov, (void *)B5, (void *)B6, (void *)cp, (void *)cr, (void *)orow,
                        (void *)ocol})
            HIPCHK(hipFree(p));
    }

    // index sampling (util.hh:459-560): uniform, signed, and from a CDF
    {
        const int64_t k = 1001, nw = 300;
        int64_t *samp = dev_alloc<int64_t>(k), *samp2 = dev_alloc<int64_t>(k), *samp3 = dev_alloc<int64_t>(k);
        double *rad = dev_alloc<double>(k), *w = dev_alloc<double>(nw);
        RNGState<> seed(42);
        RNGState<> n1 = RandBLAS::sample_indices_iid_uniform(100, k, samp, seed);
        RNGState<> n2 = RandBLAS::sample_indices_iid_uniform(2000, k, samp2, rad, seed);
        std::vector<double> hw(nw);
        for (int64_t i = 0; i < nw; ++i) hw[i] = 1.0 / std::pow((double)i + 1.0, 1.5);
        HIPCHK(hipMemcpy(w, hw.data(), nw * sizeof(double), hipMemcpyHostToDevice));
        RandBLAS::weights_to_cdf(nw, w);
        RNGState<> n3 = RandBLAS::sample_indices_iid(nw, (const double *)w, k, samp3, RNGState<>(7));
        HIPCHK(hipStreamSynchronize(stream));
        EXPECT(n1.counter[0] == (uint32_t)((k + 3) / 4) && n2.counter[0] == (uint32_t)((k + 1) / 2) &&
               n3.counter[0] == (uint32_t)((k + 3) / 4) && n1.key[0] == 42u);
        EXPECT(throws_error([&] { RandBLAS::weights_to_cdf(nw, w, 2.0); }, "weights_to_cdf"));
        dump(out, "iid_uniform", samp, k);
        dump(out, "iid_signed", samp2, k);
        dump(out, "iid_signs", rad, k);
        dump(out, "iid_cdf_samples", samp3, k);
        for (void *p : {(void *)samp, (void *)samp2, (void *)samp3, (void *)rad, (void *)w}) HIPCHK(hipFree(p));
    }

    RandBLAS::gpu::set_stream(nullptr);
    HIPCHK(hipStreamDestroy(stream));
}

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s host | gpu OUTDIR\n", argv[0]);
        return 2;
    }
    const std::string mode = argv[1];
    host_checks();
    if (mode == "gpu") {
        if (argc < 3) return 2;
        gpu_checks(argv[2]);
    }
    std::printf("cpp api (%s): %d checks, %d failed\n", mode.c_str(), n_checks, n_fail);
    return n_fail == 0 ? 0 : 1;
}
