"""GPU parity for sparse operators and sparse products against the oracle.

Bars: SASO / LASO index arrays and values bit-exact (LASO vectors with
duplicates are written in first-occurrence order by both sides); lskges /
rskges bit-exact with +-1 operators (same per-entry summation order as the
reference's fallback kernels), FP32 and FP64; rsksp3 bit-exact in the
reference's dot-product form.  The SASO left sketch of the sparse
benchmark workload (S 2000 x 100000, vec_nnz 8, key 7; A 100000 x 500) runs at
full size.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda"


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def view(buf, r, c, layout, ld):
    st = buf.itemsize
    return np.lib.stride_tricks.as_strided(buf, (r, c), (st, ld * st) if layout == "C" else (ld * st, st))


@pytest.mark.parametrize("shape,k,axis", [((2000, 100000), 8, "S"), ((300, 20), 5, "S"), ((64, 64), 64, "S"),
                                          ((30, 500), 40, "L"), ((700, 40), 9, "L"), ((1, 1), 1, "S")])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_fill_sparse_bit_exact(rb, orc, shape, k, axis, dtype):
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    D = rb.SparseDist(shape[0], shape[1], k, rb.Axis(axis))
    full = D.full_nnz
    vals = torch.empty(full, dtype=tdt, device=DEV)
    rows = torch.empty(full, dtype=torch.int64, device=DEV)
    cols = torch.empty(full, dtype=torch.int64, device=DEV)
    nnz, nxt = rb.fill_sparse_unpacked_nosub(D, vals, rows, cols, rb.RNGState(7, (3, 0, 0, 0)))
    ov, orr, oc, onnz, onxt = orc.fill_sparse(shape[0], shape[1], k, axis, [3, 0, 0, 0], [7, 0], dtype=dtype)
    assert nnz == onnz
    assert np.array_equal(host(rows)[:nnz], orr)
    assert np.array_equal(host(cols)[:nnz], oc)
    assert np.array_equal(host(vals)[:nnz], ov)
    assert list(nxt.counter) == [int(x) for x in onxt]


def _lsk(rb, orc, dtype, layout, opS, opA, Sshape, k, axis, d, n, m, ro, co, alpha, beta, prefill, seed=0):
    rng = np.random.default_rng(seed)
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    rA, cA = (m, n) if opA == "N" else (n, m)
    lda = (rA if layout == "C" else cA) + 1
    ldb = (d if layout == "C" else n) + 2
    A = rng.uniform(-1, 1, lda * (cA if layout == "C" else rA)).astype(dtype)
    B0 = rng.uniform(-1, 1, ldb * (n if layout == "C" else d)).astype(dtype)
    S = rb.SparseSkOp(rb.SparseDist(Sshape[0], Sshape[1], k, rb.Axis(axis)), rb.RNGState(7), dtype=tdt)
    if prefill:
        rb.fill_sparse(S)
    Bd = torch.from_numpy(B0.copy()).to(DEV)
    rb.lskges(layout, opS, opA, d, n, m, alpha, S, ro, co, torch.from_numpy(A).to(DEV), lda, beta, Bd, ldb)
    got = host(Bd)
    vals, rows, cols, nnz, _ = orc.fill_sparse(Sshape[0], Sshape[1], k, axis, [0, 0, 0, 0], [7, 0], dtype=dtype)
    ref = B0.copy()
    orc.left_spmm_coo("d" if dtype == np.float64 else "s", layout, opS, opA, d, n, m, alpha, Sshape, vals, rows,
                      cols, ro, co, A, lda, beta, ref, ldb)
    return view(got, d, n, layout, ldb), view(ref, d, n, layout, ldb)


@pytest.mark.parametrize("prefill", [False, True])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_lskges_saso_config_full_size(rb, orc, dtype, prefill):
    """S = SparseDist(2000, 100000, vec_nnz 8, Short), key 7; A 100000 x 500 ColMajor Uniform(-1,1).
    Unfilled (the call samples and sorts the operator), or filled and so prepared (the apply then
    reads the prepared copy's 16-bit slots, a few rows' chunks spilling past their 12 or 20 slots)."""
    got, ref = _lsk(rb, orc, dtype, "C", "N", "N", (2000, 100000), 8, "S", 2000, 500, 100000, 0, 0, 1.0, 0.0,
                    prefill=prefill)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("layout", ["C", "R"])
@pytest.mark.parametrize("opS", ["N", "T"])
@pytest.mark.parametrize("opA", ["N", "T"])
def test_lskges_variants(rb, orc, layout, opS, opA):
    Sshape = (60, 900)  # submat(S) is (d x m) for NoTrans and (m x d) for Trans
    d, m = (50, 880) if opS == "N" else (880, 50)
    got, ref = _lsk(rb, orc, np.float64, layout, opS, opA, Sshape, 4, "S", d, 33, m, 3, 7, 1.0, -0.5, prefill=True)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("axis", ["S", "L"])
def test_lskges_general_alpha(rb, orc, axis):
    got, ref = _lsk(rb, orc, np.float64, "R", "N", "N", (40, 3000), 24, axis, 40, 20, 3000, 0, 0, 0.37, 1.5,
                    prefill=False)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("layout", ["C", "R"])
@pytest.mark.parametrize("opS", ["N", "T"])
def test_rskges(rb, orc, layout, opS):
    rng = np.random.default_rng(9)
    m, d, n = 41, 30, 1200
    Sshape = (n, d) if opS == "N" else (d, n)
    lda = m if layout == "C" else n
    ldb = m if layout == "C" else d
    A = rng.uniform(-1, 1, m * n)
    B0 = rng.uniform(-1, 1, m * d)
    S = rb.SparseSkOp(rb.SparseDist(*Sshape, 6), rb.RNGState(13))
    Bd = torch.from_numpy(B0.copy()).to(DEV)
    rb.rskges(layout, "N", opS, m, d, n, 1.0, torch.from_numpy(A).to(DEV), lda, S, 0, 0, 2.0, Bd, ldb)
    vals, rows, cols, _, _ = orc.fill_sparse(*Sshape, 6, "S", [0, 0, 0, 0], [13, 0])
    ref = B0.copy()
    # right_spmm -> left_spmm with flipped layout and op (spmm_dispatch.hh:181-219)
    tl = "R" if layout == "C" else "C"
    orc.left_spmm_coo("d", tl, "T" if opS == "N" else "N", "N", d, m, n, 1.0, Sshape, vals, rows, cols, 0, 0, A,
                      lda, 2.0, ref, ldb)
    assert np.array_equal(host(Bd), ref)


def _csr(rng, nrows, ncols, mean_nnz):
    deg = rng.poisson(mean_nnz, nrows)
    rowptr = np.concatenate([[0], np.cumsum(deg)]).astype(np.int64)
    cols = np.concatenate([np.sort(rng.choice(ncols, size=min(k, ncols), replace=False)) for k in deg]).astype(np.int64)
    rowptr = np.concatenate([[0], np.cumsum(np.minimum(deg, ncols))]).astype(np.int64)
    vals = rng.standard_normal(len(cols))
    return rowptr, cols, vals


@pytest.mark.parametrize("layout", ["C", "R"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_rsksp3_csr(rb, orc, layout, dtype):
    rng = np.random.default_rng(17)
    m, n, d = 500, 3000, 64
    rowptr, colidx, vals = _csr(rng, m, n, 20)
    vals = vals.astype(dtype)
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    S = rb.DenseSkOp(rb.DenseDist(n, d), rb.RNGState(42))  # tall, natural ColMajor
    Acsr = rb.CSRMatrix(m, n, torch.from_numpy(vals).to(DEV), torch.from_numpy(rowptr).to(DEV),
                        torch.from_numpy(colidx).to(DEV))
    ldb = m if layout == "C" else d
    B = torch.zeros(m * d, dtype=tdt, device=DEV)
    rb.rsksp3(layout, "N", "N", m, d, n, 1.0, Acsr, 0, 0, S, 0, 0, 0.0, B, ldb)
    got = host(B)
    # reference: materialise S in its natural layout, then left_spmm with opS
    # flipped when layouts differ (sksp.hh:298-324)
    Sm, _ = orc.fill_dense(n, d, [0, 0, 0, 0], [42, 0], dtype=dtype)  # n x d values
    nat = orc.natural_layout(n, d, "L")
    Sbuf = np.ascontiguousarray(Sm.reshape(-1, order="F" if nat == "C" else "C"))
    lds = n if nat == "C" else d
    opS_ref = "N" if nat == layout else "T"
    ref = np.zeros(m * d, dtype)
    orc.left_spmm_csr("d" if dtype == np.float64 else "s", layout, opS_ref, m, d, n, 1.0, vals, rowptr, colidx,
                      Sbuf, lds, 0.0, ref, ldb)
    assert np.array_equal(got, ref)


def test_unsampled_right_sketch_returns(rb):
    # rskges on an operator with nnz < 0 samples it and returns (the reference's
    # skge.hh:616-620 falls through to a failing requirement instead)
    S = rb.SparseSkOp(rb.SparseDist(100, 10, 3), rb.RNGState(2))
    A = torch.randn(5 * 100, dtype=torch.float64, device=DEV)
    B = torch.empty(5 * 10, dtype=torch.float64, device=DEV)
    rb.rskges("C", "N", "N", 5, 10, 100, 1.0, A, 5, S, 0, 0, 0.0, B, 5)
    assert torch.isfinite(B).all()


def _op_submat(orc, Sshape, key, opS, ro, co, rows, cols, dtype):
    """op(submat(S)) (rows x cols) of a Gaussian DenseSkOp, from the oracle."""
    Sm, _ = orc.fill_dense(Sshape[0], Sshape[1], [0, 0, 0, 0], [key, 0], dtype=dtype)
    if opS == "N":
        return np.ascontiguousarray(Sm[ro:ro + rows, co:co + cols])
    return np.ascontiguousarray(Sm[ro:ro + cols, co:co + rows].T)


@pytest.mark.parametrize("layout", ["C", "R"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_rsksp3_materialised_operator(rb, orc, layout, dtype):
    """rsksp3 with S.buff set uses the buffer in S.layout (sksp.hh:318-321)."""
    rng = np.random.default_rng(21)
    m, n, d = 300, 2000, 48
    rowptr, colidx, vals = _csr(rng, m, n, 15)
    vals = vals.astype(dtype)
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    S = rb.DenseSkOp(rb.DenseDist(n, d), rb.RNGState(42))
    rb.fill_dense(S, dtype=tdt)
    Acsr = rb.CSRMatrix(m, n, torch.from_numpy(vals).to(DEV), torch.from_numpy(rowptr).to(DEV),
                        torch.from_numpy(colidx).to(DEV))
    ldb = m if layout == "C" else d
    B = torch.zeros(m * d, dtype=tdt, device=DEV)
    rb.rsksp3(layout, "N", "N", m, d, n, 1.0, Acsr, 0, 0, S, 0, 0, 0.0, B, ldb)
    Sm, _ = orc.fill_dense(n, d, [0, 0, 0, 0], [42, 0], dtype=dtype)
    nat = S.layout
    Sbuf = np.ascontiguousarray(Sm.reshape(-1, order="F" if nat == "C" else "C"))
    opS_ref = "N" if nat == layout else "T"
    ref = np.zeros(m * d, dtype)
    orc.left_spmm_csr("d" if dtype == np.float64 else "s", layout, opS_ref, m, d, n, 1.0, vals, rowptr, colidx,
                      Sbuf, n if nat == "C" else d, 0.0, ref, ldb)
    assert np.array_equal(host(B), ref)


@pytest.mark.parametrize("layout", ["C", "R"])
@pytest.mark.parametrize("opS", ["N", "T"])
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-13), (np.float32, 1e-5)])
@pytest.mark.parametrize("materialised", [False, True])
@pytest.mark.parametrize("d", [300, 301])   # 301: not a multiple of the vector width
def test_rsksp3_adjoint(rb, orc, layout, opS, dtype, tol, materialised, d):
    from conftest import rel_fro
    rng = np.random.default_rng(23)
    m, n = 250, 1500              # d > one LDS chunk for both precisions
    ro, co = 3, 5
    rowptr, colidx, vals = _csr(rng, m, n, 12)
    vals = vals.astype(dtype)
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    Sshape = (n + 10, d + 10) if opS == "N" else (d + 10, n + 10)
    S = rb.DenseSkOp(rb.DenseDist(*Sshape), rb.RNGState(8))
    if materialised:
        rb.fill_dense(S, dtype=tdt)
    Acsr = rb.CSRMatrix(m, n, torch.from_numpy(vals).to(DEV), torch.from_numpy(rowptr).to(DEV),
                        torch.from_numpy(colidx).to(DEV))
    ldbb = m if layout == "C" else d
    Bbar = rng.standard_normal(m * d).astype(dtype)
    dvals = torch.zeros(len(vals), dtype=tdt, device=DEV)
    rb.rsksp3_adjoint(layout, opS, m, d, n, 0.5, Acsr, S, ro, co, torch.from_numpy(Bbar).to(DEV), ldbb, dvals)
    G = _op_submat(orc, Sshape, 8, opS, ro, co, n, d, dtype)  # n x d
    brs, bcs = (1, ldbb) if layout == "C" else (ldbb, 1)
    ref = orc.sddmm_csr("d", m, d, 0.5, rowptr, colidx, Bbar.astype(np.float64), brs, bcs,
                        G.astype(np.float64).reshape(-1), d, 1)
    assert rel_fro(host(dvals), ref) <= tol
    # adjoint identity against the forward product: <vals, dvals> = 0.5 * <A op(S), Bbar>
    if dtype == np.float64:
        B = torch.zeros(m * d, dtype=tdt, device=DEV)
        rb.rsksp3(layout, "N", opS, m, d, n, 1.0, Acsr, 0, 0, S, ro, co, 0.0, B, ldbb)
        lhs = float(np.dot(vals, host(dvals)))
        rhs = 0.5 * float(np.dot(host(B), Bbar))
        assert abs(lhs - rhs) <= 1e-12 * max(abs(rhs), 1.0)


def test_rsksp3_power_law_full_size(rb, orc):
    """A 200000 x 50000 CSR, power-law column degrees (exponent 2.1, 4e6 draws,
    ~20 stored entries per row), N(0,1) FP32 values, generated on the device
    from key 17; S = DenseDist(50000, 256) key 42.
    B = A * S bit-exact against the oracle (csr_spmm_impl.hh order); the
    adjoint with respect to A's values within 1e-5 of the FP64 oracle."""
    from conftest import rel_fro
    from randblas_amd.datagen import power_law_csr
    m, n, d = 200000, 50000, 256
    Acsr = power_law_csr(m, n, 20, 2.1, 17)
    rowptr, colidx, vals = host(Acsr.rowptr), host(Acsr.colidxs), host(Acsr.vals)
    assert 3.0e6 < len(vals) <= 4.0e6
    S = rb.DenseSkOp(rb.DenseDist(n, d), rb.RNGState(42))
    B = torch.zeros(m * d, dtype=torch.float32, device=DEV)
    rb.rsksp3("C", "N", "N", m, d, n, 1.0, Acsr, 0, 0, S, 0, 0, 0.0, B, m)
    Sm, _ = orc.fill_dense(n, d, [0, 0, 0, 0], [42, 0], dtype=np.float32)  # natural ColMajor
    ref = np.zeros(m * d, np.float32)
    orc.left_spmm_csr("s", "C", "N", m, d, n, 1.0, vals, rowptr, colidx,
                      np.ascontiguousarray(Sm.reshape(-1, order="F")), n, 0.0, ref, m)
    assert np.array_equal(host(B), ref)
    Bbar = np.random.default_rng(3).standard_normal(m * d).astype(np.float32)
    dvals = torch.zeros(len(vals), dtype=torch.float32, device=DEV)
    rb.rsksp3_adjoint("C", "N", m, d, n, 1.0, Acsr, S, 0, 0, torch.from_numpy(Bbar).to(DEV), m, dvals)
    refd = orc.sddmm_csr("d", m, d, 1.0, rowptr, colidx, Bbar.astype(np.float64), 1, m,
                         Sm.astype(np.float64).reshape(-1), d, 1)
    assert rel_fro(host(dvals), refd) <= 1e-5


@pytest.mark.parametrize("layout", ["C", "R"])
@pytest.mark.parametrize("opS", ["N", "T"])
@pytest.mark.parametrize("opA", ["N", "T"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_lsksp3(rb, orc, layout, opS, opA, dtype):
    """lsksp3 bit-exact against the reference's reduction (sksp.hh:132-180):
    right_spmm -> left_spmm(opposite layout, op(A) flipped) on the operator
    submatrix in its natural layout (op(S) flipped when that differs)."""
    rng = np.random.default_rng(29)
    d, m, n = 40, 900, 70
    ro, co = 2, 3
    rowsA, colsA = (m, n) if opA == "N" else (n, m)
    rowptr, colidx, vals = _csr(rng, rowsA, colsA, 9)
    vals = vals.astype(dtype)
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    rS, cS = (d, m) if opS == "N" else (m, d)           # submat(S) before op
    Sshape = (rS + ro + 5, cS + co + 4)
    S = rb.DenseSkOp(rb.DenseDist(*Sshape), rb.RNGState(31))
    Acsr = rb.CSRMatrix(rowsA, colsA, torch.from_numpy(vals).to(DEV), torch.from_numpy(rowptr).to(DEV),
                        torch.from_numpy(colidx).to(DEV))
    ldb = d if layout == "C" else n
    B0 = rng.standard_normal(d * n).astype(dtype)
    Bd = torch.from_numpy(B0.copy()).to(DEV)
    rb.sketch_sparse(layout, opS, opA, d, n, m, 0.75, S, ro, co, Acsr, -0.5, Bd, ldb)
    # reference chain
    full, _ = orc.fill_dense(*Sshape, [0, 0, 0, 0], [31, 0], dtype=dtype)
    nat = orc.natural_layout(*Sshape, "L")
    sub = np.ascontiguousarray(full[ro:ro + rS, co:co + cS])
    Ssub = np.ascontiguousarray(sub.reshape(-1, order="F" if nat == "C" else "C"))
    lds = rS if nat == "C" else cS
    opS_ref = opS if nat == layout else ("T" if opS == "N" else "N")
    tl = "R" if layout == "C" else "C"
    ref = B0.copy()
    dt = "d" if dtype == np.float64 else "s"
    if opA == "T":   # left_spmm(tl, 'N', ...) on the CSR matrix
        orc.left_spmm_csr(dt, tl, opS_ref, n, d, m, 0.75, vals, rowptr, colidx, Ssub, lds, -0.5, ref, ldb)
    else:            # left_spmm(tl, 'T', ...): the CSC order of the transposed CSR matrix
        rows = np.repeat(np.arange(rowsA, dtype=np.int64), np.diff(rowptr))
        orc.left_spmm_coo(dt, tl, "T", opS_ref, n, d, m, 0.75, (rowsA, colsA), vals, rows, colidx, 0, 0, Ssub, lds,
                          -0.5, ref, ldb)
    assert np.array_equal(host(Bd), ref)


def _expand(ptr):
    return np.repeat(np.arange(len(ptr) - 1, dtype=np.int64), np.diff(ptr))


def _ref_left_spmm(orc, fmt, dt, layout, opA, opB, d, n, m, alpha, shape, parts, ro_a, co_a, Bm, ldb, beta, C, ldc):
    """The reference's left_spmm dispatch (spmm_dispatch.hh:72-176) for one data
    format, in oracle calls: CSR with op(A) = N and CSC with op(A) = T (whose
    transpose is that CSR matrix) run the CSR row kernels; every other case runs
    the CSC column kernels, i.e. the COO oracle."""
    vals = parts["vals"]
    if fmt == "csr":
        if opA == "N":
            orc.left_spmm_csr(dt, layout, opB, d, n, m, alpha, vals, parts["rowptr"], parts["colidx"], Bm, ldb, beta,
                              C, ldc)
        else:
            orc.left_spmm_coo(dt, layout, "T", opB, d, n, m, alpha, shape, vals, _expand(parts["rowptr"]),
                              parts["colidx"], 0, 0, Bm, ldb, beta, C, ldc)
    elif fmt == "csc":
        if opA == "T":
            orc.left_spmm_csr(dt, layout, opB, d, n, m, alpha, vals, parts["colptr"], parts["rowidx"], Bm, ldb, beta,
                              C, ldc)
        else:
            orc.left_spmm_coo(dt, layout, "N", opB, d, n, m, alpha, shape, vals, parts["rowidx"],
                              _expand(parts["colptr"]), 0, 0, Bm, ldb, beta, C, ldc)
    else:
        orc.left_spmm_coo(dt, layout, opA, opB, d, n, m, alpha, shape, vals, parts["rows"], parts["cols"], ro_a, co_a,
                          Bm, ldb, beta, C, ldc)


def _sparse_in_format(rb, rng, fmt, nrows, ncols, mean_nnz, dtype, pad=(0, 0)):
    """A random sparse matrix as (rb matrix, host parts, stored shape).  COO
    matrices get pad extra rows/columns of entries around the window (offsets
    (pad[0], pad[1])) and a shuffled entry order; CSC has ascending rows per
    column."""
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV)
    if fmt == "coo":
        R, Cc = nrows + 2 * pad[0], ncols + 2 * pad[1]
        rowptr, colidx, vals = _csr(rng, R, Cc, mean_nnz)
        rows = _expand(rowptr)
        perm = rng.permutation(len(vals))
        parts = dict(vals=vals[perm].astype(dtype), rows=rows[perm], cols=colidx[perm])
        return rb.COOMatrix(R, Cc, t(parts["vals"]), t(parts["rows"]), t(parts["cols"])), parts, (R, Cc)
    rowptr, colidx, vals = _csr(rng, nrows, ncols, mean_nnz)
    vals = vals.astype(dtype)
    if fmt == "csr":
        parts = dict(vals=vals, rowptr=rowptr, colidx=colidx)
        return rb.CSRMatrix(nrows, ncols, t(vals), t(rowptr), t(colidx)), parts, (nrows, ncols)
    rows = _expand(rowptr)
    order = np.lexsort((rows, colidx))       # by column, then row
    colptr = np.concatenate([[0], np.cumsum(np.bincount(colidx, minlength=ncols))]).astype(np.int64)
    parts = dict(vals=vals[order], rowidx=rows[order], colptr=colptr)
    return (rb.CSCMatrix(nrows, ncols, t(parts["vals"]), t(parts["rowidx"]), t(colptr)), parts, (nrows, ncols))


def _rsksp3_formats_case(rb, orc, fmt, layout, opA, opS, dtype, m, n, d):
    rng = np.random.default_rng(41)
    ro_s, co_s = 3, 2
    pad = (4, 6) if fmt == "coo" else (0, 0)
    rowsA, colsA = (m, n) if opA == "N" else (n, m)
    A, parts, shape = _sparse_in_format(rb, rng, fmt, rowsA, colsA, 7, dtype, pad)
    rS, cS = (n, d) if opS == "N" else (d, n)
    Sshape = (rS + ro_s + 2, cS + co_s + 5)
    S = rb.DenseSkOp(rb.DenseDist(*Sshape), rb.RNGState(13))
    ldb = m + 3 if layout == "C" else d + 2
    B0 = rng.standard_normal(ldb * (d if layout == "C" else m)).astype(dtype)
    Bd = torch.from_numpy(B0.copy()).to(DEV)
    rb.rsksp3(layout, opA, opS, m, d, n, 0.75, A, pad[0], pad[1], S, ro_s, co_s, -0.5, Bd, ldb)
    full, _ = orc.fill_dense(*Sshape, [0, 0, 0, 0], [13, 0], dtype=dtype)
    nat = orc.natural_layout(*Sshape, "L")
    sub = np.ascontiguousarray(full[ro_s:ro_s + rS, co_s:co_s + cS])
    Ssub = np.ascontiguousarray(sub.reshape(-1, order="F" if nat == "C" else "C"))
    lds = rS if nat == "C" else cS
    opS_ref = opS if nat == layout else ("T" if opS == "N" else "N")
    ref = B0.copy()
    dt = "d" if dtype == np.float64 else "s"
    _ref_left_spmm(orc, fmt, dt, layout, opA, opS_ref, m, d, n, 0.75, shape, parts, pad[0], pad[1], Ssub, lds, -0.5,
                   ref, ldb)
    assert np.array_equal(host(Bd), ref)


@pytest.mark.parametrize("fmt", ["csr", "csc", "coo"])
@pytest.mark.parametrize("layout", ["C", "R"])
@pytest.mark.parametrize("opA", ["N", "T"])
@pytest.mark.parametrize("opS", ["N", "T"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_rsksp3_formats(rb, orc, fmt, layout, opA, opS, dtype):
    """rsksp3 for CSR / CSC / COO data bit-exact against the reference's
    dispatch (sksp.hh:277-326 -> spmm_dispatch.hh:72-176), alpha and beta
    general; COO as a submatrix with entries outside the window."""
    _rsksp3_formats_case(rb, orc, fmt, layout, opA, opS, dtype, 260, 610, 36)


@pytest.mark.parametrize("fmt", ["csr", "csc", "coo"])
@pytest.mark.parametrize("opA", ["N", "T"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_rsksp3_formats_many_rows(rb, orc, fmt, opA, dtype):
    """More than 8192 rows of op(A) on the column-kernel cases: CSR viewed transposed and CSC take the
    stable sort on the row key alone, COO the (row, column) key; all bit-exact."""
    _rsksp3_formats_case(rb, orc, fmt, "C", opA, "N", dtype, 9000, 610, 20)


@pytest.mark.parametrize("fmt", ["csc", "coo"])
@pytest.mark.parametrize("layout", ["C", "R"])
@pytest.mark.parametrize("opA", ["N", "T"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_lsksp3_formats(rb, orc, fmt, layout, opA, dtype):
    """lsksp3 for CSC / COO data bit-exact against the reference's reduction
    (sksp.hh:132-180 -> right_spmm -> left_spmm on the opposite layout with
    op(A) flipped)."""
    rng = np.random.default_rng(43)
    d, m, n = 30, 520, 90
    opS = "N"
    ro_s, co_s = 1, 4
    pad = (2, 5) if fmt == "coo" else (0, 0)
    rowsA, colsA = (m, n) if opA == "N" else (n, m)
    A, parts, shape = _sparse_in_format(rb, rng, fmt, rowsA, colsA, 6, dtype, pad)
    rS, cS = d, m
    Sshape = (rS + ro_s + 3, cS + co_s + 1)
    S = rb.DenseSkOp(rb.DenseDist(*Sshape), rb.RNGState(17))
    ldb = d if layout == "C" else n
    B0 = rng.standard_normal(d * n).astype(dtype)
    Bd = torch.from_numpy(B0.copy()).to(DEV)
    rb.lsksp3(layout, opS, opA, d, n, m, 1.25, S, ro_s, co_s, A, pad[0], pad[1], 0.5, Bd, ldb)
    full, _ = orc.fill_dense(*Sshape, [0, 0, 0, 0], [17, 0], dtype=dtype)
    nat = orc.natural_layout(*Sshape, "L")
    sub = np.ascontiguousarray(full[ro_s:ro_s + rS, co_s:co_s + cS])
    Ssub = np.ascontiguousarray(sub.reshape(-1, order="F" if nat == "C" else "C"))
    lds = rS if nat == "C" else cS
    opS_ref = opS if nat == layout else "T"
    tl = "R" if layout == "C" else "C"
    ref = B0.copy()
    dt = "d" if dtype == np.float64 else "s"
    _ref_left_spmm(orc, fmt, dt, tl, "T" if opA == "N" else "N", opS_ref, n, d, m, 1.25, shape, parts, pad[0],
                   pad[1], Ssub, lds, 0.5, ref, ldb)
    assert np.array_equal(host(Bd), ref)


@pytest.mark.parametrize("fmt", ["csr", "csc"])
@pytest.mark.parametrize("layout", ["C", "R"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("d", [37, 300])
def test_rsksp3_row_kernel_widths(rb, orc, fmt, layout, dtype, d):
    """The row-kernel cases (CSR with op(A) = N, CSC with op(A) = T) at an odd
    width (scalar kernel) and at a width spanning two vector chunks in FP64,
    bit-exact, alpha and beta general, ldb padded."""
    rng = np.random.default_rng(47)
    m, n = 170, 900
    opA = "N" if fmt == "csr" else "T"
    rowsA, colsA = (m, n) if opA == "N" else (n, m)
    A, parts, shape = _sparse_in_format(rb, rng, fmt, rowsA, colsA, 9, dtype)
    S = rb.DenseSkOp(rb.DenseDist(n, d), rb.RNGState(19))
    ldb = m + 1 if layout == "C" else d + 3
    B0 = rng.standard_normal(ldb * (d if layout == "C" else m)).astype(dtype)
    Bd = torch.from_numpy(B0.copy()).to(DEV)
    rb.rsksp3(layout, opA, "N", m, d, n, -1.5, A, 0, 0, S, 0, 0, 0.25, Bd, ldb)
    full, _ = orc.fill_dense(n, d, [0, 0, 0, 0], [19, 0], dtype=dtype)
    nat = orc.natural_layout(n, d, "L")
    Sbuf = np.ascontiguousarray(full.reshape(-1, order="F" if nat == "C" else "C"))
    lds = n if nat == "C" else d
    opS_ref = "N" if nat == layout else "T"
    ref = B0.copy()
    _ref_left_spmm(orc, fmt, "d" if dtype == np.float64 else "s", layout, opA, opS_ref, m, d, n, -1.5, shape, parts,
                   0, 0, Sbuf, lds, 0.25, ref, ldb)
    assert np.array_equal(host(Bd), ref)


# ---------------------------------------------------------------- long rows
# Rows of the row kernels' orientation with more than max(256, 4 x mean)
# stored entries are summed in FP64 segments (deterministic, not the
# reference's sequential rounding): shorter rows stay bit-exact against the
# oracle, the long ones are checked against the FP64 product at the
# precision bars.
def _long_row_threshold(nnz, nrows):
    return max(256, 4 * -(-nnz // nrows))


def _csr_long(rng, nrows, ncols, mean, long_rows):
    """CSR with Poisson(mean) rows, rows in `long_rows` (row -> entries) long;
    columns random (repeats allowed), N(0,1) values."""
    deg = rng.poisson(mean, nrows)
    for r, k in long_rows.items():
        deg[r] = k
    rowptr = np.concatenate([[0], np.cumsum(deg)]).astype(np.int64)
    cols = rng.integers(0, ncols, int(rowptr[-1])).astype(np.int64)
    return rowptr, cols, rng.standard_normal(len(cols))


def _dense(rowptr, cols, vals, shape):
    A = np.zeros(shape)
    np.add.at(A, (np.repeat(np.arange(shape[0]), np.diff(rowptr)), cols), vals.astype(np.float64))
    return A


@pytest.mark.parametrize("layout", ["C", "R"])
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-13), (np.float32, 1e-6)])
@pytest.mark.parametrize("d", [64, 37])
def test_rsksp3_long_rows(rb, orc, layout, dtype, tol, d):
    """CSR row kernel (vector and scalar widths) with rows of 300, 5000 and
    40000 entries among Poisson(12) rows."""
    rng = np.random.default_rng(53)
    m, n = 400, 20000
    longs = {3: 300, 10: 5000, 399: 40000}
    rowptr, colidx, vals = _csr_long(rng, m, n, 12, longs)
    vals = vals.astype(dtype)
    S = rb.DenseSkOp(rb.DenseDist(n, d), rb.RNGState(61))
    A = rb.CSRMatrix(m, n, torch.from_numpy(vals).to(DEV), torch.from_numpy(rowptr).to(DEV),
                     torch.from_numpy(colidx).to(DEV))
    ldb = m if layout == "C" else d
    B0 = rng.standard_normal(m * d).astype(dtype)
    Bd = torch.from_numpy(B0.copy()).to(DEV)
    rb.rsksp3(layout, "N", "N", m, d, n, 0.75, A, 0, 0, S, 0, 0, -0.5, Bd, ldb)
    Sm, _ = orc.fill_dense(n, d, [0, 0, 0, 0], [61, 0], dtype=dtype)
    nat = orc.natural_layout(n, d, "L")
    Sbuf = np.ascontiguousarray(Sm.reshape(-1, order="F" if nat == "C" else "C"))
    opS_ref = "N" if nat == layout else "T"
    ref = B0.copy()
    dt = "d" if dtype == np.float64 else "s"
    orc.left_spmm_csr(dt, layout, opS_ref, m, d, n, 0.75, vals, rowptr, colidx, Sbuf, n if nat == "C" else d, -0.5,
                      ref, ldb)
    shape = (d, m) if layout == "C" else (m, d)
    out = host(Bd).reshape(shape)
    ref = ref.reshape(shape)
    out, ref = (out.T, ref.T) if layout == "C" else (out, ref)   # m x d
    short = np.diff(rowptr) <= _long_row_threshold(len(vals), m)
    assert short[3] and not short[10] and not short[399]
    assert np.array_equal(out[short], ref[short])
    B0m = B0.reshape(shape).T if layout == "C" else B0.reshape(shape)
    exact = 0.75 * (_dense(rowptr, colidx, vals, (m, n)) @ Sm.astype(np.float64)) - 0.5 * B0m.astype(np.float64)
    for r in longs:
        err = np.linalg.norm(out[r].astype(np.float64) - exact[r]) / np.linalg.norm(exact[r])
        assert err <= tol, (r, err)


@pytest.mark.parametrize("fmt", ["csc", "coo"])
@pytest.mark.parametrize("layout", ["C", "R"])
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-13), (np.float32, 1e-6)])
def test_lsksp3_long_columns(rb, orc, fmt, layout, dtype, tol):
    """lsksp3 of a tall sparse matrix whose first columns hold 30000, 3000
    and 400 entries (the left sketch of data with power-law column degrees):
    CSC reaches the row kernel on its transpose, COO the sorted-CSR kernel."""
    rng = np.random.default_rng(59)
    m, n, d = 40000, 60, 48
    longs = {0: 30000, 1: 3000, 7: 400}
    # built as the CSR of A^T (n x m): its rows are A's columns
    tptr, trows, tvals = _csr_long(rng, n, m, 8, longs)
    tvals = tvals.astype(dtype)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV)
    cols = np.repeat(np.arange(n, dtype=np.int64), np.diff(tptr))
    if fmt == "csc":
        A = rb.CSCMatrix(m, n, t(tvals), t(trows), t(tptr))
    else:
        perm = rng.permutation(len(tvals))
        A = rb.COOMatrix(m, n, t(tvals[perm]), t(trows[perm]), t(cols[perm]))
    S = rb.DenseSkOp(rb.DenseDist(d, m), rb.RNGState(67))
    ldb = d if layout == "C" else n
    B0 = rng.standard_normal(d * n).astype(dtype)
    Bd = torch.from_numpy(B0.copy()).to(DEV)
    rb.lsksp3(layout, "N", "N", d, n, m, 1.25, S, 0, 0, A, 0, 0, 0.5, Bd, ldb)
    shape = (n, d) if layout == "C" else (d, n)
    out = host(Bd).reshape(shape)
    out = out.T if layout == "C" else out          # d x n
    Sm, _ = orc.fill_dense(d, m, [0, 0, 0, 0], [67, 0], dtype=dtype)
    B0m = B0.reshape(shape).T if layout == "C" else B0.reshape(shape)
    exact = 1.25 * (Sm.astype(np.float64) @ _dense(tptr, trows, tvals, (n, m)).T) + 0.5 * B0m.astype(np.float64)
    for j in range(n):
        err = np.linalg.norm(out[:, j].astype(np.float64) - exact[:, j]) / np.linalg.norm(exact[:, j])
        assert err <= (tol if j in longs else max(tol, 1e-6 if dtype == np.float32 else 1e-14)), (j, err)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_long_rows_beta_zero_ignores_nan(rb, dtype):
    """beta = 0 leaves the output of the long-row segments unread."""
    rng = np.random.default_rng(61)
    m, n, d = 64, 5000, 32
    rowptr, colidx, vals = _csr_long(rng, m, n, 4, {5: 20000, 40: 3000})
    vals = vals.astype(dtype)
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    A = rb.CSRMatrix(m, n, torch.from_numpy(vals).to(DEV), torch.from_numpy(rowptr).to(DEV),
                     torch.from_numpy(colidx).to(DEV))
    S = rb.DenseSkOp(rb.DenseDist(n, d), rb.RNGState(2))
    for layout in ("C", "R"):
        B = torch.full((m * d,), float("nan"), dtype=tdt, device=DEV)
        rb.rsksp3(layout, "N", "N", m, d, n, 1.0, A, 0, 0, S, 0, 0, 0.0, B, m if layout == "C" else d)
        assert bool(torch.isfinite(B).all()), layout


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-13), (np.float32, 1e-5)])
@pytest.mark.parametrize("layout", ["C", "R"])
def test_lskges_small_sketch_dimension(rb, orc, dtype, tol, layout):
    """A SASO operator with few, long rows (d = 60, m = 200000, vec_nnz 8:
    rows of ~26700 entries) is split into segments to fill the GPU; the
    result is within the precision bar of the oracle's sequential product
    (not bit-exact; in FP32 the oracle's own sequential rounding over 26700
    terms is ~3e-6)."""
    d, m, n = 60, 200000, 40
    got, ref = _lsk(rb, orc, dtype, layout, "N", "N", (d, m), 8, "S", d, n, m, 0, 0, 1.0, 0.0, prefill=False)
    exact_err = np.linalg.norm(np.asarray(got, np.float64) - np.asarray(ref, np.float64)) / np.linalg.norm(ref)
    assert exact_err <= tol, exact_err


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("layout", ["C", "R"])
def test_lskges_shuffled_operator_with_duplicates(rb, orc, dtype, layout):
    """An operator given as arrays in random order, with duplicate (row,
    column) entries and one row of ~6500 entries: the counting-sort CSR
    builder must order each row by (column, stored position) -- sorting that
    row, longer than one 4096-entry LDS tile, by tiles and merge passes -- so
    the product equals the reference's (coo_spmm_impl.hh:52-105: CSC sort,
    then ascending columns, duplicates in stored order) bit for bit.  (The
    row stays under the row kernels' long-row threshold, 4 x the mean row,
    above which row-major products sum in FP64 segments.)"""
    rng = np.random.default_rng(71)
    d, m, n = 20, 20000, 33
    r = rng.integers(0, d, 30000)
    c = rng.integers(0, m, 30000)
    r = np.concatenate([r, np.full(5000, 3), r[:500]])        # a long row, and 500 duplicates
    c = np.concatenate([c, rng.integers(0, m, 5000), c[:500]])
    v = rng.uniform(-2, 2, len(r)).astype(dtype)
    perm = rng.permutation(len(r))
    r, c, v = r[perm].astype(np.int64), c[perm].astype(np.int64), v[perm]
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    S = rb.SparseSkOp(rb.SparseDist(d, m, 8), rb.RNGState(5), dtype=tdt)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV)
    S.vals, S.rows, S.cols, S.nnz = t(v), t(r), t(c), len(v)
    lda = m if layout == "C" else n
    ldb = d if layout == "C" else n
    A = rng.uniform(-1, 1, m * n).astype(dtype)
    B0 = rng.uniform(-1, 1, d * n).astype(dtype)
    Bd = t(B0.copy())
    rb.lskges(layout, "N", "N", d, n, m, 0.5, S, 0, 0, t(A), lda, 1.5, Bd, ldb)
    ref = B0.copy()
    orc.left_spmm_coo("d" if dtype == np.float64 else "s", layout, "N", "N", d, n, m, 0.5, (d, m), v, r, c, 0, 0, A,
                      lda, 1.5, ref, ldb)
    assert np.array_equal(host(Bd), ref)


def _lsk_run(rb, S, dtype, layout, opS, d, n, m, ro, co, A, lda, B0, ldb, alpha=1.0, beta=0.0, noalloc=False):
    Bd = torch.from_numpy(B0.copy()).to(DEV)
    rb.profile_collect()
    rb.profile_enable(True)
    f = rb.lskges_noalloc if noalloc else rb.lskges
    f(layout, opS, "N", d, n, m, alpha, S, ro, co, torch.from_numpy(A).to(DEV), lda, beta, Bd, ldb)
    torch.cuda.synchronize()
    rb.profile_enable(False)
    return host(Bd), rb.profile_collect()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_lskges_prepared_operator(rb, orc, dtype):
    """fill_sparse prepares the SASO operator (row-sorted copy + LDS entry layout); lskges over the
    whole operator then skips the sort, with results bit-identical to the oracle and to the
    per-call path; windows and opS = 'T' ignore the copy; an in-place edit followed by prepare()
    is picked up."""
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    rng = np.random.default_rng(5)
    d, m, n = 2000, 100000, 40
    S = rb.SparseSkOp(rb.SparseDist(d, m, 8, rb.Axis.Short), rb.RNGState(7), dtype=tdt)
    rb.fill_sparse(S)
    assert S._prep
    A = rng.uniform(-1, 1, m * n).astype(dtype)
    B0 = rng.uniform(-1, 1, d * n).astype(dtype)
    vals, rows, cols, _, _ = orc.fill_sparse(d, m, 8, "S", [0, 0, 0, 0], [7, 0], dtype=dtype)
    pre = "d" if dtype == np.float64 else "s"

    def oracle(vv, layout, opS, dd, mm, ro, co, alpha, beta, lda, ldb):
        ref = B0.copy()
        orc.left_spmm_coo(pre, layout, opS, "N", dd, n, mm, alpha, (d, m), vv, rows, cols, ro, co, A, lda, beta,
                          ref, ldb)
        return ref

    got, prof = _lsk_run(rb, S, dtype, "C", "N", d, n, m, 0, 0, A, m, B0, d, 0.75, -0.5)
    assert "coo_to_csr" not in prof and "csr_lds_apply_kernel" in prof
    assert np.array_equal(got, oracle(vals, "C", "N", d, m, 0, 0, 0.75, -0.5, m, d))
    got_na, prof = _lsk_run(rb, S, dtype, "C", "N", d, n, m, 0, 0, A, m, B0, d, 0.75, -0.5, noalloc=True)
    assert "coo_to_csr" not in prof and np.array_equal(got_na, got)
    # a window inside the operator: the copy does not apply
    got, prof = _lsk_run(rb, S, dtype, "C", "N", d - 3, n, m - 5000, 2, 17, A, m, B0, d)
    assert "coo_to_csr" in prof
    assert np.array_equal(got, oracle(vals, "C", "N", d - 3, m - 5000, 2, 17, 1.0, 0.0, m, d))
    # op(S) = S^T over S[500:2000, 0:1000]
    got, prof = _lsk_run(rb, S, dtype, "C", "T", 1000, n, 1500, 500, 0, A, 1500, B0, 1000)
    assert "coo_to_csr" in prof
    assert np.array_equal(got, oracle(vals, "C", "T", 1000, 1500, 500, 0, 1.0, 0.0, 1500, 1000))
    # in-place edit of the values, then prepare(): the new values are used
    S.vals.mul_(-3.0)
    S.prepare()
    got, prof = _lsk_run(rb, S, dtype, "C", "N", d, n, m, 0, 0, A, m, B0, d)
    assert "coo_to_csr" not in prof
    assert np.array_equal(got, oracle((vals * dtype(-3.0)).astype(dtype), "C", "N", d, m, 0, 0, 1.0, 0.0, m, d))
    # without the copy: the per-call sort gives the same bits
    S.release()
    got2, prof = _lsk_run(rb, S, dtype, "C", "N", d, n, m, 0, 0, A, m, B0, d)
    assert "coo_to_csr" in prof and np.array_equal(got2, got)
    # replaced arrays: the stale copy is not passed
    S.prepare()
    S.vals = S.vals.clone()
    got3, prof = _lsk_run(rb, S, dtype, "C", "N", d, n, m, 0, 0, A, m, B0, d)
    assert "coo_to_csr" in prof and np.array_equal(got3, got)


@pytest.mark.parametrize("axis,shape,k", [("S", (300, 4000), 6), ("L", (300, 4000), 40), ("L", (64, 9000), 30)])
@pytest.mark.parametrize("layout,opA", [("C", "N"), ("R", "T")])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_lskges_prepared_whole_operator(rb, orc, axis, shape, k, layout, opA, dtype):
    """The prepared copy over SASO and LASO operators (LASO rows hold merged duplicates), with op(A)
    of unit stride along the contraction in both layouts: bit-identical to the oracle, no sort."""
    d, m = shape
    n = 24
    rb.profile_collect()
    rb.profile_enable(True)
    try:
        got, ref = _lsk(rb, orc, dtype, layout, "N", opA, shape, k, axis, d, n, m, 0, 0, 0.5, 2.0, prefill=True)
        torch.cuda.synchronize()
    finally:
        rb.profile_enable(False)
    prof = rb.profile_collect()
    assert "coo_to_csr" not in prof and "csr_lds_apply_kernel" in prof
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("axis,k", [("S", 8), ("L", 40)])
@pytest.mark.parametrize("layout,opA", [("C", "N"), ("R", "T")])
@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (0.75, -0.5)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_lskges_prepared_wide(rb, orc, axis, k, layout, opA, alpha, beta, dtype):
    """Prepared operators applied to 400 columns, so the LDS apply runs 8-column blocks: a SASO's
    values are all +1 or -1, and its layout carries their signs (no value loads); a LASO's merged
    duplicates are not, and its values are read.  Both bit-identical to the oracle, in both
    summation forms, with alpha = 1 and otherwise."""
    d, m, n = 2000, 20000, 400
    rb.profile_collect()
    rb.profile_enable(True)
    try:
        got, ref = _lsk(rb, orc, dtype, layout, "N", opA, (d, m), k, axis, d, n, m, 0, 0, alpha, beta, prefill=True)
        torch.cuda.synchronize()
    finally:
        rb.profile_enable(False)
    prof = rb.profile_collect()
    assert "coo_to_csr" not in prof and "csr_lds_apply_kernel" in prof
    assert np.array_equal(got, ref)
