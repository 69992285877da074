"""GPU: the BLAS-like wrappers sketch_general / sketch_vector /
sketch_symmetric and util::require_symmetric.

Cases follow the reference's wrapper tests:
  * test/test_matmul_wrappers/test_sketch_symmetric.cc:192-420 -- left and
    right, "sketch" (d < n) and "lift" (d > n), lda = n and lda > n, beta 0 and
    -1, Short and Long operators, same and opposing layouts, checked against a
    dense symmetric product with the operator materialised by the oracle, and
    the error raised for an asymmetric A (:390-420);
  * test/test_matmul_wrappers/test_sketch_vector.cc:185-215 -- tall and wide
    operators, transposed application, checked against the matrix form;
  * sketch_general's full-operator overloads (skge.hh:1089-1095, 1191-1197):
    same result as the submatrix form at (0, 0), dimension mismatch rejected.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda"


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def as_mat(buf, r, c, layout, ld):
    st = buf.itemsize
    return np.lib.stride_tricks.as_strided(buf, (r, c), (st, ld * st) if layout == "C" else (ld * st, st))


def sym_matrix(rng, n, lda):
    X = rng.standard_normal((n, n))
    X = X + X.T
    A = np.zeros((lda, lda))
    A[:n, :n] = X
    return A  # symmetric, so its layout does not matter


@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("d,lda", [(3, 10), (3, 19), (13, 10), (50, 19)])
@pytest.mark.parametrize("beta", [0.0, -1.0])
@pytest.mark.parametrize("axis", ["S", "L"])
@pytest.mark.parametrize("opposing", [False, True])
@pytest.mark.parametrize("materialised", [True, False])
def test_sketch_symmetric(rb, orc, side, d, lda, beta, axis, opposing, materialised):
    n, alpha = 10, 0.5
    rng = np.random.default_rng(31 + d + lda)
    A = sym_matrix(rng, n, lda)
    rows_out, cols_out = (d, n) if side == "L" else (n, d)
    D = rb.DenseDist(rows_out, cols_out, rb.ScalarDist.Uniform, rb.Axis(axis))
    S = rb.DenseSkOp(D, rb.RNGState(33))
    if materialised:
        rb.fill_dense(S, dtype=torch.float64)
    Sm, _ = orc.fill_dense(rows_out, cols_out, [0, 0, 0, 0], [33, 0], family="U", major_axis=axis)
    nat = orc.natural_layout(rows_out, cols_out, axis)
    layout = ("R" if nat == "C" else "C") if opposing else nat
    ldb = rows_out if layout == "C" else cols_out
    B0 = rng.uniform(-1, 1, rows_out * cols_out)
    Bd = torch.from_numpy(B0.copy()).to(DEV)
    Ad = torch.from_numpy(A.reshape(-1).copy()).to(DEV)
    if side == "L":
        rb.sketch_symmetric(layout, d, n, alpha, S, 0, 0, Ad, lda, beta, Bd, ldb)
        expect = alpha * Sm @ A[:n, :n] + beta * as_mat(B0, rows_out, cols_out, layout, ldb)
    else:
        rb.sketch_symmetric(layout, n, d, alpha, Ad, lda, S, 0, 0, beta, Bd, ldb)
        expect = alpha * A[:n, :n] @ Sm + beta * as_mat(B0, rows_out, cols_out, layout, ldb)
    got = as_mat(host(Bd), rows_out, cols_out, layout, ldb)
    np.testing.assert_allclose(got, expect, rtol=1e-13, atol=1e-13 * np.abs(expect).max())


@pytest.mark.parametrize("full_form", ["left", "right"])
def test_sketch_symmetric_full_operator_overloads(rb, full_form):
    rng = np.random.default_rng(2)
    n, d = 12, 5
    A = sym_matrix(rng, n, n)
    Ad = torch.from_numpy(A.reshape(-1).copy()).to(DEV)
    if full_form == "left":
        S = rb.DenseSkOp(rb.DenseDist(d, n), rb.RNGState(4))
        B1 = torch.zeros(d * n, dtype=torch.float64, device=DEV)
        B2 = torch.zeros_like(B1)
        rb.sketch_symmetric("C", 1.0, S, Ad, n, 0.0, B1, d)
        rb.sketch_symmetric("C", d, n, 1.0, S, 0, 0, Ad, n, 0.0, B2, d)
    else:
        S = rb.DenseSkOp(rb.DenseDist(n, d), rb.RNGState(4))
        B1 = torch.zeros(d * n, dtype=torch.float64, device=DEV)
        B2 = torch.zeros_like(B1)
        rb.sketch_symmetric("C", 1.0, Ad, n, S, 0.0, B1, n)
        rb.sketch_symmetric("C", n, d, 1.0, Ad, n, S, 0, 0, 0.0, B2, n)
    assert torch.equal(B1, B2)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_sketch_symmetric_rejects_asymmetric(rb, dtype):
    # test_sketch_symmetric.cc:390-420: column-major 3 x 3 with A(0,1) = 5, A(1,0) = 0
    A = torch.tensor([1.0, 0.0, 0.0, 5.0, 2.0, 0.0, 0.0, 0.0, 3.0], dtype=dtype, device=DEV)
    S = rb.DenseSkOp(rb.DenseDist(2, 3, rb.ScalarDist.Uniform, rb.Axis.Short), rb.RNGState(42))
    B = torch.zeros(6, dtype=dtype, device=DEV)
    with pytest.raises(rb.Error, match=r"Symmetry check failed\. \|A\(0,1\) - A\(1,0\)\| was 5\.0+e\+00"):
        rb.sketch_symmetric("C", 2, 3, 1.0, S, 0, 0, A, 3, 0.0, B, 2)
    # a negative tolerance skips the check (util.hh:128-129)
    rb.sketch_symmetric("C", 2, 3, 1.0, S, 0, 0, A, 3, 0.0, B, 2, sym_check_tol=-1.0)


def test_require_symmetric_reports_first_pair_in_loop_order(rb):
    n = 300
    rng = np.random.default_rng(8)
    X = rng.standard_normal((n, n))
    X = X + X.T
    X[7, 200] += 1.0   # row-major loop order: (7, 200) comes before (150, 151)
    X[151, 150] += 1.0
    Ad = torch.from_numpy(X.reshape(-1).copy()).to(DEV)
    with pytest.raises(rb.Error, match=r"A\(7,200\) - A\(200,7\)"):
        rb.require_symmetric("R", Ad, n, n, 0.0)
    # relative tolerance: |a - b| <= (|a| + |b| + 1) * tol passes
    rb.require_symmetric("R", Ad, n, n, 1.0)
    # the column-major view is the transpose: same violating pairs, same first pair
    with pytest.raises(rb.Error, match=r"A\(7,200\) - A\(200,7\)"):
        rb.require_symmetric("C", Ad, n, n, 0.0)
    # lda > n: only the leading n x n block is examined
    Y = np.zeros((n, n + 5))
    Y[:, :n] = X + X.T - np.diag(np.diag(X))
    Y[:, n:] = 1e6
    Yd = torch.from_numpy(Y.reshape(-1).copy()).to(DEV)
    rb.require_symmetric("R", Yd, n, n + 5, 0.0)


@pytest.mark.parametrize("shape,opS", [((50, 200), "N"), ((200, 50), "N"), ((50, 200), "T"), ((200, 50), "T")])
@pytest.mark.parametrize("incx,incy", [(1, 1), (3, 2)])
def test_sketch_vector(rb, orc, shape, opS, incx, incy):
    # test_sketch_vector.cc:185-215 (tall, wide, transposed application)
    rng = np.random.default_rng(5)
    r, c = shape
    dlen, mlen = (r, c) if opS == "N" else (c, r)
    S = rb.DenseSkOp(rb.DenseDist(r, c), rb.RNGState(77))
    x = rng.standard_normal(mlen * incx)
    y0 = rng.standard_normal(dlen * incy)
    xd = torch.from_numpy(x).to(DEV)
    yd = torch.from_numpy(y0.copy()).to(DEV)
    rb.sketch_vector(opS, 0.75, S, xd, incx, 0.5, yd, incy)
    Sm, _ = orc.fill_dense(r, c, [0, 0, 0, 0], [77, 0])
    opSm = Sm if opS == "N" else Sm.T
    expect = 0.75 * opSm @ x[::incx] + 0.5 * y0[::incy]
    got = host(yd)
    np.testing.assert_allclose(got[::incy], expect, rtol=1e-13, atol=1e-13 * np.abs(expect).max())
    # entries between the strided ones are untouched
    mask = np.ones(len(y0), bool)
    mask[::incy] = False
    assert np.array_equal(got[mask], y0[mask])
    # submatrix form with (d, m) of submat(S) before op
    yd2 = torch.from_numpy(y0.copy()).to(DEV)
    rb.sketch_vector(opS, r, c, 0.75, S, 0, 0, xd, incx, 0.5, yd2, incy)
    assert torch.equal(yd, yd2)


@pytest.mark.parametrize("kind", ["dense", "sparse"])
def test_sketch_general_dispatch(rb, kind):
    rng = np.random.default_rng(6)
    d, m, n = 40, 300, 17
    A = torch.from_numpy(rng.standard_normal(m * n)).to(DEV)
    if kind == "dense":
        S = rb.DenseSkOp(rb.DenseDist(d, m), rb.RNGState(9))
    else:
        S = rb.SparseSkOp(rb.SparseDist(d, m, 5), rb.RNGState(9))
    B1 = torch.zeros(d * n, dtype=torch.float64, device=DEV)
    B2 = torch.zeros_like(B1)
    rb.sketch_general("C", "N", "N", d, n, m, 1.0, S, A, m, 0.0, B1, d)
    direct = rb.lskge3 if kind == "dense" else rb.lskges
    direct("C", "N", "N", d, n, m, 1.0, S, 0, 0, A, m, 0.0, B2, d)
    assert torch.equal(B1, B2)
    # right form: B (n x d) = A^T (n x m) * S^T (m x d)
    B3 = torch.zeros(n * d, dtype=torch.float64, device=DEV)
    B4 = torch.zeros_like(B3)
    rb.sketch_general("C", "T", "T", n, d, m, 1.0, A, m, S, 0.0, B3, n)
    rdirect = rb.rskge3 if kind == "dense" else rb.rskges
    rdirect("C", "T", "T", n, d, m, 1.0, A, m, S, 0, 0, 0.0, B4, n)
    assert torch.equal(B3, B4)
    # the full-operator overloads require exact dimensions
    with pytest.raises(rb.Error):
        rb.sketch_general("C", "N", "N", d - 1, n, m, 1.0, S, A, m, 0.0, B1, d)
