"""GPU: sketch-and-solve least squares, x = argmin ||S(Ax - b)||_2, and its
adjoint (the downstream use of a left sketch, as in
examples/total-least-squares/tls_dense_skop.cc:138-171).

  * the solution against numpy's lstsq of the oracle's materialised S times
    A and b (FP64 <= 1e-11 relative, FP32 inputs <= 1e-4), both layouts, with
    submatrix offsets;
  * at a large size, where the oracle is too slow, the normal equations of
    the sketched problem: (SA)^T (SAx - Sb) ~ 0 relative to ||SA|| ||SAx - Sb||;
  * the adjoint against torch's own autograd through a host QR solve of the
    oracle's sketched system (FP64 <= 1e-10) and against FP64 central finite
    differences of the device forward (Richardson-extrapolated, <= 1e-6 of the
    largest directional derivative);
  * argument errors: d < n, short leading dimensions.

The QR of the sketched system is CholeskyQR2 when the system is well
conditioned and Householder otherwise; the accuracy tests run both
(RBX_SOLVE_QR=householder forces Householder), and an ill-conditioned system
(cond 1e10) checks that the Householder fallback is taken.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda"


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.fixture(params=["auto", "householder"])
def qr(request, monkeypatch):
    if request.param == "householder":
        monkeypatch.setenv("RBX_SOLVE_QR", "householder")
    else:
        monkeypatch.delenv("RBX_SOLVE_QR", raising=False)
    return request.param


def _ran(rb, fn):
    torch.cuda.synchronize()
    rb.profile_collect()
    rb.profile_enable(True)
    fn()
    rb.profile_enable(False)
    return rb.profile_collect()


def _dev(a, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV, dtype)


def _store(A, layout):
    """A (m x n numpy) as a flat buffer in `layout` with the tight leading dim."""
    return (A.T if layout == "C" else A).ravel().copy()


def _problem(m, n, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((m, n))
    x_star = rng.standard_normal(n)
    b = A @ x_star + 0.1 * rng.standard_normal(m)
    return A, b


@pytest.mark.parametrize("layout", ["C", "R"])
@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-11), (torch.float32, 1e-4)])
@pytest.mark.parametrize("d,m,n,ro,co", [(200, 3000, 40, 0, 0), (150, 2501, 60, 3, 17), (513, 4099, 1, 1, 2)])
def test_solution_matches_lstsq(rb, qr, orc, layout, dtype, tol, d, m, n, ro, co):
    A, b = _problem(m, n, 7)
    Dr, Dc = d + ro + 5, m + co + 9
    S = rb.DenseSkOp(rb.DenseDist(Dr, Dc), rb.RNGState(key=19, counter=(4, 0, 0, 0)))
    nat = orc.natural_layout(Dr, Dc, "L")
    Sm, _ = orc.fill_dense_unpacked(nat, Dr, Dc, d, m, ro, co, [4, 0, 0, 0], [19, 0])
    x_ref = np.linalg.lstsq(Sm @ A, Sm @ b, rcond=None)[0]
    if dtype == torch.float32:  # the reference solution for the rounded inputs
        A, b = A.astype(np.float32).astype(np.float64), b.astype(np.float32).astype(np.float64)
        x_ref = np.linalg.lstsq(Sm @ A, Sm @ b, rcond=None)[0]
    x = torch.full((n,), np.nan, dtype=dtype, device=DEV)
    rb.sketch_solve(layout, d, m, n, S, ro, co, _dev(_store(A, layout), dtype), m if layout == "C" else n,
                    _dev(b, dtype), x)
    err = np.linalg.norm(host(x).astype(np.float64) - x_ref) / np.linalg.norm(x_ref)
    assert err <= tol, err


def test_normal_equations_at_size(rb, qr):
    m, n, d = 262144, 256, 1024
    A = torch.empty(m * n, dtype=torch.float64, device=DEV)
    rb.fill_dense(rb.DenseDist(m, n), A, rb.RNGState(11))  # column-major m x n
    b = torch.empty(m, dtype=torch.float64, device=DEV)
    rb.fill_dense(rb.DenseDist(m, 1), b, rb.RNGState(12))
    S = rb.DenseSkOp(rb.DenseDist(d, m), rb.RNGState(42))
    x = torch.empty(n, dtype=torch.float64, device=DEV)
    rb.sketch_solve("C", d, m, n, S, 0, 0, A, m, b, x)
    SA = torch.empty(d * n, dtype=torch.float64, device=DEV)
    rb.lskge3("C", "N", "N", d, n, m, 1.0, S, 0, 0, A, m, 0.0, SA, d)
    r = torch.empty(d, dtype=torch.float64, device=DEV)
    rb.lskge3("C", "N", "N", d, 1, m, 1.0, S, 0, 0, b, m, 0.0, r, d)
    SAm = SA.view(n, d).T
    r = SAm @ x - r
    g = SAm.T @ r
    rel = float(torch.linalg.norm(g) / (torch.linalg.norm(SAm) * torch.linalg.norm(r)))
    assert rel <= 1e-12, rel
    assert bool(torch.isfinite(x).all())


# Column-major systems whose width n fills whole wide tiles (or leaves more
# than the tail) sketch [A b] in one pass over the operator: b is the wide
# kernel's tail column (n = 200: columns 128..199 of A on a rest problem,
# b in the tail), so no one-column kernel runs.  FP64 n = 128, 200; FP32
# n = 256, 300.  The adjoint's [M w] product (K = d = 1024) likewise.
@pytest.mark.parametrize("dtype,tol,n", [(torch.float64, 1e-11, 128), (torch.float64, 1e-11, 200),
                                         (torch.float32, 1e-4, 256), (torch.float32, 1e-4, 300)])
def test_joint_sketch_of_a_and_b(rb, orc, dtype, tol, n):
    d, m = 1024, 2112
    A, b = _problem(m, n, 21)
    if dtype == torch.float32:
        A, b = A.astype(np.float32).astype(np.float64), b.astype(np.float32).astype(np.float64)
    S = rb.DenseSkOp(rb.DenseDist(d, m), rb.RNGState(27))
    Sm, _ = orc.fill_dense_unpacked(orc.natural_layout(d, m, "L"), d, m, d, m, 0, 0, [0, 0, 0, 0], [27, 0])
    x_ref = np.linalg.lstsq(Sm @ A, Sm @ b, rcond=None)[0]
    Ad, bd = _dev(_store(A, "C"), dtype), _dev(b, dtype)
    x = torch.full((n,), np.nan, dtype=dtype, device=DEV)
    ran = _ran(rb, lambda: rb.sketch_solve("C", d, m, n, S, 0, 0, Ad, m, bd, x))
    wide = "f32_wide_kernel" if dtype == torch.float32 else "f64_wide_kernel"
    assert wide in ran and "gemv_gen_kernel" not in ran, ran
    assert "solve_householder" not in ran, ran  # well conditioned: CholeskyQR2 accepted
    err = np.linalg.norm(host(x).astype(np.float64) - x_ref) / np.linalg.norm(x_ref)
    assert err <= tol, err
    # the adjoint: [Abar bbar] = S^T [M w] in one pass
    xbar = np.random.default_rng(22).standard_normal(n)
    Abar = torch.full((m * n,), np.nan, dtype=dtype, device=DEV)
    bbar = torch.full((m,), np.nan, dtype=dtype, device=DEV)
    ran = _ran(rb, lambda: rb.sketch_solve_adjoint("C", d, m, n, S, 0, 0, Ad, m, bd, _dev(xbar, dtype), Abar, m,
                                                   bbar))
    assert wide in ran and "gemv_gen_kernel" not in ran, ran
    gA_ref, gb_ref = _torch_grads(Sm, A, b, xbar)
    gA = host(Abar).astype(np.float64).reshape(n, m).T
    assert np.linalg.norm(gA - gA_ref) / np.linalg.norm(gA_ref) <= (1e-9 if dtype == torch.float64 else 1e-4)
    assert np.linalg.norm(host(bbar) - gb_ref) / np.linalg.norm(gb_ref) <= (1e-9 if dtype == torch.float64 else 1e-4)


def _torch_grads(Sm, A, b, xbar):
    """Gradient of <xbar, x(A, b)> by torch's autograd through a host FP64 QR
    solve of the materialised sketched system."""
    St = torch.from_numpy(Sm)
    At = torch.from_numpy(A).requires_grad_(True)
    bt = torch.from_numpy(b).requires_grad_(True)
    Q, R = torch.linalg.qr(St @ At)
    x = torch.linalg.solve_triangular(R, (Q.T @ (St @ bt)).unsqueeze(1), upper=True).squeeze(1)
    torch.sum(x * torch.from_numpy(xbar)).backward()
    return At.grad.numpy(), bt.grad.numpy()


@pytest.mark.parametrize("layout", ["C", "R"])
def test_adjoint_matches_torch_autograd(rb, qr, orc, layout):
    d, m, n, ro, co = 150, 2000, 30, 2, 5
    A, b = _problem(m, n, 8)
    xbar = np.random.default_rng(9).standard_normal(n)
    Dr, Dc = d + ro, m + co
    S = rb.DenseSkOp(rb.DenseDist(Dr, Dc), rb.RNGState(23))
    Sm, _ = orc.fill_dense_unpacked(orc.natural_layout(Dr, Dc, "L"), Dr, Dc, d, m, ro, co, [0, 0, 0, 0], [23, 0])
    gA_ref, gb_ref = _torch_grads(Sm, A, b, xbar)
    ld = m if layout == "C" else n
    Abar = torch.full((m * n,), np.nan, dtype=torch.float64, device=DEV)
    bbar = torch.full((m,), np.nan, dtype=torch.float64, device=DEV)
    rb.sketch_solve_adjoint(layout, d, m, n, S, ro, co, _dev(_store(A, layout)), ld, _dev(b), _dev(xbar), Abar, ld,
                            bbar)
    gA = host(Abar).reshape((n, m) if layout == "C" else (m, n))
    gA = gA.T if layout == "C" else gA
    assert np.linalg.norm(gA - gA_ref) / np.linalg.norm(gA_ref) <= 1e-10
    assert np.linalg.norm(host(bbar) - gb_ref) / np.linalg.norm(gb_ref) <= 1e-10


def test_adjoint_matches_finite_differences(rb, qr):
    m, n, d = 16384, 64, 256
    A, b = _problem(m, n, 10)
    rng = np.random.default_rng(11)
    xbar = rng.standard_normal(n)
    S = rb.DenseSkOp(rb.DenseDist(d, m), rb.RNGState(31))
    Ad, bd, xbd = _dev(_store(A, "C")), _dev(b), _dev(xbar)
    Abar = torch.empty(m * n, dtype=torch.float64, device=DEV)
    bbar = torch.empty(m, dtype=torch.float64, device=DEV)
    rb.sketch_solve_adjoint("C", d, m, n, S, 0, 0, Ad, m, bd, xbd, Abar, m, bbar)
    x = torch.empty(n, dtype=torch.float64, device=DEV)

    def L(Ap, bp):
        rb.sketch_solve("C", d, m, n, S, 0, 0, Ap, m, bp, x)
        return float(torch.dot(x, xbd))

    def D(VA, Vb, h):
        return (L(Ad + h * VA, bd + h * Vb) - L(Ad - h * VA, bd - h * Vb)) / (2 * h)

    dds, rich = [], []
    g = torch.Generator(device="cpu").manual_seed(12)
    for _ in range(6):
        VA = torch.randn(m * n, generator=g, dtype=torch.float64).to(DEV)
        Vb = torch.randn(m, generator=g, dtype=torch.float64).to(DEV)
        dds.append(float(torch.dot(Abar, VA) + torch.dot(bbar, Vb)))
        d1, d2 = D(VA, Vb, 1e-5), D(VA, Vb, 5e-6)
        rich.append((4 * d2 - d1) / 3)
    scale = max(abs(v) for v in dds)
    err = max(abs(f - v) for f, v in zip(rich, dds)) / scale
    print(f"sketch_solve adjoint: FD error / max|directional derivative| {err:.2e}")
    assert err <= 1e-6


def test_fp32_tracks_fp64(rb, qr):
    m, n, d = 20000, 50, 400
    A, b = _problem(m, n, 13)
    A = A.astype(np.float32).astype(np.float64)
    b = b.astype(np.float32).astype(np.float64)
    xbar = np.random.default_rng(14).standard_normal(n)
    S = rb.DenseSkOp(rb.DenseDist(d, m), rb.RNGState(5))
    out = {}
    for dt in (torch.float32, torch.float64):
        x = torch.empty(n, dtype=dt, device=DEV)
        Abar = torch.empty(m * n, dtype=dt, device=DEV)
        bbar = torch.empty(m, dtype=dt, device=DEV)
        rb.sketch_solve("R", d, m, n, S, 0, 0, _dev(A.ravel(), dt), n, _dev(b, dt), x)
        rb.sketch_solve_adjoint("R", d, m, n, S, 0, 0, _dev(A.ravel(), dt), n, _dev(b, dt), _dev(xbar, dt), Abar, n,
                                bbar)
        out[dt] = [host(t).astype(np.float64) for t in (x, Abar, bbar)]
    for a32, a64 in zip(out[torch.float32], out[torch.float64]):
        assert np.linalg.norm(a32 - a64) / np.linalg.norm(a64) <= 1e-4


def test_argument_errors(rb):
    S = rb.DenseSkOp(rb.DenseDist(10, 100), rb.RNGState(1))
    A = torch.zeros(100 * 20, dtype=torch.float64, device=DEV)
    b = torch.zeros(100, dtype=torch.float64, device=DEV)
    x = torch.zeros(20, dtype=torch.float64, device=DEV)
    with pytest.raises(rb.Error):
        rb.sketch_solve("C", 10, 100, 20, S, 0, 0, A, 100, b, x)  # d < n
    with pytest.raises(rb.Error):
        rb.sketch_solve("C", 10, 100, 5, S, 0, 0, A, 50, b, x)  # lda < m
    with pytest.raises(rb.Error):
        rb.sketch_solve("R", 10, 100, 5, S, 0, 0, A, 4, b, x)  # lda < n
    with pytest.raises(rb.Error):
        rb.sketch_solve("C", 10, 100, 5, S, 5, 0, A, 100, b, x)  # submatrix outside S


@pytest.mark.parametrize("cond,path", [(1e2, "solve_cholqr2"), (1e10, "solve_householder")])
def test_factorization_path(rb, orc, monkeypatch, cond, path):
    # a consistent system A x* = b with singular values 1 .. 1/cond: x within
    # cond * eps of numpy's lstsq of the oracle's sketch, and the QR chosen by
    # the conditioning (CholeskyQR2's second Gram matrix is far from I at 1e10)
    monkeypatch.delenv("RBX_SOLVE_QR", raising=False)
    d, m, n = 300, 4000, 40
    rng = np.random.default_rng(3)
    U, _ = np.linalg.qr(rng.standard_normal((m, n)))
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = (U * np.logspace(0, -np.log10(cond), n)) @ V.T
    b = A @ rng.standard_normal(n)
    S = rb.DenseSkOp(rb.DenseDist(d, m), rb.RNGState(29))
    Sm, _ = orc.fill_dense_unpacked(orc.natural_layout(d, m, "L"), d, m, d, m, 0, 0, [0, 0, 0, 0], [29, 0])
    x_ref = np.linalg.lstsq(Sm @ A, Sm @ b, rcond=None)[0]
    x = torch.empty(n, dtype=torch.float64, device=DEV)
    ran = _ran(rb, lambda: rb.sketch_solve("C", d, m, n, S, 0, 0, _dev(_store(A, "C")), m, _dev(b), x))
    assert path in ran, ran
    if path == "solve_cholqr2":
        assert "solve_householder" not in ran, ran  # CholeskyQR2 accepted, no fallback
    err = np.linalg.norm(host(x) - x_ref) / np.linalg.norm(x_ref)
    print(f"cond {cond:.0e} via {path}: {err:.2e}")
    assert err <= (1e-11 if cond < 1e3 else 1e-4), err
