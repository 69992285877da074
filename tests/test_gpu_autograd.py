"""GPU: adjoint kernels (S^T * G with S regenerated from Philox) and the
differentiable sketch-and-solve least-squares problem.

  * dot-product identity <op(S) A, G> = <A, op(S)^T G> (left) and
    <A op(S), G> = <A, G op(S)^T> (right), FP64, over layouts, transposes and
    submatrix offsets -- what the mathematics fixes for an adjoint;
  * the adjoint against the oracle's materialised operator (FP64 <= 1e-13,
    FP32 <= 1e-5);
  * sketch-and-solve at the full size of the differentiable workload: A
    65536 x 256 (iid N(0,1), key 11), x* (key 12), b = A x* + 1e-3 noise
    (key 13), S 1024 x 65536 Gaussian (key 42), x = argmin ||S(Ax - b)|| by
    FP64 QR of the 1024 x 256 sketch on the host, loss 0.5 ||x - x*||^2.
    Gradients with respect to A and b flow back through the adjoint kernel and
    are checked against FP64 central finite differences along 8 random
    directions -- Richardson-extrapolated from eps = 1e-5 and 5e-6 (which
    removes the O(eps^2) truncation term, ~2e-6 relative at eps = 1e-5), and
    plain at eps = 1e-6 -- with the error measured against the largest of the
    8 directional derivatives (one random direction can be nearly orthogonal
    to the gradient): <= 1e-6.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda"


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.mark.parametrize("layout", ["C", "R"])
@pytest.mark.parametrize("opS", ["N", "T"])
@pytest.mark.parametrize("side", ["left", "right"])
def test_adjoint_dot_product_identity(rb, layout, opS, side):
    rng = np.random.default_rng(3)
    d, m, n = 70, 530, 23
    ro, co = 5, 11
    Sshape = (d + 20, m + 20) if opS == "N" else (m + 20, d + 20)
    S = rb.DenseSkOp(rb.DenseDist(*Sshape), rb.RNGState(21))
    if side == "left":
        # B (d x n) = op(S) (d x m) * A (m x n)
        A = torch.from_numpy(rng.standard_normal(m * n)).to(DEV)
        G = torch.from_numpy(rng.standard_normal(d * n)).to(DEV)
        lda, ldg = (m, d) if layout == "C" else (n, n)
        B = torch.zeros(d * n, dtype=torch.float64, device=DEV)
        Abar = torch.zeros(m * n, dtype=torch.float64, device=DEV)
        rb.lskge3(layout, opS, "N", d, n, m, 1.0, S, ro, co, A, lda, 0.0, B, ldg)
        rb.lskge3_adjoint(layout, opS, d, n, m, 1.0, S, ro, co, G, ldg, 0.0, Abar, lda)
    else:
        # B (m x d) = A (m x n) * op(S) (n x d);  here n plays the inner role
        mm, nn, dd = n, m, d
        A = torch.from_numpy(rng.standard_normal(mm * nn)).to(DEV)
        G = torch.from_numpy(rng.standard_normal(mm * dd)).to(DEV)
        lda, ldg = (mm, mm) if layout == "C" else (nn, dd)
        Sshape_r = (nn + 20, dd + 20) if opS == "N" else (dd + 20, nn + 20)
        S = rb.DenseSkOp(rb.DenseDist(*Sshape_r), rb.RNGState(21))
        B = torch.zeros(mm * dd, dtype=torch.float64, device=DEV)
        Abar = torch.zeros(mm * nn, dtype=torch.float64, device=DEV)
        rb.rskge3(layout, "N", opS, mm, dd, nn, 1.0, A, lda, S, ro, co, 0.0, B, ldg)
        rb.rskge3_adjoint(layout, opS, mm, dd, nn, 1.0, G, ldg, S, ro, co, 0.0, Abar, lda)
    lhs = float(np.dot(host(B), host(G)))
    rhs = float(np.dot(host(A), host(Abar)))
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0), (lhs, rhs)


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-13), (torch.float32, 1e-5)])
def test_adjoint_against_oracle(rb, orc, dtype, tol):
    from conftest import rel_fro
    rng = np.random.default_rng(4)
    d, m, n = 256, 4000, 31
    S = rb.DenseSkOp(rb.DenseDist(d, m), rb.RNGState(42))
    G = rng.standard_normal(d * n)
    Gd = torch.from_numpy(G).to(DEV, dtype)
    Abar = torch.zeros(m * n, dtype=dtype, device=DEV)
    rb.lskge3_adjoint("C", "N", d, n, m, 1.0, S, 0, 0, Gd, d, 0.0, Abar, m)
    Sm, _ = orc.fill_dense(d, m, [0, 0, 0, 0], [42, 0],
                           dtype=np.float64 if dtype == torch.float64 else np.float32)
    ref = Sm.astype(np.float64).T @ host(Gd).astype(np.float64).reshape(n, d).T
    assert rel_fro(host(Abar).reshape(n, m).T, ref) <= tol


def _problem(rb):
    m, n, d = 65536, 256, 1024
    A32 = torch.empty(m * n, dtype=torch.float32, device=DEV)
    rb.fill_dense(rb.DenseDist(m, n), A32, rb.RNGState(11))  # natural ColMajor: column j at j*m
    xs = torch.empty(n, dtype=torch.float32, device=DEV)
    rb.fill_dense(rb.DenseDist(n, 1), xs, rb.RNGState(12))
    noise = torch.empty(m, dtype=torch.float32, device=DEV)
    rb.fill_dense(rb.DenseDist(m, 1), noise, rb.RNGState(13))
    A = torch.from_numpy(host(A32).astype(np.float64).reshape(n, m).T.copy())  # (m, n) on the host
    x_star = torch.from_numpy(host(xs).astype(np.float64))
    b = A @ x_star + 1e-3 * torch.from_numpy(host(noise).astype(np.float64))
    S = rb.DenseSkOp(rb.DenseDist(d, m), rb.RNGState(42))
    return A.to(DEV), b.to(DEV), x_star, S


def _loss(A, b, x_star, S):
    from randblas_amd.autograd import left_sketch
    Y = left_sketch(A, S).cpu()          # (d, n) sketch, then FP64 QR on the host
    z = left_sketch(b, S).cpu()
    Q, R = torch.linalg.qr(Y)
    x = torch.linalg.solve_triangular(R, (Q.T @ z).unsqueeze(1), upper=True).squeeze(1)
    return 0.5 * torch.sum((x - x_star) ** 2)


def test_sketch_and_solve_gradients_match_finite_differences(rb):
    A, b, x_star, S = _problem(rb)
    Ar = A.clone().requires_grad_(True)
    br = b.clone().requires_grad_(True)
    L = _loss(Ar, br, x_star, S)
    L.backward()
    gA, gb = Ar.grad, br.grad
    assert gA.shape == A.shape and gb.shape == b.shape
    assert float(L) < 1.0  # the sketched solution is close to x*
    g = torch.Generator(device="cpu").manual_seed(5)

    def D(VA, Vb, h):  # central difference
        return (float(_loss(A + h * VA, b + h * Vb, x_star, S)) -
                float(_loss(A - h * VA, b - h * Vb, x_star, S))) / (2 * h)

    dds, rich, plain = [], [], []
    with torch.no_grad():
        for k in range(8):
            VA = torch.randn(A.shape, generator=g, dtype=torch.float64).to(DEV)
            Vb = torch.randn(b.shape, generator=g, dtype=torch.float64).to(DEV)
            dds.append(float(torch.sum(gA * VA) + torch.sum(gb * Vb)))
            d1, d2 = D(VA, Vb, 1e-5), D(VA, Vb, 5e-6)
            rich.append((4 * d2 - d1) / 3)  # Richardson: removes the O(h^2) term
            plain.append(D(VA, Vb, 1e-6))
    scale = max(abs(x) for x in dds)
    err_rich = max(abs(f - d) for f, d in zip(rich, dds)) / scale
    err_plain = max(abs(f - d) for f, d in zip(plain, dds)) / scale
    print(f"sketch-and-solve: FD error / max|directional derivative|: Richardson {err_rich:.2e}, "
          f"eps 1e-6 {err_plain:.2e}")
    assert err_rich <= 1e-6
    assert err_plain <= 1e-6


def test_fp32_forward_and_adjoint_track_fp64(rb):
    from conftest import rel_fro
    from randblas_amd.autograd import left_sketch
    A, b, x_star, S = _problem(rb)
    A32 = A.float().requires_grad_(True)
    A64 = A.clone().requires_grad_(True)
    Y32 = left_sketch(A32, S)
    Y64 = left_sketch(A64, S)
    assert rel_fro(host(Y32.detach()), host(Y64.detach())) <= 1e-5
    G = torch.randn(Y64.shape, dtype=torch.float64, generator=torch.Generator().manual_seed(1)).to(DEV)
    Y32.backward(G.float())
    Y64.backward(G)
    assert rel_fro(host(A32.grad), host(A64.grad)) <= 1e-5
