"""GPU: the BF16-split left sketch (rbx_slskge3_bf16split).

The operator of this form is bf16(S): every FP32 operator entry rounded to
bfloat16 (nearest-even).  Checks:
  * bf16(S) * I == bf16(S) exactly, for every split count (the split terms of
    an identity are exact) -- pins the generator, the rounding and the MFMA
    output placement;
  * against the FP64 product of the same operator (numpy, from the oracle's
    operator): relative Frobenius error <= 1e-6 with three terms, <= 4e-6 with
    two, <= 4e-3 with one;
  * against the FP64 product of the unrounded operator: the bf16 rounding of S
    alone (~2^-9 relative per entry) dominates, ~1e-3;
  * layouts, transposes and submatrix offsets.
"""
import numpy as np
import pytest
import torch

from conftest import rel_fro

pytestmark = pytest.mark.gpu
DEV = "cuda"


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def strided(buf, rows, cols, layout, ld):
    if layout == "C":
        return np.lib.stride_tricks.as_strided(buf, (rows, cols), (buf.itemsize, ld * buf.itemsize))
    return np.lib.stride_tricks.as_strided(buf, (rows, cols), (ld * buf.itemsize, buf.itemsize))


@pytest.mark.parametrize("terms", [1, 2, 3])
def test_identity_gives_rounded_operator(rb, orc, terms):
    d, m = 100, 160
    S = rb.DenseSkOp(rb.DenseDist(d, m), rb.RNGState(42))
    I = torch.eye(m, dtype=torch.float32, device=DEV).reshape(-1)
    B = torch.empty(d * m, dtype=torch.float32, device=DEV)
    rb.lskge3_bf16split("C", "N", "N", d, m, m, 1.0, S, 0, 0, I, m, 0.0, B, d, n_terms=terms)
    got = host(B).reshape(m, d).T
    ref, _ = orc.fill_dense(d, m, [0, 0, 0, 0], [42, 0], dtype=np.float32)
    assert np.array_equal(got, orc.bf16_round(ref))


@pytest.mark.parametrize("family", ["G", "U"])
def test_identity_interior_tiles_exact(rb, orc, family):
    """Interior shapes (d multiple of 64, n of 256, m of 32) take the two-wave
    kernel and, for Gaussians, the hardware-transcendental generator with its
    exact fallback: all 8.4M entries of bf16(S) must equal the oracle's."""
    d, m = 2048, 4096
    S = rb.DenseSkOp(rb.DenseDist(d, m, rb.ScalarDist(family)), rb.RNGState(77))
    I = torch.eye(m, dtype=torch.float32, device=DEV).reshape(-1)
    B = torch.empty(d * m, dtype=torch.float32, device=DEV)
    rb.lskge3_bf16split("C", "N", "N", d, m, m, 1.0, S, 0, 0, I, m, 0.0, B, d, n_terms=3)
    got = host(B).reshape(m, d).T
    ref, _ = orc.fill_dense(d, m, [0, 0, 0, 0], [77, 0], family=family, dtype=np.float32)
    want = orc.bf16_round(ref)
    bad = np.count_nonzero(got != want)
    assert bad == 0, f"{bad} entries differ"


def _case(rb, orc, layout, opS, opA, d, n, m, terms, alpha=1.0, beta=0.0, ro=0, co=0, key=42, seed=0):
    rng = np.random.default_rng(seed)
    Sshape = (d + ro, m + co) if opS == "N" else (m + ro, d + co)
    rA, cA = (m, n) if opA == "N" else (n, m)
    lda = (rA if layout == "C" else cA) + 1
    ldb = (d if layout == "C" else n) + 2
    A = rng.standard_normal(lda * (cA if layout == "C" else rA)).astype(np.float32)
    B0 = rng.standard_normal(ldb * (n if layout == "C" else d)).astype(np.float32)
    S = rb.DenseSkOp(rb.DenseDist(*Sshape), rb.RNGState(key))
    Bd = torch.from_numpy(B0.copy()).to(DEV)
    rb.lskge3_bf16split(layout, opS, opA, d, n, m, alpha, S, ro, co, torch.from_numpy(A).to(DEV), lda, beta, Bd,
                        ldb, n_terms=terms)
    got = strided(host(Bd), d, n, layout, ldb)
    full, _ = orc.fill_dense(Sshape[0], Sshape[1], [0, 0, 0, 0], [key, 0], dtype=np.float32)
    sub = full[ro:ro + (d if opS == "N" else m), co:co + (m if opS == "N" else d)]
    opS_m = (sub if opS == "N" else sub.T).astype(np.float64)
    opS_b = orc.bf16_round(sub if opS == "N" else sub.T).astype(np.float64)
    Am = strided(A, rA, cA, layout, lda).astype(np.float64)
    opA_m = Am if opA == "N" else Am.T
    B0m = strided(B0, d, n, layout, ldb).astype(np.float64)
    same = alpha * opS_b @ opA_m + beta * B0m
    refop = alpha * opS_m @ opA_m + beta * B0m
    return got, same, refop


@pytest.mark.parametrize("terms,tol", [(3, 1e-6), (2, 4e-6), (1, 4e-3)])
def test_error_vs_fp64_same_operator(rb, orc, terms, tol):
    got, same, refop = _case(rb, orc, "C", "N", "N", 130, 70, 3000, terms)
    err = rel_fro(got, same)
    print(f"terms={terms}: rel err vs FP64 (same operator) {err:.3e}, vs reference operator "
          f"{rel_fro(got, refop):.3e}")
    assert err <= tol
    assert 1e-4 < rel_fro(got, refop) < 1e-2


@pytest.mark.parametrize("layout", ["C", "R"])
@pytest.mark.parametrize("opS", ["N", "T"])
@pytest.mark.parametrize("opA", ["N", "T"])
def test_variants(rb, orc, layout, opS, opA):
    got, same, _ = _case(rb, orc, layout, opS, opA, 67, 45, 1111, 3, alpha=0.5, beta=-2.0, ro=3, co=2)
    assert rel_fro(got, same) <= 1e-6


def test_wide_output_and_many_splits(rb, orc):
    # more than one 1024-column block and a K long enough for several splits
    got, same, _ = _case(rb, orc, "C", "N", "N", 70, 1100, 9000, 3, seed=3)
    assert rel_fro(got, same) <= 1e-6


def test_materialised_operator_rejected(rb):
    S = rb.DenseSkOp(rb.DenseDist(8, 16), rb.RNGState(1))
    rb.fill_dense(S, dtype=torch.float32)
    A = torch.zeros(16 * 4, dtype=torch.float32, device=DEV)
    B = torch.zeros(8 * 4, dtype=torch.float32, device=DEV)
    with pytest.raises(rb.Unsupported):
        rb.lskge3_bf16split("C", "N", "N", 8, 4, 16, 1.0, S, 0, 0, A, 16, 0.0, B, 8)


@pytest.mark.parametrize("n", [256, 512, 768])
def test_narrow_tiles_identity_exact(rb, orc, n):
    """Sketch widths that are multiples of 256 but not of 1024 run the
    two-wave kernel with 256- or 512-column tiles: bf16(S) exact through an
    identity of that width."""
    d = 192
    S = rb.DenseSkOp(rb.DenseDist(d, n), rb.RNGState(78))
    I = torch.eye(n, dtype=torch.float32, device=DEV).reshape(-1)
    B = torch.empty(d * n, dtype=torch.float32, device=DEV)
    rb.lskge3_bf16split("C", "N", "N", d, n, n, 1.0, S, 0, 0, I, n, 0.0, B, d, n_terms=3)
    got = host(B).reshape(n, d).T
    ref, _ = orc.fill_dense(d, n, [0, 0, 0, 0], [78, 0], dtype=np.float32)
    assert np.array_equal(got, orc.bf16_round(ref))


@pytest.mark.parametrize("n", [256, 512, 768])
def test_narrow_tiles_error(rb, orc, n):
    got, same, _ = _case(rb, orc, "C", "N", "N", 130, n, 8192, 3, alpha=0.75, beta=0.5, seed=4)
    assert rel_fro(got, same) <= 1e-6
