"""GPU: the BF16-split form of rsksp3 (rbx_srsksp3_bf16split) for CSR data.

B = alpha * A_T * op(bf16(submat S)) + beta * B: the operator rounded to
bfloat16 (the operator of the dense BF16-split sketch), A's FP32 values split
into T bfloat16 terms.  Checks:
  * an identity CSR returns bf16(S) exactly (pins the bf16 operator and the
    row kernel's output placement, both layouts, both op(S));
  * against the FP64 product of the same (bf16) operator with A's exact
    values: <= 1e-6 relative Frobenius with three terms, <= 1e-5 with two,
    <= 5e-3 with one;
  * layouts, op(S), alpha / beta, submatrix offsets, a materialised FP32
    operator;
  * the full-size workload: A 200000 x 50000 CSR with power-law column
    degrees (~3.8M stored entries), S = DenseDist(50000, 256) key 42, three
    terms, against the library's FP64 rsksp3 of the bf16-rounded operator.
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


def t_(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def _csr(rng, m, n, mean):
    deg = np.minimum(rng.poisson(mean, m), n)
    rowptr = np.concatenate([[0], np.cumsum(deg)]).astype(np.int64)
    cols = np.concatenate([np.sort(rng.choice(n, size=k, replace=False)) for k in deg]).astype(np.int64)
    vals = rng.standard_normal(len(cols)).astype(np.float32)
    return rowptr, cols, vals


def _dense(rowptr, cols, vals, shape):
    A = np.zeros(shape)
    for i in range(shape[0]):
        A[i, cols[rowptr[i]:rowptr[i + 1]]] = vals[rowptr[i]:rowptr[i + 1]]
    return A


@pytest.mark.parametrize("layout", ["C", "R"])
@pytest.mark.parametrize("opS", ["N", "T"])
def test_identity_gives_rounded_operator(rb, orc, layout, opS):
    n, d = 300, 70
    Sshape = (n, d) if opS == "N" else (d, n)
    S = rb.DenseSkOp(rb.DenseDist(*Sshape), rb.RNGState(21))
    A = rb.CSRMatrix(n, n, t_(np.ones(n, np.float32)), t_(np.arange(n + 1, dtype=np.int64)),
                     t_(np.arange(n, dtype=np.int64)))
    ldb = n if layout == "C" else d
    B = torch.empty(n * d, dtype=torch.float32, device=DEV)
    rb.rsksp3_bf16split(layout, "N", opS, n, d, n, 1.0, A, 0, 0, S, 0, 0, 0.0, B, ldb)
    got = host(B).reshape((d, n) if layout == "C" else (n, d))
    got = got.T if layout == "C" else got
    full, _ = orc.fill_dense(*Sshape, [0, 0, 0, 0], [21, 0], dtype=np.float32)
    want = orc.bf16_round(full if opS == "N" else full.T)
    assert np.array_equal(got, want)


def _case(rb, orc, layout, opS, m, d, n, terms, alpha=1.0, beta=0.0, ro=0, co=0, materialise=False, seed=0):
    rng = np.random.default_rng(seed)
    rowptr, cols, vals = _csr(rng, m, n, 9)
    A = rb.CSRMatrix(m, n, t_(vals), t_(rowptr), t_(cols))
    Sshape = (n + ro, d + co) if opS == "N" else (d + ro, n + co)
    S = rb.DenseSkOp(rb.DenseDist(*Sshape), rb.RNGState(33))
    if materialise:
        rb.fill_dense(S, dtype=torch.float32)
    ldb = (m if layout == "C" else d) + 3
    B0 = rng.standard_normal(ldb * (d if layout == "C" else m)).astype(np.float32)
    Bd = t_(B0.copy())
    rb.rsksp3_bf16split(layout, "N", opS, m, d, n, alpha, A, 0, 0, S, ro, co, beta, Bd, ldb, n_terms=terms)
    got = host(Bd)
    full, _ = orc.fill_dense(*Sshape, [0, 0, 0, 0], [33, 0], dtype=np.float32)
    sub = full[ro:, co:]
    Gb = orc.bf16_round(sub if opS == "N" else sub.T).astype(np.float64)[:n, :d]
    Ad = _dense(rowptr, cols, vals, (m, n))
    if layout == "C":
        gv = np.lib.stride_tricks.as_strided(got, (m, d), (4, ldb * 4))
        B0m = np.lib.stride_tricks.as_strided(B0, (m, d), (4, ldb * 4))
    else:
        gv = np.lib.stride_tricks.as_strided(got, (m, d), (ldb * 4, 4))
        B0m = np.lib.stride_tricks.as_strided(B0, (m, d), (ldb * 4, 4))
    ref = alpha * (Ad @ Gb) + beta * B0m.astype(np.float64)
    return np.array(gv), ref


@pytest.mark.parametrize("terms,tol", [(3, 1e-6), (2, 1e-5), (1, 5e-3)])
def test_error_vs_fp64_same_operator(rb, orc, terms, tol):
    got, ref = _case(rb, orc, "C", "N", 400, 256, 900, terms)
    err = rel_fro(got, ref)
    print(f"terms={terms}: {err:.3e}")
    assert err <= tol


@pytest.mark.parametrize("layout", ["C", "R"])
@pytest.mark.parametrize("opS", ["N", "T"])
@pytest.mark.parametrize("d", [37, 128, 256, 300, 512])
def test_variants(rb, orc, layout, opS, d):
    got, ref = _case(rb, orc, layout, opS, 150, d, 500, 3, alpha=0.75, beta=-0.5, ro=2, co=3, seed=d)
    assert rel_fro(got, ref) <= 1e-6


def test_materialised_fp32_operator(rb, orc):
    got, ref = _case(rb, orc, "R", "N", 120, 64, 400, 3, alpha=1.5, beta=0.25, materialise=True, seed=5)
    assert rel_fro(got, ref) <= 1e-6


def test_unsupported_transposed_data(rb):
    A = rb.CSRMatrix(4, 4, t_(np.ones(4, np.float32)), t_(np.arange(5, dtype=np.int64)),
                     t_(np.arange(4, dtype=np.int64)))
    S = rb.DenseSkOp(rb.DenseDist(4, 4), rb.RNGState(1))
    B = torch.zeros(16, dtype=torch.float32, device=DEV)
    with pytest.raises(rb.Unsupported):
        rb.rsksp3_bf16split("C", "T", "N", 4, 4, 4, 1.0, A, 0, 0, S, 0, 0, 0.0, B, 4)


def _bf16_round_t(x):
    u = x.view(torch.int32).to(torch.int64) & 0xFFFFFFFF
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    u = torch.where(u >= 2 ** 31, u - 2 ** 32, u)
    return u.to(torch.int32).view(torch.float32)


def test_full_size_power_law(rb):
    """A 200000 x 50000 CSR (power-law column degrees, ~20 per row), S =
    DenseDist(50000, 256) key 42, three terms: <= 1e-6 against the FP64
    product of the same bf16 operator (the library's FP64 rsksp3, bit-exact to
    the reference's order, on a materialised FP64 copy of bf16(S))."""
    from randblas_amd.datagen import power_law_csr
    m, n, d = 200000, 50000, 256
    A32 = power_law_csr(m, n, 20, 2.1, 17)
    A64 = rb.CSRMatrix(m, n, A32.vals.double(), A32.rowptr, A32.colidxs)
    S = rb.DenseSkOp(rb.DenseDist(n, d), rb.RNGState(42))
    B = torch.empty(m * d, dtype=torch.float32, device=DEV)
    rb.rsksp3_bf16split("C", "N", "N", m, d, n, 1.0, A32, 0, 0, S, 0, 0, 0.0, B, m)
    Sf = torch.empty(n * d, dtype=torch.float32, device=DEV)
    rb.fill_dense(rb.DenseDist(n, d), Sf, rb.RNGState(42))  # natural ColMajor
    S64 = rb.DenseSkOp(rb.DenseDist(n, d), rb.RNGState(42), buff=_bf16_round_t(Sf).double(), layout="C")
    B64 = torch.empty(m * d, dtype=torch.float64, device=DEV)
    rb.rsksp3("C", "N", "N", m, d, n, 1.0, A64, 0, 0, S64, 0, 0, 0.0, B64, m)
    err = float(torch.linalg.norm(B.double() - B64) / torch.linalg.norm(B64))
    print(f"C5 BF16-split (3 terms) vs FP64 with the same operator: {err:.3e}")
    assert err <= 1e-6
