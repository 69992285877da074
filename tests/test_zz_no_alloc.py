"""The allocation-free lskges (rbx_[sd]lskges_noalloc) against rbx_[sd]lskges.

The entry point keeps its workspace in a fixed per-GPU device array that
every call reuses, so each test makes two calls in a row, the second with a
smaller operator window, and checks both against the allocating entry point
bit for bit (the two share their kernels).
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _case(rb, dtype, Sshape, vec_nnz, axis, key, layout, opS, opA, d, n, m, ro_s, co_s, alpha, beta, seed):
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    g = torch.Generator(device="cpu").manual_seed(seed)
    S = rb.SparseSkOp(rb.SparseDist(*Sshape, vec_nnz, rb.Axis(axis)), rb.RNGState(key), dtype=tdt)
    rb.fill_sparse(S)
    rows_A, cols_A = (m, n) if opA == "N" else (n, m)
    lda = rows_A if layout == "C" else cols_A
    ldb = d if layout == "C" else n
    A = (torch.rand(rows_A * cols_A, generator=g, dtype=torch.float64) * 2 - 1).to(tdt).to(DEV)
    B0 = (torch.rand(d * n, generator=g, dtype=torch.float64) * 2 - 1).to(tdt).to(DEV)
    ref = B0.clone()
    got = B0.clone()
    rb.lskges(layout, opS, opA, d, n, m, alpha, S, ro_s, co_s, A, lda, beta, ref, ldb)
    rb.lskges_noalloc(layout, opS, opA, d, n, m, alpha, S, ro_s, co_s, A, lda, beta, got, ldb)
    torch.cuda.synchronize()
    return got.cpu(), ref.cpu()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_noalloc_saso_workload_then_smaller(rb, dtype):
    print(f"noalloc: SASO 2000x100000 vec_nnz 8, A 100000x500, {dtype.__name__}", flush=True)
    got, ref = _case(rb, dtype, (2000, 100000), 8, "S", 7, "C", "N", "N", 2000, 500, 100000, 0, 0, 1.0, 0.0, 1)
    assert torch.equal(got, ref)
    print("noalloc: smaller window 300x20000 (submatrix), beta = -0.5", flush=True)
    got, ref = _case(rb, dtype, (310, 20500), 4, "S", 11, "C", "N", "N", 300, 64, 20000, 7, 300, 0.75, -0.5, 2)
    assert torch.equal(got, ref)


@pytest.mark.parametrize("axis", ["S", "L"])
def test_noalloc_row_major_transposed_then_smaller(rb, axis):
    # row-major A with opA = 'T' has unit stride along the contraction
    got, ref = _case(rb, np.float64, (1500, 60000), 6, axis, 21, "R", "N", "T", 1500, 96, 60000, 0, 0, 1.0, 0.0, 3)
    assert torch.equal(got, ref)
    got, ref = _case(rb, np.float64, (200, 5000), 3, axis, 22, "R", "N", "T", 180, 33, 4000, 5, 9, 2.0, 1.0, 4)
    assert torch.equal(got, ref)


def test_noalloc_transposed_operator(rb):
    # opS = 'T': submat(S) is m x d
    got, ref = _case(rb, np.float32, (40000, 900), 5, "S", 31, "C", "T", "N", 800, 40, 39000, 100, 50, 1.0, 0.0, 5)
    assert torch.equal(got, ref)
    got, ref = _case(rb, np.float32, (9000, 120), 2, "S", 32, "C", "T", "N", 100, 17, 8000, 0, 3, 1.0, 0.0, 6)
    assert torch.equal(got, ref)


def test_noalloc_rejects_beyond_its_scratch(rb):
    cap = rb.lskges_noalloc_scratch_bytes()
    assert cap == 64 << 20
    # 1.6e6 stored entries: the FP64 CSR of the window needs ~45 MB
    S = rb.SparseSkOp(rb.SparseDist(1000, 200000, 8), rb.RNGState(5), dtype=torch.float64)
    rb.fill_sparse(S)
    S.release()  # without the operator's prepared copy, the call sorts its entries in the scratch
    A = torch.ones(200000 * 4, dtype=torch.float64, device=DEV)
    B = torch.full((1000 * 4,), 3.0, dtype=torch.float64, device=DEV)
    with pytest.raises(rb.Error, match="scratch"):
        rb.lskges_noalloc("C", "N", "N", 1000, 4, 200000, 1.0, S, 0, 0, A, 200000, 0.0, B, 1000)
    torch.cuda.synchronize()
    assert torch.all(B == 3.0)  # nothing launched
    # with the prepared copy the call needs no scratch, and agrees with lskges
    S.prepare()
    rb.lskges_noalloc("C", "N", "N", 1000, 4, 200000, 1.0, S, 0, 0, A, 200000, 0.0, B, 1000)
    B2 = torch.zeros_like(B)
    S.release()
    rb.lskges("C", "N", "N", 1000, 4, 200000, 1.0, S, 0, 0, A, 200000, 0.0, B2, 1000)
    assert torch.equal(B, B2)
    # a shape the fixed-workspace path does not take
    Bm = torch.zeros(1000 * 4, dtype=torch.float64, device=DEV)
    Sm = rb.SparseSkOp(rb.SparseDist(1000, 5000, 2), rb.RNGState(6), dtype=torch.float64)
    rb.fill_sparse(Sm)
    Am = torch.ones(5000 * 4, dtype=torch.float64, device=DEV)
    with pytest.raises(rb.Unsupported):
        rb.lskges_noalloc("R", "N", "N", 1000, 4, 5000, 1.0, Sm, 0, 0, Am, 4, 0.0, Bm, 4)
    # and the allocating path still agrees after the rejected calls
    got, ref = _case(rb, np.float64, (120, 3000), 4, "S", 8, "C", "N", "N", 120, 10, 3000, 0, 0, 1.0, 0.0, 7)
    assert torch.equal(got, ref)


def test_noalloc_alpha_zero_scales_only(rb):
    got, ref = _case(rb, np.float64, (100, 2000), 4, "S", 9, "C", "N", "N", 100, 12, 2000, 0, 0, 0.0, 2.0, 8)
    assert torch.equal(got, ref)
