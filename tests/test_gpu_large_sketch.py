"""GPU: the large fused Gaussian sketch at its full size, and the bounds the
BF16 operator's fast generator rests on.

Workload: A 262144 x 1024 FP32, column-major, G diag(10^(-6 j / 1023)) with
G = DenseDist(262144, 1024) key 5 (condition number ~1e6); S =
DenseDist(4096, 262144) Gaussian, key 42, never materialised; B = S A.

  * FP64 fused path (A promoted to FP64) against the oracle's FP64 product on
    512 rows of B in 16 blocks spread over all 4096 rows: <= 1e-13 relative
    Frobenius (the bar for FP64 sketches);
  * BF16-split, three terms, against the FP64 product with the same operator
    (every FP32 entry of S rounded to bfloat16), all 4096 rows: <= 1e-6;
  * the fast generator's per-factor error bounds, re-measured over all 2^32
    input words on this GPU, hold with a factor-of-two margin.
"""
import numpy as np
import pytest
import torch

from conftest import rel_fro

pytestmark = pytest.mark.gpu
DEV = "cuda"
D, M, N = 4096, 262144, 1024


@pytest.fixture(scope="module")
def workload(rb):
    from randblas_amd.datagen import conditioned_gaussian
    A = conditioned_gaussian(M, N, 1e6, 5, DEV)
    S = rb.DenseSkOp(rb.DenseDist(D, M), rb.RNGState(42))
    torch.cuda.synchronize()
    yield A, S
    del A
    torch.cuda.empty_cache()


def _ran(rb, fn):
    rb.profile_collect()
    rb.profile_enable(True)
    fn()
    rb.profile_enable(False)
    return rb.profile_collect()


def test_fp64_fused_full_size_vs_oracle(rb, orc, workload):
    A, S = workload
    A64 = A.double()
    B = torch.empty(D * N, dtype=torch.float64, device=DEV)
    ran = _ran(rb, lambda: rb.lskge3("C", "N", "N", D, N, M, 1.0, S, 0, 0, A64, M, 0.0, B, D))
    assert "f64_wide_kernel" in ran, ran
    got = B.view(N, D).T.cpu().numpy()
    A_h = A64.view(N, M).cpu().numpy().T  # M x N
    del A64
    errs = []
    num = den = 0.0
    for blk in range(16):
        r0 = blk * 256 + (37 * blk) % 224  # 32-row blocks spread over all rows
        Sblk, _ = orc.fill_dense_unpacked("R", D, M, 32, M, r0, 0, [0, 0, 0, 0], [42, 0], dtype=np.float64)
        ref = Sblk @ A_h
        num += np.linalg.norm(got[r0:r0 + 32] - ref) ** 2
        den += np.linalg.norm(ref) ** 2
        errs.append(rel_fro(got[r0:r0 + 32], ref))
    err = float(np.sqrt(num / den))
    print(f"FP64 fused vs oracle, 512 rows: {err:.3e} (worst block {max(errs):.3e})")
    assert err <= 1e-13


def _bf16_round(x):
    """FP32 -> bfloat16 (nearest, ties to even), returned as FP32."""
    u = x.view(torch.int32).to(torch.int64) & 0xFFFFFFFF
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    u = torch.where(u >= 2 ** 31, u - 2 ** 32, u)
    return u.to(torch.int32).view(torch.float32)


def test_bf16_split_full_size_all_rows(rb, workload):
    A, S = workload
    B = torch.empty(D * N, dtype=torch.float32, device=DEV)
    ran = _ran(rb, lambda: rb.lskge3_bf16split("C", "N", "N", D, N, M, 1.0, S, 0, 0, A, M, 0.0, B, D, n_terms=3))
    assert "bf16_skge_kernel" in ran, ran
    got = B.view(N, D).T.double()
    A64 = A.view(N, M).double()  # N x M: row j is column j of A
    num = den = 0.0
    rows = 256
    Sf = torch.empty(rows * M, dtype=torch.float32, device=DEV)
    for r0 in range(0, D, rows):
        rb.fill_dense_unpacked("R", rb.DenseDist(D, M), rows, M, r0, 0, Sf, rb.RNGState(42))
        Sb = _bf16_round(Sf).view(rows, M).double()
        ref = Sb @ A64.T
        num += float(torch.linalg.norm(got[r0:r0 + rows] - ref)) ** 2
        den += float(torch.linalg.norm(ref)) ** 2
        del Sb, ref
    err = (num / den) ** 0.5
    print(f"BF16-split (3 terms) vs FP64 with the same operator, all rows: {err:.3e}")
    assert err <= 1e-6


def test_fast_generator_bounds_exhaustive(rb):
    """The BF16 kernels keep a hardware-transcendental Gaussian only when it
    provably rounds like the portable one, given per-factor bounds; those
    bounds must hold over every input word, with a factor-of-two margin."""
    meas, assumed = rb.bf16_fastgen_bounds()
    es, ec, er, zero_mismatch = meas
    print(f"max |s'-s| {es:.3e}, max |c'-c| {ec:.3e}, max |r'-r|/r {er:.3e}, zero mismatches {zero_mismatch}; "
          f"assumed {assumed[0]:.3e} / {assumed[1]:.3e}")
    assert zero_mismatch == 0
    assert es > 0 and ec > 0 and er > 0  # the sweep measured something
    assert es <= assumed[0] / 2 and ec <= assumed[0] / 2
    assert er <= assumed[1] / 2
