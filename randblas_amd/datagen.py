"""Synthetic data for tests and benchmarks.

conditioned_gaussian(m, n, cond, key) is the dense workload of the large
fused sketch: A = G diag(cond^(-j/(n-1))) in FP32, column-major, with G the
Gaussian DenseDist(m, n) drawn by this library's generator (key `key`), so
the column scales decay geometrically to condition number ~cond.


power_law_csr(rows, cols, mean_nnz, exponent, key) is the sparse data matrix
of the sparse-data workloads, drawn on the device from Philox key `key`:
column indices from a Zipf-law popularity over the columns (column degrees
with a power-law tail of the given exponent), rows uniform, repeated entries
kept once, N(0,1) values; a CSRMatrix over device tensors, column indices
ascending within each row.
"""
from __future__ import annotations

import numpy as np


def power_law_csr(rows: int, cols: int, mean_nnz: float, exponent: float, key: int, dtype=None):
    """The power-law CSR matrix of the sparse-data workloads, generated on the
    device by rbx_[sd]power_law_csr: rows * mean_nnz (row, column) draws --
    columns from the Zipf weights (c+1)^(-1/(exponent-1)), rows uniform --
    repeated pairs kept once, N(0,1) values; all from Philox key `key`.
    Returns a CSRMatrix over device tensors (FP32 values unless dtype says
    otherwise)."""
    import randblas_amd as rb
    A, _ = rb.power_law_csr(rows, cols, int(round(rows * mean_nnz)), exponent, rb.RNGState(key), dtype)
    return A


def conditioned_gaussian(m: int, n: int, cond: float, key: int, device="cuda"):
    """A (m x n, FP32, column-major, a flat device tensor of m*n entries) =
    G * diag(cond^(-j / (n - 1))), G = DenseDist(m, n) Gaussian with key `key`
    (natural layout ColMajor for m >= n), the scales rounded to FP32."""
    import torch
    import randblas_amd as rb
    A = torch.empty(m * n, dtype=torch.float32, device=device)
    rb.fill_dense(rb.DenseDist(m, n), A, rb.RNGState(key))
    j = torch.arange(n, device=device, dtype=torch.float64)
    scales = torch.pow(torch.tensor(float(cond), dtype=torch.float64, device=device), -j / max(n - 1, 1)).float()
    A.view(n, m).mul_(scales.view(n, 1))
    return A
