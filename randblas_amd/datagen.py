"""Synthetic data for tests and benchmarks.

conditioned_gaussian(m, n, cond, key) is the dense workload of the large
fused sketch: A = G diag(cond^(-j/(n-1))) in FP32, column-major, with G the
Gaussian DenseDist(m, n) drawn by this library's generator (key `key`), so
the column scales decay geometrically to condition number ~cond.


power_law_csr(rows, cols, mean_nnz, exponent, seed) draws Poisson(mean_nnz)
column indices per row from a Zipf-law popularity over the columns, so the
column degrees have a power-law tail with the given exponent (rank exponent
1/(exponent - 1)); duplicate entries in a row are merged, values are N(0,1)
float32.  Returns (rowptr int64, colidxs int64, vals float32), zero-based,
column indices sorted within each row.
"""
from __future__ import annotations

import numpy as np


def power_law_csr(rows: int, cols: int, mean_nnz: float, exponent: float, seed: int):
    rng = np.random.default_rng(seed)
    w = (np.arange(cols) + 1.0) ** (-1.0 / (exponent - 1.0))
    cdf = np.cumsum(w) / w.sum()
    deg = rng.poisson(mean_nnz, rows)
    r = np.repeat(np.arange(rows, dtype=np.int64), deg)
    c = np.minimum(np.searchsorted(cdf, rng.random(len(r))), cols - 1).astype(np.int64)
    key = np.unique(r * cols + c)
    r, c = key // cols, key % cols
    rowptr = np.zeros(rows + 1, np.int64)
    np.add.at(rowptr, r + 1, 1)
    rowptr = np.cumsum(rowptr)
    vals = rng.standard_normal(len(c)).astype(np.float32)
    return rowptr, c, vals


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
