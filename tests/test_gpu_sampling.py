"""GPU parity for index sampling and weights_to_cdf (util.hh:459-560):
bit-exact against the oracle's sequential restatement (indices, signs, CDF
values, returned states), including counters that carry across 32-bit
words, and the reference's error behaviour."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda"


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.mark.parametrize("n,k,ctr", [(100, 37, (0, 0, 0, 0)), (7, 1, (5, 0, 0, 0)), (2 ** 40, 1000, (0xFFFFFFF0, 3, 0, 0)),
                                     (123457, 1 << 20, (9, 9, 9, 9)), (5, 0, (0, 0, 0, 0))])
def test_uniform_bit_exact(rb, orc, n, k, ctr):
    st = rb.RNGState(key=42, counter=ctr)
    s = torch.full((max(k, 1),), -1, dtype=torch.int64, device=DEV)
    nxt = rb.sample_indices_iid_uniform(n, k, s, st)
    ref, _, rnxt = orc.sample_indices_iid_uniform(n, k, list(ctr), [42, 0])
    assert np.array_equal(host(s)[:k], ref)
    assert list(nxt.counter) == [int(x) for x in rnxt] and nxt.key == 42


@pytest.mark.parametrize("signed", [False, True])
def test_uniform_unaligned_outputs(rb, orc, signed):
    """Outputs that start 8 (4) bytes past a 16-byte boundary take the per-entry stores."""
    k = 4099
    st = rb.RNGState(key=8, counter=(77, 0, 0, 0))
    sbuf = torch.full((k + 1,), -1, dtype=torch.int64, device=DEV)
    s = sbuf[1:]
    rbuf = torch.zeros(k + 1, dtype=torch.float32, device=DEV)
    r = rbuf[1:] if signed else None
    rb.sample_indices_iid_uniform(10 ** 6, k, s, st, rademachers=r)
    ref, rr, _ = orc.sample_indices_iid_uniform(10 ** 6, k, [77, 0, 0, 0], [8, 0], signed=signed)
    assert np.array_equal(host(s), ref) and host(sbuf)[0] == -1
    if signed:
        assert np.array_equal(host(r).astype(np.float64), rr) and host(rbuf)[0] == 0


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("k", [1, 2, 3, 1001, 65536])
def test_signed_bit_exact(rb, orc, dtype, k):
    st = rb.RNGState(key=3, counter=(0xFFFFFFFF, 0, 0, 0))
    s = torch.empty(k, dtype=torch.int64, device=DEV)
    r = torch.empty(k, dtype=dtype, device=DEV)
    nxt = rb.sample_indices_iid_uniform(2000, k, s, st, rademachers=r)
    ref, rr, rnxt = orc.sample_indices_iid_uniform(2000, k, [0xFFFFFFFF, 0, 0, 0], [3, 0], signed=True)
    assert np.array_equal(host(s), ref) and np.array_equal(host(r).astype(np.float64), rr)
    assert list(nxt.counter) == [int(x) for x in rnxt]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_weights_to_cdf_and_sampling_bit_exact(rb, orc, dtype):
    rng = np.random.default_rng(3)
    n = 5000
    w = (rng.random(n) / (np.arange(n) + 1.0) ** 0.8).astype(dtype)
    w[::11] = 0
    wd = torch.from_numpy(w.copy()).to(DEV)
    rb.weights_to_cdf(n, wd)
    ref, st = orc.weights_to_cdf(w)
    assert st == 0
    cdf = host(wd)
    assert np.array_equal(cdf, ref)
    k = 300001
    s = torch.empty(k, dtype=torch.int64, device=DEV)
    nxt = rb.sample_indices_iid(n, wd, k, s, rb.RNGState(key=77, counter=(1, 2, 3, 4)))
    rs, rnxt = orc.sample_indices_iid(n, ref, k, [1, 2, 3, 4], [77, 0])
    assert np.array_equal(host(s), rs)
    assert list(nxt.counter) == [int(x) for x in rnxt]


def test_degenerate_distribution_on_device(rb):
    # test_discrete.cc:153-165: a delta at 17 with a tiny negative weight clipped
    w = torch.zeros(100, dtype=torch.float32, device=DEV)
    w[17] = 99.0
    w[3] = -float(np.finfo(np.float32).eps) / 10
    rb.weights_to_cdf(100, w)
    s = torch.empty(10000, dtype=torch.int64, device=DEV)
    rb.sample_indices_iid(100, w, 10000, s, rb.RNGState(4))
    assert bool((s == 17).all())


def test_weights_to_cdf_errors_on_device(rb, orc):
    w = torch.tensor([1.0, 2.0, -1e-3, 4.0], dtype=torch.float64, device=DEV)
    with pytest.raises(rb.Error, match="w\\[2\\]"):
        rb.weights_to_cdf(4, w)
    # the reference has overwritten the entries before the failing one
    assert host(w).tolist()[:2] == [1.0, 3.0]
    with pytest.raises(rb.Error, match="sum"):
        rb.weights_to_cdf(5, torch.zeros(5, dtype=torch.float32, device=DEV))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n", [2048, 4097, 3_000_001])
def test_weights_to_cdf_blocks_bit_exact(rb, orc, dtype, n):
    """Sizes at, just past and far beyond the device's 2048-entry blocks, with
    negative-zero, zero and tiny negative weights, against the oracle's
    sequential sums."""
    rng = np.random.default_rng(n)
    w = rng.random(n).astype(dtype) ** 3
    w[::7] = 0
    w[1::13] = -0.0
    w[2::17] = -np.finfo(dtype).eps / 100
    wd = torch.from_numpy(w.copy()).to(DEV)
    rb.weights_to_cdf(n, wd)
    ref, st = orc.weights_to_cdf(w)
    assert st == 0
    assert np.array_equal(host(wd), ref)


def test_weights_to_cdf_error_in_a_later_block(rb, orc):
    """A weight below error_if_below at index 5000 (the third block): the
    entries before it hold the sequential partial sums, the rest are
    untouched, as when the reference throws there."""
    rng = np.random.default_rng(8)
    w = rng.random(7000)
    w[5000] = -1.0
    w[6500] = -2.0
    wd = torch.from_numpy(w.copy()).to(DEV)
    with pytest.raises(rb.Error, match="w\\[5000\\]"):
        rb.weights_to_cdf(7000, wd)
    got = host(wd)
    assert np.array_equal(got[:5000], np.cumsum(w[:5000]))
    assert np.array_equal(got[5000:], w[5000:])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n", [1, 2, 3, 7, 2046, 2047, 2048, 2049, 1048583])
def test_cdf_search_sizes_bit_exact(rb, orc, dtype, n):
    """Sizes around the device's 2047-node probe tree (11 levels in LDS) and
    far beyond it: indices identical to the oracle's lower_bound, including
    draws above the last CDF value (a total below 1)."""
    rng = np.random.default_rng(n + 5)
    cdf = (np.sort(rng.random(n)) * 0.999).astype(dtype)
    k = 40001
    s = torch.empty(k, dtype=torch.int64, device=DEV)
    rb.sample_indices_iid(n, torch.from_numpy(cdf).to(DEV), k, s, rb.RNGState(key=6, counter=(7, 0, 0, 0)))
    ref, _ = orc.sample_indices_iid(n, cdf, k, [7, 0, 0, 0], [6, 0])
    assert np.array_equal(host(s), ref)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_cdf_search_unsorted_array_bit_exact(rb, orc, dtype):
    """An array that is not a CDF (random order, NaNs): the device follows
    lower_bound's probe sequence exactly, as the oracle does."""
    rng = np.random.default_rng(12)
    n = 100003
    a = rng.random(n).astype(dtype)
    a[::97] = np.nan
    k = 20000
    s = torch.empty(k, dtype=torch.int64, device=DEV)
    rb.sample_indices_iid(n, torch.from_numpy(a).to(DEV), k, s, rb.RNGState(key=3))
    ref, _ = orc.sample_indices_iid(n, a, k, [0, 0, 0, 0], [3, 0])
    assert np.array_equal(host(s), ref)


# ---------------------------------------------------------------- power-law sparse test data
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("m,n,k,expo,ctr", [(1000, 500, 20000, 2.1, 0), (37, 3, 500, 1.5, 5), (1, 10, 7, 3.0, 2),
                                            (5000, 20000, 0, 2.1, 0), (300, 100000, 4001, 2.5, 2**32 - 3)])
def test_power_law_csr_bit_exact(rb, orc, dtype, m, n, k, expo, ctr):
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    st = rb.RNGState(17, (ctr, 0, 0, 0))
    A, nxt = rb.power_law_csr(m, n, k, expo, st, dtype=tdt)
    torch.cuda.synchronize()
    rowptr, cols, vals, onxt = orc.power_law_csr(m, n, k, expo, [ctr, 0, 0, 0], [17, 0], dtype=dtype)
    assert A.n_rows == m and A.n_cols == n
    assert np.array_equal(A.rowptr.cpu().numpy(), rowptr)
    assert np.array_equal(A.colidxs.cpu().numpy(), cols)
    assert np.array_equal(A.vals.cpu().numpy(), vals)
    assert list(nxt.counter) == [int(x) for x in onxt]


def test_power_law_csr_sparse_workload(rb):
    """The 200000 x 50000 power-law matrix (4e6 draws, exponent 2.1, key 17):
    structure of the CSR and the column-degree tail."""
    A, _ = rb.power_law_csr(200000, 50000, 4_000_000, 2.1, rb.RNGState(17))
    torch.cuda.synchronize()
    rp = A.rowptr.cpu().numpy()
    c = A.colidxs.cpu().numpy()
    nnz = len(c)
    assert rp[0] == 0 and rp[-1] == nnz and np.all(np.diff(rp) >= 0)
    assert 3_500_000 < nnz <= 4_000_000
    # columns strictly ascending within each row (sorted, repeats merged)
    d = np.diff(c)
    row_start = np.zeros(nnz, bool)
    row_start[rp[:-1][rp[:-1] < nnz]] = True
    assert np.all(d[~row_start[1:]] > 0)
    # popularity of column c ~ (c+1)^(-1/1.1): rank-frequency slope of the top
    # columns near -0.91 (before merging repeats flattens the very top)
    deg = np.bincount(c, minlength=50000).astype(np.float64)
    ranks = np.arange(100, 5000)
    slope = np.polyfit(np.log(ranks + 1.0), np.log(deg[ranks]), 1)[0]
    assert -1.0 < slope < -0.8, slope
