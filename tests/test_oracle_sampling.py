"""CPU: pin the oracle's index samplers and weights_to_cdf (util.hh:459-560).

Pins:
  * an independent closed form from the (KAT-pinned) Philox words: sample i
    of the unsigned sampler is floor(n * (uneg11(w) + 1) / 2) for word i mod 4
    of counter ctr + i // 4; the signed one uses words 2(i mod 2), 2(i mod 2)+1
    of counter ctr + i // 2; the CDF sampler is numpy's searchsorted(side =
    "left") (= std::lower_bound) on u = (uneg11(w) + 1) / 2 computed in T;
  * the oracle's LASO sampler draws the same indices and signs;
  * the returned state: ceil(k / 4) (signed: ceil(k / 2)) counters, and a
    split call equals one call when the first part ends on a counter boundary
    (test_discrete.cc:300-328 for the Fisher-Yates analogue);
  * weights_to_cdf: numpy's sequential cumsum in T, scaled by 1 / sum; the
    error cases; the degenerate distributions of test_discrete.cc:137-166;
  * Kolmogorov-Smirnov against the uniform and a power-law CDF
    (test_discrete.cc:100-135).
"""
import numpy as np
import pytest
from scipy import stats


def _words(orc, ctr, key, ncounters):
    return orc.philox_stream(ctr, key, ncounters).reshape(-1)


def _uneg11(words):
    # (float)(int32)w * 2^-31 + 2^-32 in float32 (the product is exact)
    return (words.view(np.int32).astype(np.float32) * np.float32(2.0 ** -31) + np.float32(2.0 ** -32)).astype(
        np.float32)


@pytest.mark.parametrize("n,k,ctr", [(100, 37, [0, 0, 0, 0]), (7, 1, [5, 0, 0, 0]), (2 ** 40, 64, [0xFFFFFFFE, 3, 0, 0]),
                                     (1, 9, [1, 2, 3, 4])])
def test_uniform_closed_form(orc, n, k, ctr):
    s, _, nxt = orc.sample_indices_iid_uniform(n, k, ctr, [42, 0])
    w = _words(orc, ctr, [42, 0], (k + 3) // 4)[:k]
    u = (_uneg11(w).astype(np.float64) + 1.0) / 2.0
    ref = (float(n) * u).astype(np.int64)
    assert np.array_equal(s, ref)
    assert np.all((s >= 0) & (s < n))
    assert list(nxt) == list(orc.ctr_incr(ctr, (k + 3) // 4))


@pytest.mark.parametrize("n,k", [(1000, 33), (3, 2), (50, 64)])
def test_signed_closed_form_and_laso(orc, n, k):
    ctr = [9, 0, 0, 0]
    s, r, nxt = orc.sample_indices_iid_uniform(n, k, ctr, [3, 0], signed=True)
    w = _words(orc, ctr, [3, 0], (k + 1) // 2)[:2 * k]
    u = (_uneg11(w[0::2]).astype(np.float64) + 1.0) / 2.0
    assert np.array_equal(s, (float(n) * u).astype(np.int64))
    assert np.array_equal(r, np.where(_uneg11(w[1::2]) >= 0, 1.0, -1.0))
    assert list(nxt) == list(orc.ctr_incr(ctr, (k + 1) // 2))


def test_signed_sampler_is_the_laso_sampler(orc):
    # one long-axis vector of a LASO operator: vec_nnz draws from
    # sample_indices_iid_uniform<signed>, merged; pick a seed with no repeats
    n, k = 5000, 6
    vals, rows, cols, nnz, _ = orc.fill_sparse(n, 1, k, "L", [0, 0, 0, 0], [11, 0])
    s, r, _ = orc.sample_indices_iid_uniform(n, k, [0, 0, 0, 0], [11, 0], signed=True)
    assert len(set(s.tolist())) == k and nnz == k
    assert np.array_equal(np.asarray(rows)[:k], s)   # a tall operator: rows are the major indices
    assert np.array_equal(np.asarray(vals)[:k], r)


def test_split_call_equals_one_call(orc):
    ctr, key = [100, 0, 0, 0], [5, 0]
    s_all, _, nxt_all = orc.sample_indices_iid_uniform(77, 40, ctr, key)
    s1, _, nxt1 = orc.sample_indices_iid_uniform(77, 24, ctr, key)
    s2, _, nxt2 = orc.sample_indices_iid_uniform(77, 16, nxt1, key)
    assert np.array_equal(np.concatenate([s1, s2]), s_all) and list(nxt2) == list(nxt_all)
    # a first part that ends inside a counter abandons that counter's other words
    t1, _, m1 = orc.sample_indices_iid_uniform(77, 5, ctr, key)
    assert list(m1) == list(orc.ctr_incr(ctr, 2))


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_weights_to_cdf_closed_form(orc, dt):
    rng = np.random.default_rng(1)
    w = rng.random(1000).astype(dt)
    w[::7] = 0
    cdf, st = orc.weights_to_cdf(w)
    assert st == 0
    cs = np.cumsum(w, dtype=dt)            # sequential in dt
    ref = cs * (dt(1.0) / cs[-1])
    assert np.array_equal(cdf, ref.astype(dt))
    assert cdf[-1] <= 1.0 + np.finfo(dt).eps and np.all(np.diff(cdf) >= 0)


def test_weights_to_cdf_errors(orc):
    w = np.array([1.0, 2.0, -1e-3, 4.0], np.float32)
    cdf, st = orc.weights_to_cdf(w)
    assert st == 1 + 2 and list(cdf[:2]) == [1.0, 3.0] and cdf[2] == np.float32(-1e-3)
    z, st = orc.weights_to_cdf(np.zeros(9))
    assert st == -1
    # a negative weight within the default tolerance is clipped
    c, st = orc.weights_to_cdf(np.array([0.0, -1e-9, 2.0]))
    assert st == 0 and list(c) == [0.0, 0.0, 1.0]


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_cdf_sampler_closed_form(orc, dt):
    rng = np.random.default_rng(2)
    w = (1.0 / (np.arange(300) + 1.0) ** 1.5).astype(dt)
    cdf, _ = orc.weights_to_cdf(w)
    k = 203
    s, nxt = orc.sample_indices_iid(300, cdf, k, [7, 0, 0, 0], [13, 0])
    words = _words(orc, [7, 0, 0, 0], [13, 0], (k + 3) // 4)[:k]
    u = ((_uneg11(words).astype(dt) + dt(1.0)) / dt(2.0)).astype(dt)
    assert np.array_equal(s, np.searchsorted(cdf, u, side="left"))
    assert list(nxt) == list(orc.ctr_incr([7, 0, 0, 0], (k + 3) // 4))


def test_degenerate_distributions(orc):
    # test_discrete.cc:137-166
    N = 100
    w = np.zeros(N, np.float32)
    w[0::2] = 1.0 / (np.arange(0, N, 2, dtype=np.float32) + 1.0)
    w[10] = 0.0
    cdf, st = orc.weights_to_cdf(w)
    assert st == 0
    s, _ = orc.sample_indices_iid(N, cdf, N * N, [0, 0, 0, 0], [4, 0])
    assert not np.any((s == 10) | (s % 2 == 1))
    w = np.zeros(N, np.float32)
    w[17] = 99.0
    w[3] = -np.finfo(np.float32).eps / 10
    cdf, st = orc.weights_to_cdf(w)
    assert st == 0 and cdf[17] >= 0
    s, _ = orc.sample_indices_iid(N, cdf, N * N, [0, 0, 0, 0], [4, 0])
    assert np.all(s == 17)


def test_kolmogorov_smirnov(orc):
    n, k = 1000, 200000
    s, _, _ = orc.sample_indices_iid_uniform(n, k, [0, 0, 0, 0], [21, 0])
    assert stats.kstest(s + 0.5, stats.uniform(loc=0, scale=n).cdf).pvalue > 1e-3
    w = (1.0 / (np.arange(n) + 1.0) ** 1.2).astype(np.float64)
    cdf, _ = orc.weights_to_cdf(w)
    t, _ = orc.sample_indices_iid(n, cdf, k, [0, 0, 0, 0], [22, 0])
    emp = np.cumsum(np.bincount(t, minlength=n)) / k
    crit = 1.95 / np.sqrt(k)   # 0.1% two-sided KS critical value
    assert np.max(np.abs(emp - cdf)) < crit


def test_power_law_csr_structure_and_streams(orc):
    """The power-law test data: a CSR with columns strictly ascending per row,
    counters consumed as the three samplers state (k/4 for the columns, k/4 for
    the rows, nnz/4 for the values), and values equal to the Box-Muller pairs
    of the Philox words after the index draws (entry i: pair (i % 4) // 2 of
    counter c2 + i // 4)."""
    m, n, k = 400, 2000, 9000
    rp, c, v, nxt = orc.power_law_csr(m, n, k, 2.1, [11, 0, 0, 0], [17, 0], dtype=np.float32)
    nnz = len(c)
    assert rp[0] == 0 and rp[-1] == nnz and np.all(np.diff(rp) >= 0)
    for i in range(m):
        seg = c[rp[i]:rp[i + 1]]
        assert np.all(np.diff(seg) > 0)
    assert 0 < nnz <= k
    c2 = 11 + 2 * ((k + 3) // 4)
    assert list(nxt) == [c2 + (nnz + 3) // 4, 0, 0, 0]
    words = orc.philox_stream([c2, 0, 0, 0], [17, 0], (nnz + 3) // 4).reshape(-1)
    g = orc.boxmuller(words)  # two Gaussians per pair of words, in word order
    assert np.array_equal(v, g[:nnz].astype(np.float32))
    # the column draws follow the normalised Zipf weights (about one draw per
    # row here, so merging repeats removes almost nothing)
    m, k = 40000, 20000
    rp, c, v, nxt = orc.power_law_csr(m, n, k, 2.1, [0, 0, 0, 0], [17, 0], dtype=np.float64)
    w = (np.arange(n) + 1.0) ** (-1 / 1.1)
    freq = np.bincount(c, minlength=n)
    assert k - 200 < len(c) <= k
    assert freq[0] > freq[10] > freq[1000]
    assert abs(freq[:10].sum() / len(c) - w[:10].sum() / w.sum()) < 0.02
