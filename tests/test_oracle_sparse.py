"""CPU: pin the oracle's sparse operators and sparse products.

Pins:
  * SASO structure (test_sparseskop.cc:63-91): vec_nnz distinct in-range
    indices per major vector, +-1 values, compute_next_state;
  * an independent numpy restatement of repeated_fisher_yates
    (sparse_skops.hh:51-106) on small cases, index for index;
  * split-call == one-call (test_discrete.cc:300-328): consecutive blocks of
    vectors from next_state equal the corresponding part of one big sample;
  * LASO: iid index map and duplicate merging (sparse_skops.hh:453-491);
  * left_spmm on the COO operator equals the dense product with the
    operator materialised (+-1 values: exact products).
"""
import numpy as np
import pytest


def fy_numpy(orc, ctr, key, k, dim_major, dim_minor):
    """Partial Fisher-Yates per vector on the identity, written from the
    published algorithm with a dict as the work array."""
    idx = np.zeros(dim_minor * k, np.int64)
    val = np.zeros(dim_minor * k)
    words = orc.philox_stream(ctr, key, dim_minor * k)
    for i in range(dim_minor):
        work = {}
        for j in range(k):
            w = words[i * k + j]
            ell = j + int(w[0]) % (dim_major - j)
            a, b = work.get(j, j), work.get(ell, ell)
            work[j], work[ell] = b, a
            idx[i * k + j] = b
            val[i * k + j] = 1.0 if int(w[1]) % 2 == 0 else -1.0
    return idx, val


@pytest.mark.parametrize("shape,k", [((20, 300), 3), ((300, 20), 5), ((8, 8), 8), ((1, 50), 1), ((64, 2000), 8)])
def test_saso_matches_numpy_restatement(orc, shape, k):
    vals, rows, cols, nnz, nxt = orc.fill_sparse(shape[0], shape[1], k, "S", [5, 0, 0, 0], [7, 0])
    dim_major, dim_minor = min(shape), max(shape)
    assert nnz == k * dim_minor
    idx, val = fy_numpy(orc, [5, 0, 0, 0], [7, 0], k, dim_major, dim_minor)
    major = rows if shape[0] <= shape[1] else cols
    minor = cols if shape[0] <= shape[1] else rows
    assert np.array_equal(major, idx)
    assert np.array_equal(minor, np.repeat(np.arange(dim_minor), k))
    assert np.array_equal(vals, val)
    assert list(nxt) == list(orc.ctr_incr([5, 0, 0, 0], dim_minor * k))
    assert list(nxt) == list(orc.sparse_next_state(shape[0], shape[1], k, "S", [5, 0, 0, 0]))


def test_saso_structure(orc):
    d, m, k = 2000, 10000, 8
    vals, rows, cols, nnz, _ = orc.fill_sparse(d, m, k, "S", [0, 0, 0, 0], [7, 0])
    r = rows.reshape(m, k)
    assert (np.sort(r, axis=1)[:, 1:] != np.sort(r, axis=1)[:, :-1]).all()  # distinct per column
    assert r.min() >= 0 and r.max() < d
    assert set(np.unique(vals)) <= {-1.0, 1.0}
    # each row index is roughly equally likely
    counts = np.bincount(rows, minlength=d)
    assert abs(counts.mean() - k * m / d) < 1e-9 and counts.std() < 3 * np.sqrt(k * m / d)


def test_saso_split_call_equals_one_call(orc):
    d, m1, m2, k = 50, 300, 200, 6
    v, r, c, _, _ = orc.fill_sparse(d, m1 + m2, k, "S", [0, 0, 0, 0], [3, 0])
    v1, r1, c1, _, nxt1 = orc.fill_sparse(d, m1, k, "S", [0, 0, 0, 0], [3, 0])
    v2, r2, c2, _, _ = orc.fill_sparse(d, m2, k, "S", nxt1, [3, 0])
    assert np.array_equal(np.concatenate([r1, r2]), r)
    assert np.array_equal(np.concatenate([c1, c2 + m1]), c)
    assert np.array_equal(np.concatenate([v1, v2]), v)


def test_laso(orc):
    n_rows, n_cols, k = 30, 500, 40  # LASO: rows are long-axis vectors of length 500
    vals, rows, cols, nnz, nxt = orc.fill_sparse(n_rows, n_cols, k, "L", [0, 0, 0, 0], [11, 0])
    assert nnz <= k * n_rows
    assert list(nxt) == list(orc.sparse_next_state(n_rows, n_cols, k, "L", [0, 0, 0, 0]))
    words = orc.philox_stream([0, 0, 0, 0], [11, 0], n_rows * ((k + 1) // 2)).reshape(n_rows, -1)
    L = orc.lib()
    for i in range(n_rows):
        sel = rows == i
        got = dict(zip(cols[sel].tolist(), vals[sel].tolist()))
        w = words[i].reshape(-1)
        exp_cnt, exp_sign = {}, {}
        for s in range(k):
            u = (float(L.orc_uneg11(int(w[2 * s]))) + 1.0) / 2.0
            j = int(float(n_cols) * u)
            sg = 1.0 if L.orc_uneg11(int(w[2 * s + 1])) >= 0 else -1.0
            exp_cnt[j] = exp_cnt.get(j, 0) + 1
            exp_sign.setdefault(j, sg)
        assert set(got) == set(exp_cnt)
        for j in got:
            assert got[j] == np.sqrt(exp_cnt[j]) * exp_sign[j]


def _dense(vals, rows, cols, shape):
    S = np.zeros(shape)
    np.add.at(S, (rows, cols), vals)
    return S


@pytest.mark.parametrize("layout", ["C", "R"])
@pytest.mark.parametrize("opS", ["N", "T"])
@pytest.mark.parametrize("opB", ["N", "T"])
def test_left_spmm_matches_dense(orc, layout, opS, opB):
    rng = np.random.default_rng(4)
    Sr, Sc = 40, 300
    vals, rows, cols, nnz, _ = orc.fill_sparse(Sr, Sc, 4, "S", [0, 0, 0, 0], [1, 0])
    Sd = _dense(vals, rows, cols, (Sr, Sc))
    ro, co = 2, 7
    d, m = (30, 250) if opS == "N" else (250, 30)
    n = 17
    sub = Sd[ro:ro + (d if opS == "N" else m), co:co + (m if opS == "N" else d)]
    opSm = sub if opS == "N" else sub.T
    rB, cB = (m, n) if opB == "N" else (n, m)
    ldb = (rB if layout == "C" else cB) + 1
    ldc = (d if layout == "C" else n) + 2
    B = rng.integers(-8, 8, ldb * max(rB, cB) + 4).astype(np.float64)  # integer data: exact sums
    C = rng.integers(-8, 8, ldc * max(d, n) + 4).astype(np.float64)

    def view(buf, r, c, ld):
        st = buf.itemsize
        return np.lib.stride_tricks.as_strided(buf, (r, c), (st, ld * st) if layout == "C" else (ld * st, st))
    C0 = view(C, d, n, ldc).copy()
    opBm = view(B, rB, cB, ldb) if opB == "N" else view(B, rB, cB, ldb).T
    orc.left_spmm_coo("d", layout, opS, opB, d, n, m, 2.0, (Sr, Sc), vals, rows, cols, ro, co, B, ldb, -1.0, C, ldc)
    assert np.array_equal(view(C, d, n, ldc), 2.0 * opSm @ opBm - C0)


def test_left_spmm_csr_matches_dense(orc):
    rng = np.random.default_rng(5)
    d, m, n = 50, 80, 9
    A = (rng.random((d, m)) < 0.1) * rng.standard_normal((d, m))
    rowptr = np.concatenate([[0], np.cumsum((A != 0).sum(1))]).astype(np.int64)
    colidx = np.nonzero(A)[1].astype(np.int64)
    vals = A[A != 0]
    B = rng.standard_normal(m * n)
    C = np.zeros(d * n)
    orc.left_spmm_csr("d", "C", "N", d, n, m, 1.5, vals, rowptr, colidx, B, m, 0.0, C, d)
    np.testing.assert_allclose(C.reshape(n, d).T, 1.5 * A @ B.reshape(n, m).T, rtol=1e-13, atol=1e-13)


def test_sddmm_oracle_pins(orc):
    """The oracle's SDDMM equals the dense formula alpha*(Bbar @ Sm^T) sampled at
    A's nonzeros, and satisfies the adjoint identity
    sum_k vals_k * dvals_k = alpha * <A @ Sm, Bbar>."""
    rng = np.random.default_rng(2)
    m, n, d = 40, 70, 9
    A = (rng.random((m, n)) < 0.15) * rng.standard_normal((m, n))
    rowptr = np.concatenate([[0], np.cumsum((A != 0).sum(1))]).astype(np.int64)
    col = np.nonzero(A)[1].astype(np.int64)
    vals = A[A != 0]
    Bb = rng.standard_normal((m, d))
    Sm = rng.standard_normal((n, d))
    # Bbar column-major, Sm row-major
    BbF = np.asfortranarray(Bb)
    out = orc.sddmm_csr("d", m, d, 0.75, rowptr, col, BbF.reshape(-1, order="F"), 1, m,
                        np.ascontiguousarray(Sm).reshape(-1), d, 1)
    np.testing.assert_allclose(out, 0.75 * (Bb @ Sm.T)[np.nonzero(A)], rtol=1e-14, atol=1e-14)
    assert abs(np.dot(vals, out) - 0.75 * np.sum((A @ Sm) * Bb)) <= 1e-12 * np.abs(out).sum()
