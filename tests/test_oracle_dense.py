"""CPU: pin the oracle's dense-operator sampling and dense sketches.

Pins (reference file:line of the invariant restated):
  * entry (i, j) = word (b mod 4) of Philox(counter + a*P + b/4), with the
    natural-layout mapping of dense_skops.hh:96-199 -- checked against an
    independent numpy evaluation of that closed form;
  * submatrix == slice of the full matrix (test_denseskop.cc:162-296);
  * next_state chaining / concatenation (test_denseskop.cc:405-489);
  * wide vs tall autotranspose (test_denseskop.cc:344-403);
  * moments (test_denseskop.cc:97-159);
  * S * I == S (linop_common.hh:308-358) and sketches against numpy GEMM on
    the materialised operator with the componentwise bound of
    linop_common.hh:260-266.
"""
import numpy as np
import pytest


def closed_form(orc, D_rows, D_cols, major_axis, family, counter, key):
    """Independent numpy restatement of the counter -> entry map."""
    lng = major_axis == "L"
    dim_major = max(D_rows, D_cols) if lng else min(D_rows, D_cols)
    dim_minor = min(D_rows, D_cols) if lng else max(D_rows, D_cols)
    P = -(-dim_major // 4)
    words = orc.philox_stream(counter, key, dim_minor * P).reshape(dim_minor, P, 4)
    if family == "G":
        vals = orc.boxmuller(words.reshape(-1), 0).reshape(dim_minor, P * 4)
    else:
        w = words.reshape(-1).astype(np.int64)
        s = np.where(w >= 2**31, w - 2**32, w).astype(np.float32)
        vals = (s * np.float32(2.0**-31) + np.float32(2.0**-32)).astype(np.float32).reshape(dim_minor, P * 4)
    nat = vals[:, :dim_major]  # row a of the natural-layout matrix
    layout = orc.natural_layout(D_rows, D_cols, major_axis)
    return nat if layout == "R" else nat.T


@pytest.mark.parametrize("shape,axis,fam", [((7, 13), "L", "G"), ((13, 7), "L", "G"), ((9, 21), "S", "G"),
                                             ((21, 9), "S", "U"), ((5, 5), "L", "G"), ((3, 40), "L", "U"),
                                             ((1, 9), "S", "G"), ((12, 1), "L", "G")])
def test_fill_matches_closed_form(orc, shape, axis, fam):
    ctr, key = [3, 0, 0, 0], [42, 7]
    ref = closed_form(orc, shape[0], shape[1], axis, fam, ctr, key)
    mat, nxt = orc.fill_dense(shape[0], shape[1], ctr, key, family=fam, major_axis=axis, dtype=np.float32)
    if fam == "U":
        ref = (ref * np.float32(np.sqrt(3.0))).astype(np.float32)
    assert np.array_equal(mat, ref)
    assert np.array_equal(nxt, orc.dense_next_state(shape[0], shape[1], axis, ctr))


@pytest.mark.parametrize("layout", ["C", "R"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_submatrix_is_slice(orc, layout, dtype):
    rng = np.random.default_rng(1)
    for shape, axis in [((20, 37), "L"), ((37, 20), "L"), ((20, 37), "S"), ((37, 20), "S")]:
        full, _ = orc.fill_dense_unpacked(layout, shape[0], shape[1], shape[0], shape[1], 0, 0, [0, 0, 0, 0],
                                          [9, 0], "G", axis, dtype)
        for _ in range(12):
            r = int(rng.integers(1, shape[0] + 1)); c = int(rng.integers(1, shape[1] + 1))
            ro = int(rng.integers(0, shape[0] - r + 1)); co = int(rng.integers(0, shape[1] - c + 1))
            sub, _ = orc.fill_dense_unpacked(layout, shape[0], shape[1], r, c, ro, co, [0, 0, 0, 0], [9, 0],
                                             "G", axis, dtype)
            assert np.array_equal(sub, full[ro:ro + r, co:co + c])


def test_next_state_concatenation(orc):
    # wide, Long-axis: rows are major vectors; D(r1,c) then D(r2,c) from next_state
    # are the rows of D(r1+r2, c) (test_denseskop.cc:405-489)
    c = 29
    full, nxt_full = orc.fill_dense(10, c, [0, 0, 0, 0], [5, 0])
    top, nxt1 = orc.fill_dense(4, c, [0, 0, 0, 0], [5, 0])
    bot, nxt2 = orc.fill_dense(6, c, nxt1, [5, 0])
    assert np.array_equal(np.vstack([top, bot]), full)
    assert np.array_equal(nxt2, nxt_full)
    assert list(nxt1) == [4 * 8, 0, 0, 0]  # P = ceil(29/4) = 8 counters per row


def test_autotranspose(orc):
    # DenseDist(m, n) and DenseDist(n, m) with the same axis are transposes
    a, _ = orc.fill_dense(11, 30, [0, 0, 0, 0], [1, 0])
    b, _ = orc.fill_dense(30, 11, [0, 0, 0, 0], [1, 0])
    assert np.array_equal(a, b.T)


def test_submat_next_state_matches_reference_formula(orc):
    # fill_dense_submat_impl returns seed + ctr_mat_start + n_srows*P
    # (dense_skops.hh:109-169); natural RowMajor 40 x 100, P = 25
    _, nxt = orc.fill_dense_unpacked("R", 40, 100, 5, 7, 3, 9, [0, 0, 0, 0], [1, 0])
    assert list(nxt) == [3 * 25 + 9 // 4 + 5 * 25, 0, 0, 0]


@pytest.mark.parametrize("fam,std", [("G", 1.0), ("U", 1.0)])
def test_moments(orc, fam, std):
    for key in (0, 1, 2):
        m, _ = orc.fill_dense(203, 503, [0, 0, 0, 0], [key, 0], family=fam)
        assert abs(m.mean()) < 1e-2
        assert abs(m.std(ddof=1) - std) < 1e-2


def _strided(buf, rows, cols, layout, ld):
    if layout == "C":
        return np.lib.stride_tricks.as_strided(buf, (rows, cols), (buf.itemsize, ld * buf.itemsize))
    return np.lib.stride_tricks.as_strided(buf, (rows, cols), (ld * buf.itemsize, buf.itemsize))


@pytest.mark.parametrize("layout", ["C", "R"])
@pytest.mark.parametrize("opS", ["N", "T"])
@pytest.mark.parametrize("opA", ["N", "T"])
def test_lskge3_oracle_vs_numpy(orc, layout, opS, opA):
    rng = np.random.default_rng(7)
    d, n, m = 13, 9, 31
    ro, co = 2, 3
    Sr, Sc = (d + ro + 4, m + co + 1) if opS == "N" else (m + ro + 2, d + co + 5)
    full, _ = orc.fill_dense(Sr, Sc, [0, 0, 0, 0], [42, 0])
    sub = full[ro:ro + (d if opS == "N" else m), co:co + (m if opS == "N" else d)]
    opSm = sub if opS == "N" else sub.T
    rA, cA = (m, n) if opA == "N" else (n, m)
    lda = (rA if layout == "C" else cA) + 3
    ldb = (d if layout == "C" else n) + 2
    Abuf = rng.standard_normal(lda * max(rA, cA) + 8)
    Am = _strided(Abuf, rA, cA, layout, lda)
    opAm = Am if opA == "N" else Am.T
    Bbuf = rng.standard_normal(ldb * max(d, n) + 8)
    B0 = _strided(Bbuf.copy(), d, n, layout, ldb).copy()
    alpha, beta = 1.5, -0.5
    orc.lskge3("d", layout, opS, opA, d, n, m, alpha, (Sr, Sc), [0, 0, 0, 0], [42, 0], ro, co, Abuf, lda, beta,
               Bbuf, ldb)
    got = _strided(Bbuf, d, n, layout, ldb)
    ref = alpha * opSm @ opAm + beta * B0
    bound = 2 * m * np.finfo(float).eps * abs(alpha) * (np.abs(opSm) @ np.abs(opAm)) + \
        np.finfo(float).eps * abs(beta) * np.abs(B0)
    assert np.all(np.abs(got - ref) <= bound + 1e-300)


def test_sketch_of_identity_is_operator(orc):
    # S * I == S exactly (linop_common.hh:308-358)
    d, m = 17, 23
    S, _ = orc.fill_dense(d, m, [0, 0, 0, 0], [3, 0])
    I = np.eye(m).reshape(-1, order="F").copy()
    B = np.zeros(d * m)
    orc.lskge3("d", "C", "N", "N", d, m, m, 1.0, (d, m), [0, 0, 0, 0], [3, 0], 0, 0, I, m, 0.0, B, d)
    assert np.array_equal(B.reshape(m, d).T, S)


@pytest.mark.parametrize("layout", ["C", "R"])
@pytest.mark.parametrize("opS", ["N", "T"])
def test_rskge3_oracle_vs_numpy(orc, layout, opS):
    rng = np.random.default_rng(8)
    m, d, n = 11, 7, 29
    Sr, Sc = (n, d) if opS == "N" else (d, n)
    full, _ = orc.fill_dense(Sr, Sc, [0, 0, 0, 0], [4, 0])
    opSm = full if opS == "N" else full.T
    lda = (m if layout == "C" else n)
    A = rng.standard_normal(m * n)
    Am = _strided(A, m, n, layout, lda)
    ldb = (m if layout == "C" else d)
    B = np.zeros(m * d)
    orc.rskge3("d", layout, "N", opS, m, d, n, 1.0, A, lda, (Sr, Sc), [0, 0, 0, 0], [4, 0], 0, 0, 0.0, B, ldb)
    np.testing.assert_allclose(_strided(B, m, d, layout, ldb), Am @ opSm, rtol=1e-13, atol=1e-13)
