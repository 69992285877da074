"""GPU: sparse-data format conversions (sparse_data/conversions.hh:49-126,
sparse_data/base.hh:142-220) against numpy: COO -> CSR / CSC is a stable
sort by (row, column) / (column, row) (numpy.lexsort is stable), CSR / CSC
-> COO expands the pointers; determine_sort classifies as the reference's
increasing_by_csc / increasing_by_csr do.  Includes empty rows and columns,
duplicate entries (kept in stored order) and nnz = 0."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda"


def t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def host(x):
    torch.cuda.synchronize()
    return x.cpu().numpy()


def random_coo(rng, m, n, nnz, dup=True):
    r = rng.integers(0, m, nnz).astype(np.int64)
    c = rng.integers(0, n, nnz).astype(np.int64)
    if dup and nnz > 4:
        r[-3:], c[-3:] = r[:3], c[:3]   # duplicates of earlier entries
    v = rng.standard_normal(nnz)
    return r, c, v


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("m,n,nnz", [(50, 40, 300), (1000, 3, 2000), (7, 9000, 500), (5, 5, 0), (1, 1, 1)])
@pytest.mark.parametrize("given", ["shuffled", "by_row", "by_col"])
def test_coo_to_csr_and_csc(rb, dtype, m, n, nnz, given):
    """Entries given in any order, ordered by (row, column) (CSR order: the
    conversion to CSR copies, to CSC sorts on the column alone) or by
    (column, row); duplicates stay in stored order."""
    rng = np.random.default_rng(m + n + nnz)
    r, c, v = random_coo(rng, m, n, nnz)
    if given != "shuffled":
        o = np.lexsort((c, r)) if given == "by_row" else np.lexsort((r, c))
        r, c, v = r[o], c[o], v[o]
    v = v.astype(dtype)
    coo = rb.COOMatrix(m, n, t(v), t(r), t(c))
    # CSR
    o = np.lexsort((c, r))               # by row, then column; stable
    rp = torch.empty(m + 1, dtype=torch.int64, device=DEV)
    ci = torch.empty(max(nnz, 1), dtype=torch.int64, device=DEV)
    cv = torch.empty(max(nnz, 1), dtype=torch.float64 if dtype == np.float64 else torch.float32, device=DEV)
    rb.coo_to_csr(coo, rb.CSRMatrix(m, n, cv[:nnz] if nnz else cv, rp, ci))
    assert np.array_equal(host(rp), np.searchsorted(r[o], np.arange(m + 1), side="left"))
    assert np.array_equal(host(ci)[:nnz], c[o]) and np.array_equal(host(cv)[:nnz], v[o])
    # CSC
    o = np.lexsort((r, c))
    cp = torch.empty(n + 1, dtype=torch.int64, device=DEV)
    ri = torch.empty(max(nnz, 1), dtype=torch.int64, device=DEV)
    rv = torch.empty_like(cv)
    rb.coo_to_csc(coo, rb.CSCMatrix(m, n, rv[:nnz] if nnz else rv, ri, cp))
    assert np.array_equal(host(cp), np.searchsorted(c[o], np.arange(n + 1), side="left"))
    assert np.array_equal(host(ri)[:nnz], r[o]) and np.array_equal(host(rv)[:nnz], v[o])


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_compressed_to_coo_round_trip(rb, dtype):
    rng = np.random.default_rng(5)
    m, n, nnz = 300, 200, 4000
    r, c, v = random_coo(rng, m, n, nnz)
    o = np.lexsort((c, r))
    r, c, v = r[o], c[o], v[o].astype(dtype)
    rowptr = np.searchsorted(r, np.arange(m + 1), side="left")
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    out = rb.COOMatrix(m, n, torch.empty(nnz, dtype=tdt, device=DEV), torch.empty(nnz, dtype=torch.int64, device=DEV),
                       torch.empty(nnz, dtype=torch.int64, device=DEV))
    rb.csr_to_coo(rb.CSRMatrix(m, n, t(v), t(rowptr), t(c)), out)
    assert np.array_equal(host(out.rows), r) and np.array_equal(host(out.cols), c) and np.array_equal(host(out.vals), v)
    assert rb.coo_determine_sort(out) == "R"
    # CSC of the same matrix -> COO is in CSC order
    o = np.lexsort((r, c))
    colptr = np.searchsorted(c[o], np.arange(n + 1), side="left")
    rb.csc_to_coo(rb.CSCMatrix(m, n, t(v[o]), t(r[o]), t(colptr)), out)
    assert rb.coo_determine_sort(out) == "C"
    assert np.array_equal(host(out.rows), r[o]) and np.array_equal(host(out.cols), c[o])
    # in-place sorts
    rb.coo_sort(out, "R")
    assert rb.coo_determine_sort(out) == "R" and np.array_equal(host(out.vals), v)
    perm = rng.permutation(nnz)
    shuffled = rb.COOMatrix(m, n, t(v[perm]), t(r[perm]), t(c[perm]))
    assert rb.coo_determine_sort(shuffled) == "N"
    rb.coo_sort(shuffled, "C")
    oc = np.lexsort((r[perm], c[perm]))
    assert np.array_equal(host(shuffled.vals), v[perm][oc])


def test_conversion_argument_checks(rb):
    coo = rb.COOMatrix(3, 4, t(np.ones(2)), t(np.array([0, 1])), t(np.array([2, 3])))
    with pytest.raises(rb.Error):
        rb.coo_to_csr(coo, rb.CSRMatrix(4, 4, t(np.ones(2)), t(np.zeros(5, np.int64)), t(np.zeros(2, np.int64))))
    with pytest.raises(rb.Error):
        rb.coo_sort(coo, "X")
