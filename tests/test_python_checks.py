"""CPU: the Python layer rejects mismatched operands before any pointer
reaches the library (no GPU needed: every case raises before a launch).

The C ABI takes raw device pointers, so a float64 operator applied to a
float32 matrix, an int32 index array or a buffer shorter than the strided
extent would otherwise be read as other bits or out of bounds.  The
reference's typed DenseSkOp<T> / SparseSkOp<T, RNG, int64_t> rule these out
at compile time (dense_skops.hh:357-478, sparse_skops.hh:289-450)."""
import pytest
import torch

import randblas_amd as rb


def f32(n):
    return torch.zeros(n, dtype=torch.float32)


def f64(n):
    return torch.zeros(n, dtype=torch.float64)


def i64(n):
    return torch.zeros(n, dtype=torch.int64)


def test_extent():
    assert rb._extent(3, 4, "C", 5) == 3 * 5 + 3
    assert rb._extent(3, 4, "R", 6) == 2 * 6 + 4
    assert rb._extent(0, 4, "C", 5) == 0


def test_materialised_operator_dtype_mismatch():
    # fill_dense(S) defaults to float64; a float32 sketch with it must raise
    S = rb.DenseSkOp(rb.DenseDist(8, 16), rb.RNGState(1), buff=f64(8 * 16))
    with pytest.raises(rb.Error, match="S.buff: dtype"):
        rb.lskge3("C", "N", "N", 8, 4, 16, 1.0, S, 0, 0, f32(16 * 4), 16, 0.0, f32(8 * 4), 8)
    with pytest.raises(rb.Error, match="S.buff: dtype"):
        rb.rskge3("C", "N", "N", 4, 8, 16, 1.0, f32(4 * 16), 4, rb.DenseSkOp(rb.DenseDist(16, 8), rb.RNGState(1),
                                                                          buff=f64(128)), 0, 0, 0.0, f32(32), 4)


def test_operator_dtype_fixed():
    S = rb.DenseSkOp(rb.DenseDist(8, 16), rb.RNGState(1), dtype=torch.float32)
    assert S.dtype == torch.float32
    with pytest.raises(rb.Error):
        rb.DenseSkOp(rb.DenseDist(8, 16), rb.RNGState(1), buff=f64(128), dtype=torch.float32)
    with pytest.raises(rb.Error, match="holds"):
        rb.fill_dense(S, dtype=torch.float64)


def test_operand_dtype_mismatch():
    S = rb.DenseSkOp(rb.DenseDist(8, 16), rb.RNGState(1))
    with pytest.raises(rb.Error, match="A: dtype"):
        rb.lskge3("C", "N", "N", 8, 4, 16, 1.0, S, 0, 0, f64(64), 16, 0.0, f32(32), 8)
    with pytest.raises(rb.Error, match="Bbar: dtype"):
        rb.lskge3_adjoint("C", "N", 8, 4, 16, 1.0, S, 0, 0, f32(32), 8, 0.0, f64(64), 16)
    with pytest.raises(rb.Error, match="A: dtype"):
        rb.lskge3_bf16split("C", "N", "N", 8, 4, 16, 1.0, S, 0, 0, f64(64), 16, 0.0, f32(32), 8)


def test_short_buffers():
    S = rb.DenseSkOp(rb.DenseDist(8, 16), rb.RNGState(1))
    # A is 16 x 4 column-major with lda 20: (4 - 1) * 20 + 16 = 76 elements
    with pytest.raises(rb.Error, match="A: 75 elements"):
        rb.lskge3("C", "N", "N", 8, 4, 16, 1.0, S, 0, 0, f32(75), 20, 0.0, f32(32), 8)
    with pytest.raises(rb.Error, match="B: 31 elements"):
        rb.lskge3("C", "N", "N", 8, 4, 16, 1.0, S, 0, 0, f32(76), 20, 0.0, f32(31), 8)
    # op(A) = T: A stored 4 x 16 row-major with lda 16
    with pytest.raises(rb.Error, match="A: 63 elements"):
        rb.lskge3("R", "N", "T", 8, 4, 16, 1.0, S, 0, 0, f32(63), 16, 0.0, f32(32), 4)
    # alpha = 0: A is never read, so it is not checked
    with pytest.raises(rb.Error, match="operates on GPU"):
        rb.lskge3("C", "N", "N", 8, 4, 16, 0.0, S, 0, 0, None, 16, 0.0, f32(32), 8)


def test_non_contiguous():
    S = rb.DenseSkOp(rb.DenseDist(8, 16), rb.RNGState(1))
    A = f32(2 * 64)[::2]
    with pytest.raises(rb.Error, match="contiguous"):
        rb.lskge3("C", "N", "N", 8, 4, 16, 1.0, S, 0, 0, A, 16, 0.0, f32(32), 8)


def test_sparse_operator_and_data_checks():
    D = rb.SparseDist(8, 16, 2)
    S = rb.SparseSkOp(D, rb.RNGState(1), dtype=torch.float64)
    S.vals, S.rows, S.cols, S.nnz = f64(32), i64(32), i64(32), 32
    with pytest.raises(rb.Error, match="holds"):
        rb.lskges("C", "N", "N", 8, 4, 16, 1.0, S, 0, 0, f32(64), 16, 0.0, f32(32), 8)
    S.rows = torch.zeros(32, dtype=torch.int32)
    with pytest.raises(rb.Error, match="S.rows: dtype"):
        rb.lskges("C", "N", "N", 8, 4, 16, 1.0, S, 0, 0, f64(64), 16, 0.0, f64(32), 8)
    with pytest.raises(rb.Error, match="rows: dtype"):
        rb.fill_sparse_unpacked_nosub(D, f64(32), torch.zeros(32, dtype=torch.int32), i64(32), rb.RNGState(1))
    with pytest.raises(rb.Error, match="vals: 31 elements"):
        rb.fill_sparse_unpacked_nosub(D, f64(31), i64(32), i64(32), rb.RNGState(1))
    A = rb.CSRMatrix(10, 16, f32(5), i64(10), i64(5))  # rowptr needs 11
    Sd = rb.DenseSkOp(rb.DenseDist(16, 4), rb.RNGState(2))
    with pytest.raises(rb.Error, match="A.rowptr: 10 elements"):
        rb.rsksp3("C", "N", "N", 10, 4, 16, 1.0, A, 0, 0, Sd, 0, 0, 0.0, f32(40), 10)
    A = rb.COOMatrix(10, 16, f64(5), i64(5), torch.zeros(5, dtype=torch.int32))
    with pytest.raises(rb.Error, match="A.cols: dtype"):
        rb.rsksp3("C", "N", "N", 10, 4, 16, 1.0, A, 0, 0, Sd, 0, 0, 0.0, f64(40), 10)
    with pytest.raises(rb.Error, match="A.vals: dtype"):
        rb.rsksp3("C", "N", "N", 10, 4, 16, 1.0, A, 0, 0, Sd, 0, 0, 0.0, f32(40), 10)


def test_sampling_and_solve_checks():
    with pytest.raises(rb.Error, match="samples: dtype"):
        rb.sample_indices_iid_uniform(100, 8, torch.zeros(8, dtype=torch.int32), rb.RNGState(1))
    with pytest.raises(rb.Error, match="samples: 7 elements"):
        rb.sample_indices_iid_uniform(100, 8, i64(7), rb.RNGState(1))
    with pytest.raises(rb.Error, match="cdf: 9 elements"):
        rb.sample_indices_iid(10, f64(9), 4, i64(4), rb.RNGState(1))
    S = rb.DenseSkOp(rb.DenseDist(16, 64), rb.RNGState(1))
    with pytest.raises(rb.Error, match="b: dtype"):
        rb.sketch_solve("C", 16, 64, 4, S, 0, 0, f32(256), 64, f64(64), f32(4))
    with pytest.raises(rb.Error, match="x: 3 elements"):
        rb.sketch_solve("C", 16, 64, 4, S, 0, 0, f32(256), 64, f32(64), f32(3))


def test_conversion_checks():
    src = rb.CSRMatrix(4, 5, f32(6), i64(5), i64(6))
    dst = rb.COOMatrix(4, 5, f32(6), i64(6), i64(5))
    with pytest.raises(rb.Error, match="dst.cols: 5 elements"):
        rb.csr_to_coo(src, dst)
    dst = rb.CSCMatrix(4, 5, f64(6), i64(6), i64(6))
    with pytest.raises(rb.Error, match="dst.vals: dtype"):
        rb.coo_to_csc(rb.COOMatrix(4, 5, f32(6), i64(6), i64(6)), dst)
