"""The C++ layer (randblas_amd/cpp/RandBLAS.hh), driven by tests/cpp/test_cpp_api.cc.

CPU: RNGState / DenseDist / SparseDist / DenseSkOp / SparseSkOp semantics and
argument errors (the program's own checks).
GPU: the program runs sketch_general (dense left and right, sparse left),
sketch_vector, the adjoint, fill_dense(S), fill_sparse(S), sketch_symmetric's
symmetry error and sketch_sparse (rsksp3) through the C++ API and writes the
raw results; they are checked here against the CPU oracle -- bit-exact for
generated operators, sparse operators and sparse products, relative Frobenius
error <= 1e-13 for dense FP64 products.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, rel_fro

CPP_DIR = os.path.join(ROOT, "tests", "cpp")
EXE = os.path.join(CPP_DIR, "_build", "test_cpp_api")


def _run(args, timeout):
    r = subprocess.run([EXE] + args, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def test_cpp_layer_host_semantics():
    if not os.path.exists(EXE):
        subprocess.run(["make", "-C", CPP_DIR], check=True, capture_output=True)
    out = _run(["host"], 120)
    assert "0 failed" in out


@pytest.mark.gpu
def test_cpp_layer_on_gpu(rb, orc, tmp_path):
    assert os.path.exists(EXE), f"{EXE} is missing: build with `make -C tests/cpp`"
    out = _run(["gpu", str(tmp_path)], 300)
    assert "0 failed" in out

    def load(name, dt):
        return np.fromfile(tmp_path / f"{name}.bin", dtype=dt)

    d, m, n = 400, 2000, 200
    # A from DenseDist(2000, 200) key 1 in its natural (column-major) layout
    Am, _ = orc.fill_dense(m, n, [0, 0, 0, 0], [1, 0])
    A = load("dense_A", np.float64)
    assert np.array_equal(A, Am.reshape(-1, order="F"))
    ref = np.zeros(d * n)
    orc.lskge3("d", "C", "N", "N", d, n, m, 1.0, (d, m), [0, 0, 0, 0], [42, 0], 0, 0, A, m, 0.0, ref, d)
    B = load("dense_B", np.float64)
    assert rel_fro(B, ref) <= 1e-13
    Bt = load("dense_Bt", np.float64).reshape(d, n)  # n x d column-major = its transpose row-major
    assert rel_fro(Bt.reshape(-1), ref.reshape(n, d).T.reshape(-1)) <= 1e-13
    y = load("vector_y", np.float64)
    assert rel_fro(y, ref[:d]) <= 1e-13
    Sm, _ = orc.fill_dense(d, m, [0, 0, 0, 0], [42, 0])
    Abar = load("adjoint_Abar", np.float64)
    assert rel_fro(Abar, (Sm.T @ B.reshape(n, d).T).reshape(-1, order="F")) <= 1e-13

    Sf, _ = orc.fill_dense(64, 100, [0, 0, 0, 0], [5, 0], dtype=np.float32)  # natural RowMajor
    assert np.array_equal(load("dense_fill_f32", np.float32), Sf.reshape(-1))

    sd, sm, sn = 200, 5000, 50
    vals, rows, cols, nnz, _ = orc.fill_sparse(sd, sm, 8, "S", [0, 0, 0, 0], [7, 0])
    assert np.array_equal(load("sparse_vals", np.float64), vals)
    assert np.array_equal(load("sparse_rows", np.int64), rows)
    assert np.array_equal(load("sparse_cols", np.int64), cols)
    A2m, _ = orc.fill_dense(sm, sn, [0, 0, 0, 0], [3, 0], family="U")
    A2 = load("sparse_A", np.float64)
    assert np.array_equal(A2, A2m.reshape(-1, order="F"))
    ref2 = np.zeros(sd * sn)
    orc.left_spmm_coo("d", "C", "N", "N", sd, sn, sm, 1.0, (sd, sm), vals, rows, cols, 0, 0, A2, sm, 0.0, ref2, sd)
    assert np.array_equal(load("sparse_B", np.float64), ref2)

    cm, cn, cd = 50, 300, 16
    cv = load("csr_vals", np.float64)
    rp = load("csr_rowptr", np.int64)
    ci = load("csr_colidx", np.int64)
    S4, _ = orc.fill_dense(cn, cd, [0, 0, 0, 0], [42, 0])  # natural ColMajor
    assert orc.natural_layout(cn, cd, "L") == "C"
    ref4 = np.zeros(cm * cd)
    orc.left_spmm_csr("d", "C", "N", cm, cd, cn, 1.0, cv, rp, ci, np.ascontiguousarray(S4.reshape(-1, order="F")),
                      cn, 0.0, ref4, cm)
    assert np.array_equal(load("csr_B", np.float64), ref4)

    # index sampling through the C++ layer
    k = 1001
    s1, _, _ = orc.sample_indices_iid_uniform(100, k, [0, 0, 0, 0], [42, 0])
    assert np.array_equal(load("iid_uniform", np.int64), s1)
    s2, r2, _ = orc.sample_indices_iid_uniform(2000, k, [0, 0, 0, 0], [42, 0], signed=True)
    assert np.array_equal(load("iid_signed", np.int64), s2) and np.array_equal(load("iid_signs", np.float64), r2)
    cdf, st = orc.weights_to_cdf(1.0 / (np.arange(300) + 1.0) ** 1.5)
    s3, _ = orc.sample_indices_iid(300, cdf, k, [0, 0, 0, 0], [7, 0])
    assert st == 0 and np.array_equal(load("iid_cdf_samples", np.int64), s3)
