"""ctypes front end of the CPU oracle (oracle/rbx_oracle.c).

TEST INFRASTRUCTURE ONLY.  This package is the checker for the HIP
implementation: tests/, __graft_entry__.smoke() and the CPU-baseline leg of
bench.py may import it; the product (randblas_amd) never does.

Every helper here is a thin wrapper; the restated algorithms and their
RandBLAS citations live in rbx_oracle.c.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")

_lib = None


def build() -> None:
    """Compile liboracle.so (gcc + OpenMP)."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.orc_sample_indices_iid_uniform.argtypes = [C.c_int64, C.c_int64, _i64p, C.c_void_p, _u32p, _u32p, _u32p]
        L.orc_sample_indices_iid.argtypes = [C.c_char, C.c_int64, C.c_void_p, C.c_int64, _i64p, _u32p, _u32p,
                                             _u32p]
        L.orc_weights_to_cdf.argtypes = [C.c_char, C.c_int64, C.c_void_p, C.c_double]
        L.orc_weights_to_cdf.restype = C.c_int64
        L.orc_power_law_weights.argtypes = [C.c_int64, C.c_double, C.c_void_p]
        L.orc_philox_stream.argtypes = [_u32p, _u32p, C.c_int64, _u32p]
        L.orc_ctr_incr.argtypes = [_u32p, C.c_uint64]
        L.orc_uneg11.argtypes = [C.c_uint32]
        L.orc_uneg11.restype = C.c_float
        L.orc_u01.argtypes = [C.c_uint32]
        L.orc_u01.restype = C.c_float
        L.orc_sinpi_arg.argtypes = [C.c_uint32]
        L.orc_sinpi_arg.restype = C.c_float
        L.orc_boxmuller.argtypes = [C.c_uint32, C.c_uint32, C.c_int, _f32p]
        L.orc_boxmuller_bulk.argtypes = [C.c_int64, _u32p, C.c_int, _f32p]
        L.orc_factor_bulk.argtypes = [C.c_int64, _u32p, C.c_int, C.c_int, _f32p]
        L.orc_dense_next_state.argtypes = [C.c_int64, C.c_int64, C.c_char, _u32p, _u32p]
        L.orc_fill_dense_unpacked.argtypes = [
            C.c_char, C.c_int64, C.c_int64, C.c_char, C.c_char, C.c_int64, C.c_int64,
            C.c_int64, C.c_int64, C.c_char, C.c_void_p, _u32p, _u32p, _u32p, C.c_int, C.c_int]
        L.orc_gemm_d.argtypes = [
            C.c_int64, C.c_int64, C.c_int64, C.c_double, C.c_void_p, C.c_int64, C.c_int64,
            C.c_void_p, C.c_int64, C.c_int64, C.c_double, C.c_void_p, C.c_int64, C.c_int64, C.c_int]
        L.orc_lskge3.argtypes = [
            C.c_char, C.c_char, C.c_char, C.c_char, C.c_int64, C.c_int64, C.c_int64, C.c_double,
            C.c_int64, C.c_int64, C.c_char, C.c_char, _u32p, _u32p, C.c_int64, C.c_int64,
            C.c_void_p, C.c_int64, C.c_double, C.c_void_p, C.c_int64, C.c_int, C.c_int]
        L.orc_rskge3.argtypes = [
            C.c_char, C.c_char, C.c_char, C.c_char, C.c_int64, C.c_int64, C.c_int64, C.c_double,
            C.c_void_p, C.c_int64, C.c_int64, C.c_int64, C.c_char, C.c_char, _u32p, _u32p,
            C.c_int64, C.c_int64, C.c_double, C.c_void_p, C.c_int64, C.c_int, C.c_int]
        L.orc_sparse_next_state.argtypes = [C.c_int64, C.c_int64, C.c_int64, C.c_char, _u32p, _u32p]
        L.orc_fill_sparse.argtypes = [
            C.c_int64, C.c_int64, C.c_int64, C.c_char, C.c_char, C.POINTER(C.c_int64),
            C.c_void_p, _i64p, _i64p, _u32p, _u32p, _u32p]
        L.orc_left_spmm_coo.argtypes = [
            C.c_char, C.c_char, C.c_char, C.c_char, C.c_int64, C.c_int64, C.c_int64, C.c_double,
            C.c_int64, C.c_int64, C.c_int64, C.c_void_p, _i64p, _i64p, C.c_int64, C.c_int64,
            C.c_void_p, C.c_int64, C.c_double, C.c_void_p, C.c_int64, C.c_int]
        L.orc_left_spmm_csr.argtypes = [
            C.c_char, C.c_char, C.c_char, C.c_int64, C.c_int64, C.c_int64, C.c_double,
            C.c_void_p, _i64p, _i64p, C.c_void_p, C.c_int64, C.c_double, C.c_void_p,
            C.c_int64, C.c_int]
        L.orc_sddmm_csr.argtypes = [
            C.c_char, C.c_int64, C.c_int64, C.c_double, _i64p, _i64p, C.c_void_p, C.c_int64,
            C.c_int64, C.c_void_p, C.c_int64, C.c_int64, C.c_void_p, C.c_int]
        for f in ("orc_dense_next_state", "orc_fill_dense_unpacked", "orc_lskge3", "orc_rskge3",
                  "orc_sparse_next_state", "orc_fill_sparse", "orc_left_spmm_coo",
                  "orc_left_spmm_csr", "orc_sddmm_csr"):
            getattr(L, f).restype = C.c_int
        _lib = L
    return _lib


def _b(c: str) -> bytes:
    return c.encode()


def _ctr(counter) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(counter, dtype=np.uint32).reshape(4))


def _key(key) -> np.ndarray:
    k = np.asarray(key, dtype=np.uint64).reshape(-1)
    if k.size == 1:  # RNGState(uint64) constructor: key = {k mod 2^32, k >> 32}
        v = int(k[0])
        return np.array([v & 0xFFFFFFFF, v >> 32], dtype=np.uint32)
    return np.ascontiguousarray(k.astype(np.uint32).reshape(2))


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise ValueError(f"oracle {what}: invalid argument")


# ---------------------------------------------------------------- Philox
def philox(counter, key) -> np.ndarray:
    out = np.zeros(4, np.uint32)
    lib().orc_philox4x32_10(_ctr(counter), _key(key), out)
    return out


def philox_stream(counter, key, n_ctr: int) -> np.ndarray:
    out = np.zeros(4 * n_ctr, np.uint32)
    lib().orc_philox_stream(_ctr(counter), _key(key), n_ctr, out)
    return out.reshape(n_ctr, 4)


def ctr_incr(counter, n: int) -> np.ndarray:
    c = _ctr(counter).copy()
    lib().orc_ctr_incr(c, n)
    return c


def boxmuller(words: np.ndarray, mathmode: int = 0) -> np.ndarray:
    w = np.ascontiguousarray(words, dtype=np.uint32).reshape(-1)
    assert w.size % 2 == 0
    out = np.zeros(w.size, np.float32)
    lib().orc_boxmuller_bulk(w.size // 2, w, mathmode, out)
    return out


def factor(words: np.ndarray, which: int, mathmode: int = 0) -> np.ndarray:
    """which: 0 sin(pi*uneg11(w)), 1 cos(pi*uneg11(w)), 2 sqrt(-2 log u01(w))."""
    w = np.ascontiguousarray(words, dtype=np.uint32).reshape(-1)
    out = np.zeros(w.size, np.float32)
    lib().orc_factor_bulk(w.size, w, which, mathmode, out)
    return out


def uneg11(words: np.ndarray) -> np.ndarray:
    L = lib()
    return np.array([L.orc_uneg11(int(x)) for x in np.asarray(words).reshape(-1)], np.float32)


# ---------------------------------------------------------------- dense
def dense_next_state(n_rows, n_cols, major_axis, counter) -> np.ndarray:
    out = np.zeros(4, np.uint32)
    _check(lib().orc_dense_next_state(n_rows, n_cols, _b(major_axis), _ctr(counter), out),
           "dense_next_state")
    return out


def fill_dense_unpacked(layout, D_rows, D_cols, n_rows, n_cols, ro_s, co_s, counter, key,
                        family="G", major_axis="L", dtype=np.float64, mathmode=0, nthreads=0):
    """Returns (matrix as a 2-D numpy array of shape (n_rows, n_cols), next_counter)."""
    dt = np.dtype(dtype)
    buf = np.zeros(n_rows * n_cols, dt)
    nxt = np.zeros(4, np.uint32)
    rc = lib().orc_fill_dense_unpacked(
        _b(layout), D_rows, D_cols, _b(family), _b(major_axis), n_rows, n_cols, ro_s, co_s,
        b"d" if dt == np.float64 else b"s", buf.ctypes.data, _ctr(counter), _key(key), nxt,
        mathmode, nthreads)
    _check(rc, "fill_dense_unpacked")
    if layout == "C":
        mat = buf.reshape(n_cols, n_rows).T
    else:
        mat = buf.reshape(n_rows, n_cols)
    return mat, nxt


def fill_dense(D_rows, D_cols, counter, key, family="G", major_axis="L", dtype=np.float64,
               mathmode=0, nthreads=0):
    """fill_dense(D, buff, seed) in the natural layout; returns (matrix, next_counter)."""
    lay = natural_layout(D_rows, D_cols, major_axis)
    return fill_dense_unpacked(lay, D_rows, D_cols, D_rows, D_cols, 0, 0, counter, key,
                               family, major_axis, dtype, mathmode, nthreads)


def natural_layout(n_rows, n_cols, major_axis) -> str:
    wide = n_rows < n_cols
    flong = major_axis == "L"
    if wide and flong:
        return "R"
    if wide:
        return "C"
    if flong:
        return "C"
    return "R"


def _ptr(a):
    return None if a is None else a.ctypes.data


def lskge3(dtype, layout, opS, opA, d, n, m, alpha, S_shape, counter, key, ro_s, co_s,
           A: np.ndarray, lda, beta, B: np.ndarray, ldb, family="G", major_axis="L",
           mathmode=0, nthreads=0):
    """B (flat buffer, modified in place) = alpha*op(submat S)*op(A) + beta*B."""
    rc = lib().orc_lskge3(_b(dtype), _b(layout), _b(opS), _b(opA), d, n, m, alpha,
                          S_shape[0], S_shape[1], _b(family), _b(major_axis), _ctr(counter),
                          _key(key), ro_s, co_s, _ptr(A), lda, beta, _ptr(B), ldb, mathmode,
                          nthreads)
    _check(rc, "lskge3")


def rskge3(dtype, layout, opA, opS, m, d, n, alpha, A, lda, S_shape, counter, key, ro_s, co_s,
           beta, B, ldb, family="G", major_axis="L", mathmode=0, nthreads=0):
    rc = lib().orc_rskge3(_b(dtype), _b(layout), _b(opA), _b(opS), m, d, n, alpha, _ptr(A), lda,
                          S_shape[0], S_shape[1], _b(family), _b(major_axis), _ctr(counter),
                          _key(key), ro_s, co_s, beta, _ptr(B), ldb, mathmode, nthreads)
    _check(rc, "rskge3")


def gemm_d(M, N, K, alpha, X, xrs, xcs, Y, yrs, ycs, beta, Cm, crs, ccs, nthreads=0):
    lib().orc_gemm_d(M, N, K, alpha, _ptr(X), xrs, xcs, _ptr(Y), yrs, ycs, beta, _ptr(Cm),
                     crs, ccs, nthreads)


# ---------------------------------------------------------------- sparse
def sparse_next_state(n_rows, n_cols, vec_nnz, major_axis, counter) -> np.ndarray:
    out = np.zeros(4, np.uint32)
    _check(lib().orc_sparse_next_state(n_rows, n_cols, vec_nnz, _b(major_axis), _ctr(counter),
                                       out), "sparse_next_state")
    return out


def fill_sparse(n_rows, n_cols, vec_nnz, major_axis, counter, key, dtype=np.float64):
    """Returns (vals, rows, cols, nnz, next_counter); arrays are trimmed to nnz."""
    dim_major = min(n_rows, n_cols) if major_axis == "S" else max(n_rows, n_cols)
    dim_minor = n_rows + n_cols - dim_major
    full = vec_nnz * dim_minor
    dt = np.dtype(dtype)
    vals = np.zeros(full, dt)
    rows = np.zeros(full, np.int64)
    cols = np.zeros(full, np.int64)
    nnz = C.c_int64(0)
    nxt = np.zeros(4, np.uint32)
    rc = lib().orc_fill_sparse(n_rows, n_cols, vec_nnz, _b(major_axis),
                               b"d" if dt == np.float64 else b"s", C.byref(nnz),
                               vals.ctypes.data, rows, cols, _ctr(counter), _key(key), nxt)
    _check(rc, "fill_sparse")
    k = nnz.value
    return vals[:k].copy(), rows[:k].copy(), cols[:k].copy(), k, nxt


def left_spmm_coo(dtype, layout, opA, opB, d, n, m, alpha, A_shape, vals, rows, cols, ro_a, co_a,
                  B, ldb, beta, Cbuf, ldc, nthreads=0):
    rc = lib().orc_left_spmm_coo(_b(dtype), _b(layout), _b(opA), _b(opB), d, n, m, alpha,
                                 A_shape[0], A_shape[1], len(rows), _ptr(vals),
                                 np.ascontiguousarray(rows, np.int64),
                                 np.ascontiguousarray(cols, np.int64), ro_a, co_a, _ptr(B), ldb,
                                 beta, _ptr(Cbuf), ldc, nthreads)
    _check(rc, "left_spmm_coo")


def left_spmm_csr(dtype, layout, opB, d, n, m, alpha, vals, rowptr, colidxs, B, ldb, beta, Cbuf,
                  ldc, nthreads=0):
    rc = lib().orc_left_spmm_csr(_b(dtype), _b(layout), _b(opB), d, n, m, alpha, _ptr(vals),
                                 np.ascontiguousarray(rowptr, np.int64),
                                 np.ascontiguousarray(colidxs, np.int64), _ptr(B), ldb, beta,
                                 _ptr(Cbuf), ldc, nthreads)
    _check(rc, "left_spmm_csr")


def sddmm_csr(dtype, m, d, alpha, rowptr, colidxs, Bbar, brs, bcs, Sm, srs, scs, nthreads=0):
    """dvals[k] = alpha * sum_j Bbar[i_k, j] * Sm[c_k, j] (returns a new array)."""
    dt = np.float64 if dtype == "d" else np.float32
    out = np.zeros(int(rowptr[-1]), dt)
    rc = lib().orc_sddmm_csr(_b(dtype), m, d, alpha, np.ascontiguousarray(rowptr, np.int64),
                             np.ascontiguousarray(colidxs, np.int64), _ptr(Bbar), brs, bcs, _ptr(Sm), srs,
                             scs, _ptr(out), nthreads)
    _check(rc, "sddmm_csr")
    return out


# ---------------------------------------------------------------- index sampling
def sample_indices_iid_uniform(n, k, counter, key, signed=False):
    """(samples int64[k], rademachers float64[k] or None, next counter)."""
    s = np.zeros(max(k, 1), np.int64)
    r = np.zeros(max(k, 1), np.float64) if signed else None
    nxt = np.zeros(4, np.uint32)
    lib().orc_sample_indices_iid_uniform(n, k, s, r.ctypes.data if signed else None, _ctr(counter), _key(key), nxt)
    return s[:k], (r[:k] if signed else None), nxt


def sample_indices_iid(n, cdf, k, counter, key):
    """(samples int64[k], next counter); cdf float32 or float64."""
    c = np.ascontiguousarray(cdf)
    dt = "d" if c.dtype == np.float64 else "s"
    s = np.zeros(max(k, 1), np.int64)
    nxt = np.zeros(4, np.uint32)
    lib().orc_sample_indices_iid(_b(dt), n, c.ctypes.data, k, s, _ctr(counter), _key(key), nxt)
    return s[:k], nxt


def weights_to_cdf(w, error_if_below=None):
    """(cdf array, status): status 0, 1 + i (w[i] below error_if_below) or -1."""
    a = np.ascontiguousarray(w).copy()
    dt = "d" if a.dtype == np.float64 else "s"
    if error_if_below is None:
        error_if_below = -float(np.sqrt(np.finfo(a.dtype).eps))
    st = lib().orc_weights_to_cdf(_b(dt), a.size, a.ctypes.data, error_if_below)
    return a, int(st)


def power_law_csr(m, n, k, exponent, counter, key, dtype=np.float32):
    """The library's power-law sparse test data restated from the oracle's own
    samplers: Zipf weights by libm pow (rounded to dtype), weights_to_cdf,
    k column draws by sample_indices_iid from `counter`, k row draws by
    sample_indices_iid_uniform from the counter after them, the (row, column)
    pairs sorted with repeats kept once, and N(0,1) values from a
    DenseDist(nnz, 1) at the counter after the row draws.
    Returns (rowptr, colidxs, vals, next_counter)."""
    dt = np.dtype(dtype)
    w = np.zeros(n, np.float64)
    lib().orc_power_law_weights(n, float(exponent), w.ctypes.data)
    cdf, st = weights_to_cdf(w.astype(dt), -1.0)
    assert st == 0, st
    cols, c1 = sample_indices_iid(n, cdf, k, counter, key)
    rows, _, c2 = sample_indices_iid_uniform(m, k, c1, key)
    keys = np.unique(rows.astype(np.int64) * n + cols.astype(np.int64))
    r, c = keys // n, keys % n
    rowptr = np.zeros(m + 1, np.int64)
    np.add.at(rowptr, r + 1, 1)
    rowptr = np.cumsum(rowptr)
    nnz = len(keys)
    if nnz == 0:
        return rowptr, c, np.zeros(0, dt), c2
    vals, nxt = fill_dense(nnz, 1, c2, key, "G", "L", dt)
    return rowptr, c, np.ascontiguousarray(vals[:, 0]), nxt


# ---------------------------------------------------------------- bfloat16 helpers
def bf16_round(x) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest, ties to even); float32 result."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def bf16_split(a, terms: int):
    """A = A_1 + ... + A_terms + residual with each A_t the bfloat16 rounding of
    the previous remainder (remainders are exact in float32)."""
    r = np.ascontiguousarray(a, dtype=np.float32)
    out = []
    for _ in range(terms):
        h = bf16_round(r)
        out.append(h)
        r = (r - h).astype(np.float32)
    return out
