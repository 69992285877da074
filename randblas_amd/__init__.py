"""randblas_amd -- MI355X (gfx950) sketching hot path of RandBLAS.

Python mirror of the reference's C++ API over the C ABI in
include/randblas_amd.h (librandblas_amd.so, loaded with ctypes).  Names,
argument order and error behaviour follow namespace RandBLAS:

    RNGState, DenseDist, DenseSkOp, SparseDist, SparseSkOp, ScalarDist, Axis,
    fill_dense, fill_dense_unpacked, fill_sparse, lskge3, rskge3, lskges,
    rskges, rsksp3, sketch_general, sketch_vector, sketch_symmetric

Matrices are torch tensors on the GPU, passed as flat buffers with explicit
leading dimensions exactly like the reference's raw pointers; torch only
provides device memory and the current HIP stream.  Errors raised by the
library come back as randblas_amd.Error (RandBLAS::Error) or OverflowError
(std::overflow_error).  There is no CPU fallback: if the shared library is
missing every call raises.
"""
from __future__ import annotations

import ctypes as C
import enum
import math
import os
from dataclasses import dataclass, field
from typing import Optional, Sequence

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "librandblas_amd.so")


class Error(RuntimeError):
    """RandBLAS::Error (RandBLAS/exceptions.hh:57-95)."""


class Unsupported(Error):
    pass


# ---------------------------------------------------------------- enums
class Layout(str, enum.Enum):
    ColMajor = "C"
    RowMajor = "R"


class Op(str, enum.Enum):
    NoTrans = "N"
    Trans = "T"


class ScalarDist(str, enum.Enum):
    Gaussian = "G"
    Uniform = "U"


class Axis(str, enum.Enum):
    Short = "S"
    Long = "L"


def _c(x) -> bytes:
    if isinstance(x, enum.Enum):
        x = x.value
    return str(x).encode()


def _flip(op) -> str:
    return "T" if _c(op) == b"N" else "N"


# ---------------------------------------------------------------- C structs
class c_rng_state(C.Structure):
    _fields_ = [("counter", C.c_uint32 * 4), ("key", C.c_uint32 * 2)]


class c_dense_dist(C.Structure):
    _fields_ = [("n_rows", C.c_int64), ("n_cols", C.c_int64), ("family", C.c_char),
                ("major_axis", C.c_char)]


class c_dense_skop(C.Structure):
    _fields_ = [("dist", c_dense_dist), ("seed_state", c_rng_state), ("next_state", c_rng_state),
                ("buff", C.c_void_p), ("layout", C.c_char), ("own_memory", C.c_int)]


class c_sparse_dist(C.Structure):
    _fields_ = [("n_rows", C.c_int64), ("n_cols", C.c_int64), ("vec_nnz", C.c_int64),
                ("major_axis", C.c_char)]


class c_sparse_skop(C.Structure):
    _fields_ = [("dist", c_sparse_dist), ("seed_state", c_rng_state), ("next_state", c_rng_state),
                ("nnz", C.c_int64), ("vals", C.c_void_p), ("rows", C.c_void_p),
                ("cols", C.c_void_p), ("own_memory", C.c_int), ("prep", C.c_void_p)]


class c_csr_matrix(C.Structure):
    _fields_ = [("n_rows", C.c_int64), ("n_cols", C.c_int64), ("nnz", C.c_int64),
                ("vals", C.c_void_p), ("rowptr", C.c_void_p), ("colidxs", C.c_void_p)]


class c_csc_matrix(C.Structure):
    _fields_ = [("n_rows", C.c_int64), ("n_cols", C.c_int64), ("nnz", C.c_int64),
                ("vals", C.c_void_p), ("rowidxs", C.c_void_p), ("colptr", C.c_void_p)]


class c_coo_matrix(C.Structure):
    _fields_ = [("n_rows", C.c_int64), ("n_cols", C.c_int64), ("nnz", C.c_int64),
                ("vals", C.c_void_p), ("rows", C.c_void_p), ("cols", C.c_void_p)]


RBX_OK, RBX_INVALID_ARG, RBX_OVERFLOW, RBX_HIP_ERROR, RBX_UNSUPPORTED, RBX_OUT_OF_MEMORY = range(6)

# symbol -> argtypes (restype is always rbx_status unless listed in _RESTYPE)
_I64, _F, _D, _P, _CH = C.c_int64, C.c_float, C.c_double, C.c_void_p, C.c_char
_RS, _DD, _DS = C.POINTER(c_rng_state), C.POINTER(c_dense_dist), C.POINTER(c_dense_skop)
_SD, _SS, _CSR = C.POINTER(c_sparse_dist), C.POINTER(c_sparse_skop), C.POINTER(c_csr_matrix)
_CSC, _COO = C.POINTER(c_csc_matrix), C.POINTER(c_coo_matrix)
_SIG = {
    "rbx_last_error_message": [],
    "rbx_version": [],
    "rbx_philox4x32_10": [_RS, _I64, _P, _P],
    "rbx_rng_state_from_key": [C.c_uint64, _RS],
    "rbx_rng_state_incr": [_RS, C.c_uint64],
    "rbx_dense_skop_init": [_DD, _RS, _DS],
    "rbx_dense_dist_props": [_DD, C.POINTER(_I64), C.POINTER(_I64), C.c_char_p, C.POINTER(_D)],
    "rbx_sfill_dense_unpacked": [_CH, _DD, _I64, _I64, _I64, _I64, _P, _RS, _RS, _P],
    "rbx_dfill_dense_unpacked": [_CH, _DD, _I64, _I64, _I64, _I64, _P, _RS, _RS, _P],
    "rbx_sfill_dense_skop": [_DS, _P],
    "rbx_dfill_dense_skop": [_DS, _P],
    "rbx_dense_skop_free": [_DS],
    "rbx_slskge3": [_CH, _CH, _CH, _I64, _I64, _I64, _F, _DS, _I64, _I64, _P, _I64, _F, _P, _I64, _P],
    "rbx_dlskge3": [_CH, _CH, _CH, _I64, _I64, _I64, _D, _DS, _I64, _I64, _P, _I64, _D, _P, _I64, _P],
    "rbx_srskge3": [_CH, _CH, _CH, _I64, _I64, _I64, _F, _P, _I64, _DS, _I64, _I64, _F, _P, _I64, _P],
    "rbx_drskge3": [_CH, _CH, _CH, _I64, _I64, _I64, _D, _P, _I64, _DS, _I64, _I64, _D, _P, _I64, _P],
    "rbx_slskge3_adjoint": [_CH, _CH, _I64, _I64, _I64, _F, _DS, _I64, _I64, _P, _I64, _F, _P, _I64, _P],
    "rbx_dlskge3_adjoint": [_CH, _CH, _I64, _I64, _I64, _D, _DS, _I64, _I64, _P, _I64, _D, _P, _I64, _P],
    "rbx_srskge3_adjoint": [_CH, _CH, _I64, _I64, _I64, _F, _P, _I64, _DS, _I64, _I64, _F, _P, _I64, _P],
    "rbx_drskge3_adjoint": [_CH, _CH, _I64, _I64, _I64, _D, _P, _I64, _DS, _I64, _I64, _D, _P, _I64, _P],
    "rbx_slskge3_bf16split": [_CH, _CH, _CH, _I64, _I64, _I64, _F, _DS, _I64, _I64, _P, _I64, _F, _P,
                              _I64, C.c_int, _P],
    "rbx_sparse_skop_init": [_SD, _RS, _SS],
    "rbx_sparse_dist_props": [_SD, C.POINTER(_I64), C.POINTER(_I64), C.POINTER(_I64), C.POINTER(_D)],
    "rbx_sfill_sparse_unpacked_nosub": [_SD, C.POINTER(_I64), _P, _P, _P, _RS, _RS, _P],
    "rbx_dfill_sparse_unpacked_nosub": [_SD, C.POINTER(_I64), _P, _P, _P, _RS, _RS, _P],
    "rbx_sfill_sparse_skop": [_SS, _P],
    "rbx_dfill_sparse_skop": [_SS, _P],
    "rbx_sparse_skop_free": [_SS],
    "rbx_ssparse_skop_prepare": [_SS, _P],
    "rbx_dsparse_skop_prepare": [_SS, _P],
    "rbx_slskges": [_CH, _CH, _CH, _I64, _I64, _I64, _F, _SS, _I64, _I64, _P, _I64, _F, _P, _I64, _P],
    "rbx_dlskges": [_CH, _CH, _CH, _I64, _I64, _I64, _D, _SS, _I64, _I64, _P, _I64, _D, _P, _I64, _P],
    "rbx_srskges": [_CH, _CH, _CH, _I64, _I64, _I64, _F, _P, _I64, _SS, _I64, _I64, _F, _P, _I64, _P],
    "rbx_drskges": [_CH, _CH, _CH, _I64, _I64, _I64, _D, _P, _I64, _SS, _I64, _I64, _D, _P, _I64, _P],
    "rbx_slskges_noalloc": [_CH, _CH, _CH, _I64, _I64, _I64, _F, _SS, _I64, _I64, _P, _I64, _F, _P, _I64, _P],
    "rbx_dlskges_noalloc": [_CH, _CH, _CH, _I64, _I64, _I64, _D, _SS, _I64, _I64, _P, _I64, _D, _P, _I64, _P],
    "rbx_lskges_noalloc_scratch_bytes": [],
    "rbx_srsksp3": [_CH, _CH, _CH, _I64, _I64, _I64, _F, _CSR, _I64, _I64, _DS, _I64, _I64, _F, _P,
                    _I64, _P],
    "rbx_drsksp3": [_CH, _CH, _CH, _I64, _I64, _I64, _D, _CSR, _I64, _I64, _DS, _I64, _I64, _D, _P,
                    _I64, _P],
    "rbx_slsksp3": [_CH, _CH, _CH, _I64, _I64, _I64, _F, _DS, _I64, _I64, _CSR, _I64, _I64, _F, _P, _I64, _P],
    "rbx_dlsksp3": [_CH, _CH, _CH, _I64, _I64, _I64, _D, _DS, _I64, _I64, _CSR, _I64, _I64, _D, _P, _I64, _P],
    **{f"rbx_{p}rsksp3{sfx}": [_CH, _CH, _CH, _I64, _I64, _I64, sc, mat, _I64, _I64, _DS, _I64, _I64, sc, _P, _I64, _P]
       for p, sc in (("s", _F), ("d", _D)) for sfx, mat in (("_csc", _CSC), ("_coo", _COO))},
    **{f"rbx_{p}lsksp3{sfx}": [_CH, _CH, _CH, _I64, _I64, _I64, sc, _DS, _I64, _I64, mat, _I64, _I64, sc, _P, _I64, _P]
       for p, sc in (("s", _F), ("d", _D)) for sfx, mat in (("_csc", _CSC), ("_coo", _COO))},
    "rbx_srsksp3_adjoint": [_CH, _CH, _I64, _I64, _I64, _F, _CSR, _DS, _I64, _I64, _P, _I64, _P, _P],
    "rbx_drsksp3_adjoint": [_CH, _CH, _I64, _I64, _I64, _D, _CSR, _DS, _I64, _I64, _P, _I64, _P, _P],
    "rbx_srequire_symmetric": [_CH, _P, _I64, _I64, _F, _P],
    **{f"rbx_{p}{a}_to_{b}": [sa, sb, _P]
       for p in ("s", "d")
       for a, b, sa, sb in (("csr", "coo", _CSR, _COO), ("csc", "coo", _CSC, _COO), ("coo", "csr", _COO, _CSR),
                            ("coo", "csc", _COO, _CSC))},
    "rbx_scoo_sort": [_CH, _COO, _P],
    "rbx_dcoo_sort": [_CH, _COO, _P],
    "rbx_coo_determine_sort": [_COO, C.POINTER(C.c_char), _P],
    "rbx_sample_indices_iid_uniform": [_I64, _I64, _P, _RS, _RS, _P],
    "rbx_ssample_indices_iid_uniform_signed": [_I64, _I64, _P, _P, _RS, _RS, _P],
    "rbx_dsample_indices_iid_uniform_signed": [_I64, _I64, _P, _P, _RS, _RS, _P],
    "rbx_ssample_indices_iid": [_I64, _P, _I64, _P, _RS, _RS, _P],
    "rbx_dsample_indices_iid": [_I64, _P, _I64, _P, _RS, _RS, _P],
    "rbx_sweights_to_cdf": [_I64, _P, _F, _P],
    "rbx_spower_law_csr": [_I64, _I64, _I64, _D, _RS, C.POINTER(c_csr_matrix), _RS, _P],
    "rbx_dpower_law_csr": [_I64, _I64, _I64, _D, _RS, C.POINTER(c_csr_matrix), _RS, _P],
    "rbx_dweights_to_cdf": [_I64, _P, _D, _P],
    "rbx_drequire_symmetric": [_CH, _P, _I64, _I64, _D, _P],
    **{f"rbx_{p}sketch_solve": [_CH, _I64, _I64, _I64, _DS, _I64, _I64, _P, _I64, _P, _P, _P] for p in ("s", "d")},
    **{f"rbx_{p}sketch_solve_adjoint": [_CH, _I64, _I64, _I64, _DS, _I64, _I64, _P, _I64, _P, _P, _P, _I64, _P, _P]
       for p in ("s", "d")},
    "rbx_bf16_fastgen_bounds": [_P, _P, _P],
    "rbx_srsksp3_bf16split": [_CH, _CH, _CH, _I64, _I64, _I64, _F, _CSR, _I64, _I64, _DS, _I64, _I64, _F, _P, _I64,
                              C.c_int, _P],
    "rbx_release_workspace": [_P],
    "rbx_release_all_workspaces": [],
    "rbx_profile_enable": [C.c_int],
    "rbx_profile_collect": [C.c_char_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double), C.c_int],
}
_RESTYPE = {"rbx_last_error_message": C.c_char_p, "rbx_version": C.c_char_p, "rbx_profile_collect": C.c_int,
            "rbx_lskges_noalloc_scratch_bytes": C.c_size_t}

_lib: Optional[C.CDLL] = None


def lib() -> C.CDLL:
    """Load librandblas_amd.so.  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise Error(f"{LIB_PATH} is missing: build it with `python -m randblas_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, args in _SIG.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _RESTYPE.get(name, C.c_int)
        _lib = L
    return _lib


def _call(name: str, *args) -> None:
    rc = getattr(lib(), name)(*args)
    if rc != RBX_OK:
        msg = lib().rbx_last_error_message().decode()
        if rc == RBX_OVERFLOW:
            raise OverflowError(msg)
        if rc == RBX_UNSUPPORTED:
            raise Unsupported(msg)
        raise Error(f"{name} -> status {rc}: {msg}")


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t) -> C.c_void_p:
    if t is None:
        return C.c_void_p(0)
    if isinstance(t, int):
        return C.c_void_p(t)
    if not t.is_cuda:
        raise Error("randblas_amd operates on GPU tensors (device pointers)")
    return C.c_void_p(t.data_ptr())


def _prefix(t) -> str:
    import torch
    if not isinstance(t, torch.Tensor):
        raise Error(f"expected a torch tensor, got {type(t).__name__}")
    if t.dtype == torch.float64:
        return "d"
    if t.dtype == torch.float32:
        return "s"
    raise Error(f"unsupported dtype {t.dtype}")


# ---------------------------------------------------------------- operand checks
# The C ABI takes raw device pointers, so a wrong dtype or a short buffer
# would be read as other bits or out of bounds.  Every wrapper checks its
# operands here, before any pointer reaches the library (the reference's
# typed DenseSkOp<T> / SparseSkOp<T> rule out the dtype mismatch at compile
# time).
def _extent(rows, cols, layout, ld) -> int:
    """Elements spanned by a rows x cols matrix stored in `layout` with
    leading dimension ld (0 for an empty matrix)."""
    rows, cols, ld = int(rows), int(cols), int(ld)
    if rows <= 0 or cols <= 0:
        return 0
    return (cols - 1) * ld + rows if _c(layout) == b"C" else (rows - 1) * ld + cols


def _need(name, t, dtype, numel) -> None:
    """t must be a contiguous CUDA tensor of `dtype` (a dtype or a tuple of
    them) holding at least `numel` elements."""
    import torch
    numel = int(numel)
    if t is None:
        if numel > 0:
            raise Error(f"{name}: a buffer of {numel} elements is required")
        return
    if not isinstance(t, torch.Tensor):
        raise Error(f"{name}: expected a torch tensor, got {type(t).__name__}")
    allowed = dtype if isinstance(dtype, tuple) else (dtype,)
    if t.dtype not in allowed:
        raise Error(f"{name}: dtype {t.dtype}, expected {' or '.join(str(d) for d in allowed)}")
    if not t.is_contiguous():
        raise Error(f"{name}: the tensor must be contiguous")
    if t.numel() < numel:
        raise Error(f"{name}: {t.numel()} elements, the call addresses {numel}")
    # the device itself is checked by _ptr, when the pointer is taken


def _fp(prefix):
    import torch
    return torch.float64 if prefix == "d" else torch.float32


def _i64():
    import torch
    return torch.int64


def _need_dense_op(S: "DenseSkOp", prefix: str) -> None:
    if S.buff is not None:
        _need("S.buff", S.buff, _fp(prefix), S.n_rows * S.n_cols)


def _need_sparse_op(S: "SparseSkOp", prefix: str) -> None:
    if S.vals is None or S.nnz < 0:
        return  # unsampled: the library samples it into its own scratch
    _need("S.vals", S.vals, _fp(prefix), S.nnz)
    _need("S.rows", S.rows, _i64(), S.nnz)
    _need("S.cols", S.cols, _i64(), S.nnz)


def _opdims(op, r, c):
    """(rows, cols) of the stored matrix whose op() is r x c."""
    return (r, c) if _c(op) == b"N" else (c, r)


# ---------------------------------------------------------------- RNGState
@dataclass
class RNGState:
    """RNGState<Philox4x32> (RandBLAS/base.hh:64-164).  RNGState(k) sets the key
    to {k mod 2^32, k >> 32} and the counter to zero."""
    key: int = 0
    counter: Sequence[int] = (0, 0, 0, 0)

    def c(self) -> c_rng_state:
        s = c_rng_state()
        k = int(self.key)
        s.key[0], s.key[1] = k & 0xFFFFFFFF, (k >> 32) & 0xFFFFFFFF
        for i in range(4):
            s.counter[i] = int(self.counter[i]) & 0xFFFFFFFF
        return s

    @staticmethod
    def from_c(s: c_rng_state) -> "RNGState":
        return RNGState(key=int(s.key[0]) | (int(s.key[1]) << 32), counter=tuple(int(x) for x in s.counter))

    @property
    def key_words(self):
        k = int(self.key)
        return (k & 0xFFFFFFFF, k >> 32)

    def incr(self, n: int = 1) -> "RNGState":
        s = self.c()
        _call("rbx_rng_state_incr", C.byref(s), n)
        return RNGState.from_c(s)

    def counter_int(self) -> int:
        return sum(int(w) << (32 * i) for i, w in enumerate(self.counter))


def philox4x32_10(state: RNGState, n_counters: int, out) -> None:
    """Raw words for counters state.counter + i into a uint32/int32 GPU tensor of 4*n words."""
    import torch
    _need("out", out, (torch.int32, torch.uint32), 4 * n_counters)
    s = state.c()
    _call("rbx_philox4x32_10", C.byref(s), n_counters, _ptr(out), _stream())


# ---------------------------------------------------------------- dense
@dataclass
class DenseDist:
    """DenseDist (RandBLAS/dense_skops.hh:231-350)."""
    n_rows: int
    n_cols: int
    family: ScalarDist = ScalarDist.Gaussian
    major_axis: Axis = Axis.Long

    def c(self) -> c_dense_dist:
        return c_dense_dist(self.n_rows, self.n_cols, _c(self.family), _c(self.major_axis))

    def _props(self):
        dmaj, dmin, iso = C.c_int64(), C.c_int64(), C.c_double()
        lay = C.create_string_buffer(1)
        _call("rbx_dense_dist_props", C.byref(self.c()), C.byref(dmaj), C.byref(dmin), lay, C.byref(iso))
        return dmaj.value, dmin.value, lay.raw.decode(), iso.value

    @property
    def dim_major(self):
        return self._props()[0]

    @property
    def dim_minor(self):
        return self._props()[1]

    @property
    def natural_layout(self) -> str:
        return self._props()[2]

    @property
    def isometry_scale(self) -> float:
        return self._props()[3]


class DenseSkOp:
    """DenseSkOp<T> (RandBLAS/dense_skops.hh:357-478).  buff is None until
    fill_dense(S) materialises it; sketches with buff None generate the operator
    inside the GEMM kernels.  dtype plays the role of T: it is the dtype of
    buff (float64 when neither is given), and a materialised operator can only
    be applied to matrices of that dtype."""

    def __init__(self, dist: DenseDist, seed_state: RNGState, buff=None, layout: Optional[str] = None,
                 dtype=None):
        import torch
        if buff is not None and dtype is not None and buff.dtype != dtype:
            raise Error(f"DenseSkOp: buff has dtype {buff.dtype}, dtype says {dtype}")
        self.dtype = buff.dtype if buff is not None else (dtype or torch.float64)
        self._dtype_given = buff is not None or dtype is not None
        self.dist = dist
        self.seed_state = seed_state
        out = c_dense_skop()
        _call("rbx_dense_skop_init", C.byref(dist.c()), C.byref(seed_state.c()), C.byref(out))
        self.next_state = RNGState.from_c(out.next_state)
        self.n_rows, self.n_cols = dist.n_rows, dist.n_cols
        self.layout = layout if layout is not None else out.layout.decode()
        self.buff = buff

    def c(self) -> c_dense_skop:
        s = c_dense_skop()
        s.dist = self.dist.c()
        s.seed_state = self.seed_state.c()
        s.next_state = self.next_state.c()
        s.buff = None if self.buff is None else self.buff.data_ptr()
        s.layout = _c(self.layout)
        s.own_memory = 0
        return s


def fill_dense_unpacked(layout, D: DenseDist, n_rows, n_cols, ro_s, co_s, buff, seed: RNGState) -> RNGState:
    """fill_dense_unpacked (dense_skops.hh:563-606); buff is a GPU tensor."""
    nxt = c_rng_state()
    _need("buff", buff, _fp(_prefix(buff)), int(n_rows) * int(n_cols))
    _call(f"rbx_{_prefix(buff)}fill_dense_unpacked", _c(layout), C.byref(D.c()), n_rows, n_cols, ro_s, co_s,
          _ptr(buff), C.byref(seed.c()), C.byref(nxt), _stream())
    return RNGState.from_c(nxt)


def fill_dense(D_or_S, buff=None, seed: Optional[RNGState] = None, dtype=None):
    """fill_dense(D, buff, seed) -> next state, or fill_dense(S) which attaches
    a torch buffer to S in its natural layout (dense_skops.hh:623-658)."""
    import torch
    if isinstance(D_or_S, DenseSkOp):
        S = D_or_S
        if dtype is not None and not S._dtype_given:
            S.dtype = dtype
        if dtype is not None and dtype != S.dtype:
            raise Error(f"fill_dense: the operator holds {S.dtype} entries, not {dtype}")
        if S.buff is None:
            S.buff = torch.empty(S.n_rows * S.n_cols, dtype=S.dtype, device="cuda")
        fill_dense_unpacked(S.layout, S.dist, S.n_rows, S.n_cols, 0, 0, S.buff, S.seed_state)
        return None
    D = D_or_S
    return fill_dense_unpacked(D.natural_layout, D, D.n_rows, D.n_cols, 0, 0, buff, seed)


def lskge3(layout, opS, opA, d, n, m, alpha, S: DenseSkOp, ro_s, co_s, A, lda, beta, B, ldb):
    """lskge3 (skge.hh:154-202)."""
    p = _prefix(B)
    _need_dense_op(S, p)
    if alpha != 0:
        _need("A", A, _fp(p), _extent(*_opdims(opA, m, n), layout, lda))
    _need("B", B, _fp(p), _extent(d, n, layout, ldb))
    _call(f"rbx_{_prefix(B)}lskge3", _c(layout), _c(opS), _c(opA), d, n, m, alpha, C.byref(S.c()), ro_s, co_s,
          _ptr(A), lda, beta, _ptr(B), ldb, _stream())


def rskge3(layout, opA, opS, m, d, n, alpha, A, lda, S: DenseSkOp, ro_s, co_s, beta, B, ldb):
    """rskge3 (skge.hh:307-355)."""
    p = _prefix(B)
    _need_dense_op(S, p)
    if alpha != 0:
        _need("A", A, _fp(p), _extent(*_opdims(opA, m, n), layout, lda))
    _need("B", B, _fp(p), _extent(m, d, layout, ldb))
    _call(f"rbx_{_prefix(B)}rskge3", _c(layout), _c(opA), _c(opS), m, d, n, alpha, _ptr(A), lda, C.byref(S.c()),
          ro_s, co_s, beta, _ptr(B), ldb, _stream())


def lskge3_adjoint(layout, opS, d, n, m, alpha, S: DenseSkOp, ro_s, co_s, Bbar, ldbb, beta, Abar, ldab):
    """Abar = alpha * op(submat(S))^T * Bbar + beta * Abar (operator regenerated)."""
    p = _prefix(Abar)
    _need_dense_op(S, p)
    if alpha != 0:
        _need("Bbar", Bbar, _fp(p), _extent(d, n, layout, ldbb))
    _need("Abar", Abar, _fp(p), _extent(m, n, layout, ldab))
    _call(f"rbx_{_prefix(Abar)}lskge3_adjoint", _c(layout), _c(opS), d, n, m, alpha, C.byref(S.c()), ro_s, co_s,
          _ptr(Bbar), ldbb, beta, _ptr(Abar), ldab, _stream())


def rskge3_adjoint(layout, opS, m, d, n, alpha, Bbar, ldbb, S: DenseSkOp, ro_s, co_s, beta, Abar, ldab):
    """Abar = alpha * Bbar * op(submat(S))^T + beta * Abar (operator regenerated)."""
    p = _prefix(Abar)
    _need_dense_op(S, p)
    if alpha != 0:
        _need("Bbar", Bbar, _fp(p), _extent(m, d, layout, ldbb))
    _need("Abar", Abar, _fp(p), _extent(m, n, layout, ldab))
    _call(f"rbx_{_prefix(Abar)}rskge3_adjoint", _c(layout), _c(opS), m, d, n, alpha, _ptr(Bbar), ldbb,
          C.byref(S.c()), ro_s, co_s, beta, _ptr(Abar), ldab, _stream())


def lskge3_bf16split(layout, opS, opA, d, n, m, alpha, S: DenseSkOp, ro_s, co_s, A, lda, beta, B, ldb,
                     n_terms: int = 3):
    """B = alpha * op(bf16(submat S)) * op(A) + beta * B on the BF16 matrix cores."""
    import torch
    if alpha != 0:
        _need("A", A, torch.float32, _extent(*_opdims(opA, m, n), layout, lda))
    _need("B", B, torch.float32, _extent(d, n, layout, ldb))
    _call("rbx_slskge3_bf16split", _c(layout), _c(opS), _c(opA), d, n, m, alpha, C.byref(S.c()), ro_s, co_s,
          _ptr(A), lda, beta, _ptr(B), ldb, n_terms, _stream())


# ---------------------------------------------------------------- sparse
@dataclass
class SparseDist:
    """SparseDist (RandBLAS/sparse_skops.hh:131-246)."""
    n_rows: int
    n_cols: int
    vec_nnz: int = 4
    major_axis: Axis = Axis.Short

    def c(self) -> c_sparse_dist:
        return c_sparse_dist(self.n_rows, self.n_cols, self.vec_nnz, _c(self.major_axis))

    def _props(self):
        a, b, f, iso = C.c_int64(), C.c_int64(), C.c_int64(), C.c_double()
        _call("rbx_sparse_dist_props", C.byref(self.c()), C.byref(a), C.byref(b), C.byref(f), C.byref(iso))
        return a.value, b.value, f.value, iso.value

    @property
    def dim_major(self):
        return self._props()[0]

    @property
    def dim_minor(self):
        return self._props()[1]

    @property
    def full_nnz(self):
        return self._props()[2]

    @property
    def isometry_scale(self):
        return self._props()[3]


class SparseSkOp:
    """SparseSkOp<T, Philox4x32, int64_t> (RandBLAS/sparse_skops.hh:289-450)."""

    def __init__(self, dist: SparseDist, seed_state: RNGState, dtype=None):
        import torch
        self.dist = dist
        self.seed_state = seed_state
        out = c_sparse_skop()
        _call("rbx_sparse_skop_init", C.byref(dist.c()), C.byref(seed_state.c()), C.byref(out))
        self.next_state = RNGState.from_c(out.next_state)
        self.n_rows, self.n_cols = dist.n_rows, dist.n_cols
        self.nnz = -1
        self.dtype = dtype or torch.float64
        self.vals = self.rows = self.cols = None
        self._prep = None  # the library's prepared copy of the sampled entries

    def prepare(self) -> None:
        """Rebuild the library's prepared copy of the entries (rbx_[sd]sparse_skop_prepare):
        lskges over the whole operator then skips its per-call sort.  fill_sparse
        calls it; call it again after writing vals / rows / cols in place."""
        import torch
        self.release()
        s = self.c()
        pre = "d" if self.dtype == torch.float64 else "s"
        try:
            _call(f"rbx_{pre}sparse_skop_prepare", C.byref(s), _stream())
        finally:
            self._prep = s.prep
            self._prep_key = self._key()

    def release(self) -> None:
        """Free the prepared copy (rbx_sparse_skop_free on arrays this object does not own)."""
        prep = getattr(self, "_prep", None)
        if prep:
            s = self.c()
            s.own_memory = 0
            s.prep = prep
            self._prep = None
            lib().rbx_sparse_skop_free(C.byref(s))

    def _key(self):
        return (self.nnz, None if self.vals is None else self.vals.data_ptr(),
                None if self.rows is None else self.rows.data_ptr(),
                None if self.cols is None else self.cols.data_ptr())

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass

    def c(self) -> c_sparse_skop:
        s = c_sparse_skop()
        s.dist = self.dist.c()
        s.seed_state = self.seed_state.c()
        s.next_state = self.next_state.c()
        s.nnz = self.nnz
        s.vals = None if self.vals is None else self.vals.data_ptr()
        s.rows = None if self.rows is None else self.rows.data_ptr()
        s.cols = None if self.cols is None else self.cols.data_ptr()
        s.own_memory = 0
        # the prepared copy goes along only while the arrays are the ones it was built from
        prep = getattr(self, "_prep", None)
        s.prep = prep if prep and getattr(self, "_prep_key", None) == self._key() else None
        return s


def fill_sparse(S: SparseSkOp) -> None:
    """fill_sparse(S) (sparse_skops.hh:587-603) into torch buffers."""
    import torch
    full = S.dist.full_nnz
    if S.vals is None:
        S.vals = torch.empty(full, dtype=S.dtype, device="cuda")
        S.rows = torch.empty(full, dtype=torch.int64, device="cuda")
        S.cols = torch.empty(full, dtype=torch.int64, device="cuda")
    nnz = C.c_int64()
    nxt = c_rng_state()
    pre = "d" if S.dtype == torch.float64 else "s"
    _need("S.vals", S.vals, S.dtype, full)
    _need("S.rows", S.rows, torch.int64, full)
    _need("S.cols", S.cols, torch.int64, full)
    S.release()
    _call(f"rbx_{pre}fill_sparse_unpacked_nosub", C.byref(S.dist.c()), C.byref(nnz), _ptr(S.vals), _ptr(S.rows),
          _ptr(S.cols), C.byref(S.seed_state.c()), C.byref(nxt), _stream())
    S.nnz = nnz.value
    S.prepare()


def fill_sparse_unpacked_nosub(D: SparseDist, vals, rows, cols, seed: RNGState):
    """Returns (nnz, next_state)."""
    full = D.full_nnz
    _need("vals", vals, _fp(_prefix(vals)), full)
    _need("rows", rows, _i64(), full)
    _need("cols", cols, _i64(), full)
    nnz = C.c_int64()
    nxt = c_rng_state()
    _call(f"rbx_{_prefix(vals)}fill_sparse_unpacked_nosub", C.byref(D.c()), C.byref(nnz), _ptr(vals), _ptr(rows),
          _ptr(cols), C.byref(seed.c()), C.byref(nxt), _stream())
    return nnz.value, RNGState.from_c(nxt)


def lskges(layout, opS, opA, d, n, m, alpha, S: SparseSkOp, ro_s, co_s, A, lda, beta, B, ldb):
    """lskges (skge.hh:465-492)."""
    p = _prefix(B)
    if S.vals is not None and S.dtype != _fp(p):
        raise Error(f"lskges: the operator holds {S.dtype} values, B is {B.dtype}")
    _need_sparse_op(S, p)
    if alpha != 0:
        _need("A", A, _fp(p), _extent(*_opdims(opA, m, n), layout, lda))
    _need("B", B, _fp(p), _extent(d, n, layout, ldb))
    _call(f"rbx_{_prefix(B)}lskges", _c(layout), _c(opS), _c(opA), d, n, m, alpha, C.byref(S.c()), ro_s, co_s,
          _ptr(A), lda, beta, _ptr(B), ldb, _stream())


def lskges_noalloc(layout, opS, opA, d, n, m, alpha, S: SparseSkOp, ro_s, co_s, A, lda, beta, B, ldb):
    """lskges without device allocation (no reference counterpart): S must be
    filled, op(A) unit-stride along the contraction, and the operator window
    within the library's fixed scratch (lskges_noalloc_scratch_bytes())."""
    p = _prefix(B)
    if S.vals is None or S.nnz is None or S.nnz < 0:
        raise Error("lskges_noalloc: fill the operator first (fill_sparse)")
    if S.dtype != _fp(p):
        raise Error(f"lskges_noalloc: the operator holds {S.dtype} values, B is {B.dtype}")
    _need_sparse_op(S, p)
    if alpha != 0:
        _need("A", A, _fp(p), _extent(*_opdims(opA, m, n), layout, lda))
    _need("B", B, _fp(p), _extent(d, n, layout, ldb))
    _call(f"rbx_{p}lskges_noalloc", _c(layout), _c(opS), _c(opA), d, n, m, alpha, C.byref(S.c()), ro_s, co_s,
          _ptr(A), lda, beta, _ptr(B), ldb, _stream())


def lskges_noalloc_scratch_bytes() -> int:
    """Size of the fixed per-GPU scratch array behind lskges_noalloc."""
    return int(lib().rbx_lskges_noalloc_scratch_bytes())


def rskges(layout, opA, opS, m, d, n, alpha, A, lda, S: SparseSkOp, ro_s, co_s, beta, B, ldb):
    """rskges (skge.hh:598-626)."""
    p = _prefix(B)
    if S.vals is not None and S.dtype != _fp(p):
        raise Error(f"rskges: the operator holds {S.dtype} values, B is {B.dtype}")
    _need_sparse_op(S, p)
    if alpha != 0:
        _need("A", A, _fp(p), _extent(*_opdims(opA, m, n), layout, lda))
    _need("B", B, _fp(p), _extent(m, d, layout, ldb))
    _call(f"rbx_{_prefix(B)}rskges", _c(layout), _c(opA), _c(opS), m, d, n, alpha, _ptr(A), lda, C.byref(S.c()),
          ro_s, co_s, beta, _ptr(B), ldb, _stream())


@dataclass
class CSRMatrix:
    """CSRMatrix<T, int64_t> (sparse_data/csr_matrix.hh:73-245) over GPU tensors."""
    n_rows: int
    n_cols: int
    vals: object
    rowptr: object
    colidxs: object

    def c(self) -> c_csr_matrix:
        return c_csr_matrix(self.n_rows, self.n_cols, int(self.vals.numel()), self.vals.data_ptr(),
                            self.rowptr.data_ptr(), self.colidxs.data_ptr())


@dataclass
class CSCMatrix:
    """CSCMatrix<T, int64_t> (sparse_data/csc_matrix.hh:74-244) over GPU tensors."""
    n_rows: int
    n_cols: int
    vals: object
    rowidxs: object
    colptr: object

    def c(self) -> c_csc_matrix:
        return c_csc_matrix(self.n_rows, self.n_cols, int(self.vals.numel()), self.vals.data_ptr(),
                            self.rowidxs.data_ptr(), self.colptr.data_ptr())


@dataclass
class COOMatrix:
    """COOMatrix<T, int64_t> (sparse_data/coo_matrix.hh:83-308) over GPU tensors;
    entries in any order."""
    n_rows: int
    n_cols: int
    vals: object
    rows: object
    cols: object

    def c(self) -> c_coo_matrix:
        return c_coo_matrix(self.n_rows, self.n_cols, int(self.vals.numel()), self.vals.data_ptr(),
                            self.rows.data_ptr(), self.cols.data_ptr())


def _need_sparse_data(A, prefix: str, name: str = "A") -> None:
    """Arrays of a CSR / CSC / COO matrix: values of the call's precision,
    int64 indices, pointer arrays of n + 1 entries."""
    nnz = int(A.vals.numel()) if A.vals is not None else 0
    _need(f"{name}.vals", A.vals, _fp(prefix), nnz)
    if isinstance(A, CSRMatrix):
        _need(f"{name}.rowptr", A.rowptr, _i64(), A.n_rows + 1)
        _need(f"{name}.colidxs", A.colidxs, _i64(), nnz)
    elif isinstance(A, CSCMatrix):
        _need(f"{name}.colptr", A.colptr, _i64(), A.n_cols + 1)
        _need(f"{name}.rowidxs", A.rowidxs, _i64(), nnz)
    else:
        _need(f"{name}.rows", A.rows, _i64(), nnz)
        _need(f"{name}.cols", A.cols, _i64(), nnz)


def _sp_suffix(A) -> str:
    if isinstance(A, CSRMatrix):
        return ""
    if isinstance(A, CSCMatrix):
        return "_csc"
    if isinstance(A, COOMatrix):
        return "_coo"
    raise TypeError(f"expected CSRMatrix, CSCMatrix or COOMatrix, got {type(A).__name__}")


def rsksp3(layout, opA, opS, m, d, n, alpha, A, ro_a, co_a, S: DenseSkOp, ro_s, co_s, beta, B, ldb):
    """rsksp3 (sparse_data/sksp.hh:277-326) for a CSRMatrix, CSCMatrix or COOMatrix A."""
    p = _prefix(B)
    _sp_suffix(A)
    _need_sparse_data(A, p)
    _need_dense_op(S, p)
    _need("B", B, _fp(p), _extent(m, d, layout, ldb))
    _call(f"rbx_{_prefix(B)}rsksp3{_sp_suffix(A)}", _c(layout), _c(opA), _c(opS), m, d, n, alpha, C.byref(A.c()),
          ro_a, co_a, C.byref(S.c()), ro_s, co_s, beta, _ptr(B), ldb, _stream())


def rsksp3_bf16split(layout, opA, opS, m, d, n, alpha, A: "CSRMatrix", ro_a, co_a, S: DenseSkOp, ro_s, co_s, beta,
                     B, ldb, n_terms: int = 3):
    """B = alpha * A_T * op(bf16(submat S)) + beta * B for FP32 CSR data: the
    operator rounded to bfloat16, A's values split into n_terms bfloat16 terms."""
    import torch
    if not isinstance(A, CSRMatrix):
        raise TypeError(f"rsksp3_bf16split: expected a CSRMatrix, got {type(A).__name__}")
    _need_sparse_data(A, "s")
    if S.buff is not None:
        _need("S.buff", S.buff, torch.float32, S.n_rows * S.n_cols)
    _need("B", B, torch.float32, _extent(m, d, layout, ldb))
    _call("rbx_srsksp3_bf16split", _c(layout), _c(opA), _c(opS), m, d, n, alpha, C.byref(A.c()), ro_a, co_a,
          C.byref(S.c()), ro_s, co_s, beta, _ptr(B), ldb, n_terms, _stream())


# ---------------------------------------------------------------- dispatching wrappers
def lsksp3(layout, opS, opA, d, n, m, alpha, S: DenseSkOp, ro_s, co_s, A, ro_a, co_a, beta, B, ldb):
    """lsksp3 (sparse_data/sksp.hh:132-180): B = alpha * op(submat(S)) * op(A) + beta * B,
    A a CSRMatrix, CSCMatrix or COOMatrix."""
    p = _prefix(B)
    _sp_suffix(A)
    _need_sparse_data(A, p)
    _need_dense_op(S, p)
    _need("B", B, _fp(p), _extent(d, n, layout, ldb))
    _call(f"rbx_{_prefix(B)}lsksp3{_sp_suffix(A)}", _c(layout), _c(opS), _c(opA), d, n, m, alpha, C.byref(S.c()),
          ro_s, co_s, C.byref(A.c()), ro_a, co_a, beta, _ptr(B), ldb, _stream())


def sketch_sparse(layout, op1, op2, a1, a2, a3, alpha, X, *rest):
    """sketch_sparse (sparse_data/sksp.hh:418-539):

      sketch_sparse(layout, opS, opA, d, n, m, alpha, S, ro_s, co_s, A, beta, B, ldb)   -> lsksp3
      sketch_sparse(layout, opA, opS, m, d, n, alpha, A, S, ro_s, co_s, beta, B, ldb)   -> rsksp3"""
    if isinstance(X, DenseSkOp):
        ro_s, co_s, A, beta, B, ldb = rest
        return lsksp3(layout, op1, op2, a1, a2, a3, alpha, X, ro_s, co_s, A, 0, 0, beta, B, ldb)
    S, ro_s, co_s, beta, B, ldb = rest
    return rsksp3(layout, op1, op2, a1, a2, a3, alpha, X, 0, 0, S, ro_s, co_s, beta, B, ldb)


def rsksp3_adjoint(layout, opS, m, d, n, alpha, A: CSRMatrix, S: DenseSkOp, ro_s, co_s, Bbar, ldbb, dvals):
    """Gradient of B = alpha * A * op(submat(S)) with respect to A's stored values:
    dvals[k] = alpha * sum_j Bbar[i_k, j] * op(submat(S))[c_k, j]."""
    p = _prefix(dvals)
    if not isinstance(A, CSRMatrix):
        raise TypeError(f"rsksp3_adjoint: expected a CSRMatrix, got {type(A).__name__}")
    _need_sparse_data(A, p)
    _need_dense_op(S, p)
    _need("Bbar", Bbar, _fp(p), _extent(m, d, layout, ldbb))
    _need("dvals", dvals, _fp(p), int(A.vals.numel()))
    _call(f"rbx_{_prefix(dvals)}rsksp3_adjoint", _c(layout), _c(opS), m, d, n, alpha, C.byref(A.c()),
          C.byref(S.c()), ro_s, co_s, _ptr(Bbar), ldbb, _ptr(dvals), _stream())


def sketch_general(layout, op1, op2, a1, a2, a3, alpha, *rest):
    """sketch_general overloads (skge.hh:756-1199), dispatched on which argument
    is the sketching operator, with or without (ro_s, co_s):

      sketch_general(layout, opS, opA, d, n, m, alpha, S, [ro_s, co_s,] A, lda, beta, B, ldb)
      sketch_general(layout, opA, opS, m, d, n, alpha, A, lda, S, [ro_s, co_s,] beta, B, ldb)
    """
    if isinstance(rest[0], (DenseSkOp, SparseSkOp)):
        S = rest[0]
        if len(rest) == 6:
            A, lda, beta, B, ldb = rest[1:]
            d, n, m = a1, a2, a3
            _check_full(S, op1, d, m)
            ro_s = co_s = 0
        else:
            ro_s, co_s, A, lda, beta, B, ldb = rest[1:]
        f = lskge3 if isinstance(S, DenseSkOp) else lskges
        return f(layout, op1, op2, a1, a2, a3, alpha, S, ro_s, co_s, A, lda, beta, B, ldb)
    A, lda, S = rest[0], rest[1], rest[2]
    if len(rest) == 6:
        beta, B, ldb = rest[3:]
        m, d, n = a1, a2, a3
        _check_full(S, op2, n, d)
        ro_s = co_s = 0
    else:
        ro_s, co_s, beta, B, ldb = rest[3:]
    f = rskge3 if isinstance(S, DenseSkOp) else rskges
    return f(layout, op1, op2, a1, a2, a3, alpha, A, lda, S, ro_s, co_s, beta, B, ldb)


def _check_full(S, opS, r, c):
    # full-operator overloads require exact dimensions (skge.hh:1089-1095, 1191-1197)
    if _c(opS) == b"N":
        ok = S.n_rows == r and S.n_cols == c
    else:
        ok = S.n_rows == c and S.n_cols == r
    if not ok:
        raise Error("sketch_general: operator dimensions do not match the full-operator overload")


def sketch_vector(opS, *args):
    """sketch_vector (skve.hh:140-164, 232-246).

      sketch_vector(opS, d, m, alpha, S, ro_s, co_s, x, incx, beta, y, incy)
      sketch_vector(opS, alpha, S, x, incx, beta, y, incy)

    (d, m) are the dimensions of submat(S) before op; the call reduces to
    sketch_general(RowMajor, opS, NoTrans, d', 1, m', ...) with the vector
    increments as leading dimensions."""
    if len(args) == 7:
        alpha, S, x, incx, beta, y, incy = args
        d, m, ro_s, co_s = S.n_rows, S.n_cols, 0, 0
    else:
        d, m, alpha, S, ro_s, co_s, x, incx, beta, y, incy = args
    _d, _m = (m, d) if _c(opS) == b"T" else (d, m)
    return sketch_general(Layout.RowMajor, opS, Op.NoTrans, _d, 1, _m, alpha, S, ro_s, co_s, x, incx, beta, y,
                          incy)


def sketch_symmetric(layout, *args, sym_check_tol=0.0):
    """sketch_symmetric (sksy.hh:158-177, 293-312, 406-423, 513-530).

      sketch_symmetric(layout, d, n, alpha, S, ro_s, co_s, A, lda, beta, B, ldb)   B = alpha*S*A + beta*B
      sketch_symmetric(layout, n, d, alpha, A, lda, S, ro_s, co_s, beta, B, ldb)   B = alpha*A*S + beta*B
      sketch_symmetric(layout, alpha, S, A, lda, beta, B, ldb)
      sketch_symmetric(layout, alpha, A, lda, S, beta, B, ldb)

    A (n x n, stored as a general matrix) is first checked for symmetry with
    util::require_symmetric (util.hh:127-148)."""
    if len(args) == 7:
        if isinstance(args[1], (DenseSkOp, SparseSkOp)):
            alpha, S, A, lda, beta, B, ldb = args
            d, n = S.n_rows, S.n_cols
            require_symmetric(layout, A, n, lda, sym_check_tol)
            return sketch_general(layout, Op.NoTrans, Op.NoTrans, d, n, n, alpha, S, 0, 0, A, lda, beta, B, ldb)
        alpha, A, lda, S, beta, B, ldb = args
        n, d = S.n_rows, S.n_cols
        require_symmetric(layout, A, n, lda, sym_check_tol)
        return sketch_general(layout, Op.NoTrans, Op.NoTrans, n, d, n, alpha, A, lda, S, 0, 0, beta, B, ldb)
    if isinstance(args[3], (DenseSkOp, SparseSkOp)):
        d, n, alpha, S, ro_s, co_s, A, lda, beta, B, ldb = args
        require_symmetric(layout, A, n, lda, sym_check_tol)
        return sketch_general(layout, Op.NoTrans, Op.NoTrans, d, n, n, alpha, S, ro_s, co_s, A, lda, beta, B, ldb)
    n, d, alpha, A, lda, S, ro_s, co_s, beta, B, ldb = args
    require_symmetric(layout, A, n, lda, sym_check_tol)
    return sketch_general(layout, Op.NoTrans, Op.NoTrans, n, d, n, alpha, A, lda, S, ro_s, co_s, beta, B, ldb)


def raise_if(cond: bool, msg: str):
    if cond:
        raise Error(msg)


def require_symmetric(layout, A, n, lda, tol):
    """util::require_symmetric (util.hh:127-148), evaluated on the device by
    rbx_[sd]require_symmetric; raises Error with the reference's message."""
    _need("A", A, _fp(_prefix(A)), _extent(n, n, layout, lda))
    _call(f"rbx_{_prefix(A)}require_symmetric", _c(layout), _ptr(A), n, lda, tol, _stream())


def _conv(name, src, dst):
    p = _prefix(src.vals)
    _need_sparse_data(src, p, "src")
    nnz = int(src.vals.numel())
    _need("dst.vals", dst.vals, _fp(p), nnz)
    if isinstance(dst, CSRMatrix):
        _need("dst.rowptr", dst.rowptr, _i64(), src.n_rows + 1)
        _need("dst.colidxs", dst.colidxs, _i64(), nnz)
    elif isinstance(dst, CSCMatrix):
        _need("dst.colptr", dst.colptr, _i64(), src.n_cols + 1)
        _need("dst.rowidxs", dst.rowidxs, _i64(), nnz)
    else:
        _need("dst.rows", dst.rows, _i64(), nnz)
        _need("dst.cols", dst.cols, _i64(), nnz)
    a, b = src.c(), dst.c()
    _call(f"rbx_{_prefix(src.vals)}{name}", C.byref(a), C.byref(b), _stream())


def csr_to_coo(csr: "CSRMatrix", coo: "COOMatrix") -> None:
    """csr_to_coo (sparse_data/conversions.hh:63-75) into caller-allocated COO arrays."""
    _conv("csr_to_coo", csr, coo)


def csc_to_coo(csc: "CSCMatrix", coo: "COOMatrix") -> None:
    """csc_to_coo (sparse_data/conversions.hh:49-61) into caller-allocated COO arrays."""
    _conv("csc_to_coo", csc, coo)


def coo_to_csr(coo: "COOMatrix", csr: "CSRMatrix") -> None:
    """coo_to_csr (sparse_data/conversions.hh:101-126): stable (row, column) order."""
    _conv("coo_to_csr", coo, csr)


def coo_to_csc(coo: "COOMatrix", csc: "CSCMatrix") -> None:
    """coo_to_csc (sparse_data/conversions.hh:79-99): stable (column, row) order."""
    _conv("coo_to_csc", coo, csc)


def coo_sort(coo: "COOMatrix", order: str) -> None:
    """COOMatrix::sort_arrays (coo_matrix.hh:222-226) in place: order "R" (CSR) or "C" (CSC)."""
    _need_sparse_data(coo, _prefix(coo.vals), "coo")
    a = coo.c()
    _call(f"rbx_{_prefix(coo.vals)}coo_sort", _c(order), C.byref(a), _stream())


def coo_determine_sort(coo: "COOMatrix") -> str:
    """coo_arrays_determine_sort (sparse_data/base.hh:142-167): "C", "R" or "N"."""
    _need_sparse_data(coo, _prefix(coo.vals), "coo")
    out = C.create_string_buffer(1)
    _call("rbx_coo_determine_sort", C.byref(coo.c()), out, _stream())
    return out.raw.decode()


def sample_indices_iid_uniform(n, k, samples, state: RNGState, rademachers=None) -> RNGState:
    """sample_indices_iid_uniform (util.hh:516-560): k iid indices in [0, n) into
    the int64 device tensor `samples`, and (when given) Rademacher signs into
    `rademachers`; returns the advanced state."""
    _need("samples", samples, _i64(), k)
    if rademachers is not None:
        _need("rademachers", rademachers, _fp(_prefix(rademachers)), k)
    nxt = c_rng_state()
    if rademachers is None:
        _call("rbx_sample_indices_iid_uniform", n, k, _ptr(samples), C.byref(state.c()), C.byref(nxt), _stream())
    else:
        _call(f"rbx_{_prefix(rademachers)}sample_indices_iid_uniform_signed", n, k, _ptr(samples), _ptr(rademachers),
              C.byref(state.c()), C.byref(nxt), _stream())
    return RNGState.from_c(nxt)


def sample_indices_iid(n, cdf, k, samples, state: RNGState) -> RNGState:
    """sample_indices_iid (util.hh:493-514): k iid draws from the CDF `cdf`
    (a float32 / float64 device tensor of n entries); returns the advanced state."""
    _need("cdf", cdf, _fp(_prefix(cdf)), n)
    _need("samples", samples, _i64(), k)
    nxt = c_rng_state()
    _call(f"rbx_{_prefix(cdf)}sample_indices_iid", n, _ptr(cdf), k, _ptr(samples), C.byref(state.c()),
          C.byref(nxt), _stream())
    return RNGState.from_c(nxt)


def weights_to_cdf(n, w, error_if_below=None) -> None:
    """weights_to_cdf (util.hh:459-472) in place on a float32 / float64 device
    tensor; error_if_below defaults to -sqrt(epsilon) of w's type."""
    _need("w", w, _fp(_prefix(w)), n)
    if error_if_below is None:
        import numpy as _np
        error_if_below = -float(_np.sqrt(_np.finfo(_np.float32 if _prefix(w) == "s" else _np.float64).eps))
    _call(f"rbx_{_prefix(w)}weights_to_cdf", n, _ptr(w), error_if_below, _stream())


def power_law_csr(m, n, k, exponent, state: RNGState, dtype=None):
    """Power-law sparse test data on the device (no reference counterpart; the
    reference's random_csr, sparse_data/random_matrix.hh:136-200, draws a
    uniform-density matrix from one sequential stream): k column draws from the
    Zipf weights (c+1)^(-1/(exponent-1)) by sample_indices_iid, k row draws by
    sample_indices_iid_uniform, repeated (row, column) pairs kept once, N(0,1)
    values from a DenseDist(nnz, 1).  Returns (CSRMatrix over tensors of nnz
    entries, next RNGState)."""
    import torch
    dtype = torch.float32 if dtype is None else dtype
    pre = "d" if dtype == torch.float64 else "s"
    kk = max(int(k), 1)
    vals = torch.empty(kk, dtype=dtype, device="cuda")
    colidxs = torch.empty(kk, dtype=torch.int64, device="cuda")
    rowptr = torch.empty(int(m) + 1, dtype=torch.int64, device="cuda")
    A = c_csr_matrix(m, n, 0, vals.data_ptr(), rowptr.data_ptr(), colidxs.data_ptr())
    nxt = c_rng_state()
    _call(f"rbx_{pre}power_law_csr", m, n, k, float(exponent), C.byref(state.c()), C.byref(A), C.byref(nxt),
          _stream())
    nnz = int(A.nnz)
    return CSRMatrix(m, n, vals[:nnz], rowptr, colidxs[:nnz]), RNGState.from_c(nxt)


def sketch_solve(layout, d, m, n, S: DenseSkOp, ro_s, co_s, A, lda, b, x) -> None:
    """x = argmin || submat(S) (A x - b) ||_2 on the device, submat(S) d x m at
    (ro_s, co_s), A m x n in `layout`: the sketch of [A b] as in
    examples/total-least-squares/tls_dense_skop.cc:138-171 with a QR solve of
    the sketched system (FP64 for both precisions)."""
    p = _prefix(x)
    _need_dense_op(S, p)
    _need("A", A, _fp(p), _extent(m, n, layout, lda))
    _need("b", b, _fp(p), m)
    _need("x", x, _fp(p), n)
    _call(f"rbx_{_prefix(x)}sketch_solve", _c(layout), d, m, n, C.byref(S.c()), ro_s, co_s, _ptr(A), lda, _ptr(b),
          _ptr(x), _stream())


def sketch_solve_adjoint(layout, d, m, n, S: DenseSkOp, ro_s, co_s, A, lda, b, xbar, Abar, ldab, bbar) -> None:
    """Gradient of <xbar, sketch_solve(A, b)> with respect to A (into Abar, m x n
    in `layout`) and b (into bbar); the operator is regenerated, never stored."""
    p = _prefix(Abar)
    _need_dense_op(S, p)
    _need("A", A, _fp(p), _extent(m, n, layout, lda))
    _need("b", b, _fp(p), m)
    _need("xbar", xbar, _fp(p), n)
    _need("Abar", Abar, _fp(p), _extent(m, n, layout, ldab))
    _need("bbar", bbar, _fp(p), m)
    _call(f"rbx_{_prefix(Abar)}sketch_solve_adjoint", _c(layout), d, m, n, C.byref(S.c()), ro_s, co_s, _ptr(A), lda,
          _ptr(b), _ptr(xbar), _ptr(Abar), ldab, _ptr(bbar), _stream())


def bf16_fastgen_bounds():
    """Sweep of all 2^32 words through the BF16 operator's fast generator:
    returns (measured, assumed) with measured = (max |s' - s|, max |c' - c|,
    max |r' - r| / r, words where exactly one of r, r' is zero) and assumed =
    (the sin / cos bound, the radius bound) the generator's exactness rests on."""
    meas = (C.c_double * 4)()
    assumed = (C.c_double * 2)()
    _call("rbx_bf16_fastgen_bounds", meas, assumed, _stream())
    return tuple(meas), tuple(assumed)


def release_workspace(stream=None) -> None:
    """Frees the library's scratch and rocBLAS handle kept for `stream` (a
    torch.cuda.Stream; default the current stream), after the work already
    enqueued there.  Call it before a stream the library has used is destroyed."""
    _call("rbx_release_workspace", C.c_void_p(stream.cuda_stream) if stream is not None else _stream())


def release_all_workspaces() -> None:
    """Synchronises and frees every workspace (no library work may be in flight)."""
    _call("rbx_release_all_workspaces")


def profile_enable(on: bool = True) -> None:
    """Bracket every library kernel launch with HIP events (diagnostics)."""
    _call("rbx_profile_enable", 1 if on else 0)


def profile_collect() -> dict:
    """{kernel name: (launch count, total ms)} since the last collect."""
    n_max, width = 64, 64
    names = C.create_string_buffer(n_max * width)
    counts = (C.c_int * n_max)()
    tot = (C.c_double * n_max)()
    n = lib().rbx_profile_collect(names, width, counts, tot, n_max)
    if n < 0:
        raise Error(lib().rbx_last_error_message().decode())
    raw = names.raw
    return {raw[i * width:(i + 1) * width].split(b"\0")[0].decode(): (counts[i], tot[i]) for i in range(n)}


def version() -> str:
    return lib().rbx_version().decode()
