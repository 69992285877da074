"""CPU: the C-ABI library loads, exports every symbol include/*.h declares,
and its host-only entry points (state arithmetic, descriptors, argument
validation) behave like the reference -- no device work is issued here.
"""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

from conftest import ROOT

import randblas_amd as rb


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names |= set(re.findall(r"\b(rbx_\w+)\s*\(", txt))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    L = rb.lib()
    syms = declared_symbols()
    assert len(syms) >= 30
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    # and the Python mirror binds each of them with a signature
    assert not [s for s in syms if s not in rb._SIG], [s for s in syms if s not in rb._SIG]


def test_version_string():
    assert "gfx950" in rb.version()


def test_rng_state_incr_matches_oracle(orc):
    rng = np.random.default_rng(0)
    for _ in range(50):
        c = [int(x) for x in rng.integers(0, 2**32, 4, dtype=np.uint64)]
        n = int(rng.integers(0, 2**62)) * 4 + 3
        s = rb.RNGState(key=5, counter=c).incr(n)
        assert list(s.counter) == [int(x) for x in orc.ctr_incr(c, n)]


def test_rng_state_from_key():
    s = rb.c_rng_state()
    rb._call("rbx_rng_state_from_key", (2**32) * 3 + 11, C.byref(s))
    assert list(s.key) == [11, 3] and list(s.counter) == [0, 0, 0, 0]


@pytest.mark.parametrize("shape,axis", [((400, 2000), "L"), ((2000, 400), "L"), ((7, 9), "S"), ((9, 7), "S"),
                                        ((4096, 262144), "L"), ((50000, 256), "L")])
def test_dense_skop_init_matches_oracle(orc, shape, axis):
    D = rb.DenseDist(shape[0], shape[1], rb.ScalarDist.Gaussian, rb.Axis(axis))
    S = rb.DenseSkOp(D, rb.RNGState(42, (1, 2, 3, 4)))
    assert list(S.next_state.counter) == [int(x) for x in orc.dense_next_state(shape[0], shape[1], axis, [1, 2, 3, 4])]
    assert S.layout == orc.natural_layout(shape[0], shape[1], axis)
    assert D.isometry_scale == pytest.approx(min(shape) ** -0.5 if axis == "L" else max(shape) ** -0.5)


def _lskge3(layout="C", opS="N", opA="N", d=4, n=3, m=5, S=None, ro=0, co=0, lda=5, ldb=4):
    if S is None:
        S = rb.DenseSkOp(rb.DenseDist(4, 5), rb.RNGState(1)).c()
    return rb.lib().rbx_dlskge3(layout.encode(), opS.encode(), opA.encode(), d, n, m, 1.0, C.byref(S), ro, co,
                                None, lda, 0.0, None, ldb, None)


def test_validation_errors_without_device_work():
    # each of these is a randblas_require failure in lskge3 (skge.hh:183-193)
    assert _lskge3(d=5) == rb.RBX_INVALID_ARG              # submatrix too tall
    assert "S->dist.n_rows" in rb.lib().rbx_last_error_message().decode()
    assert _lskge3(ro=1) == rb.RBX_INVALID_ARG
    assert _lskge3(lda=4) == rb.RBX_INVALID_ARG            # lda < rows(A)
    assert _lskge3(ldb=3) == rb.RBX_INVALID_ARG            # ldb < d
    assert _lskge3(layout="X") == rb.RBX_INVALID_ARG
    assert _lskge3(opA="Q") == rb.RBX_INVALID_ARG
    assert _lskge3(layout="R", lda=2) == rb.RBX_INVALID_ARG  # lda < cols(A) = 3


def test_overflow_is_reported():
    # safe_int_product (base.hh:260-274): dim_minor * P overflows int64
    D = rb.c_dense_dist(2**62, 2**62, b"G", b"S")
    out = rb.c_dense_skop()
    rc = rb.lib().rbx_dense_skop_init(C.byref(D), C.byref(rb.RNGState(0).c()), C.byref(out))
    assert rc == rb.RBX_OVERFLOW
    assert "Overflow" in rb.lib().rbx_last_error_message().decode()


def test_invalid_dist():
    D = rb.c_dense_dist(0, 5, b"G", b"L")
    out = rb.c_dense_skop()
    assert rb.lib().rbx_dense_skop_init(C.byref(D), C.byref(rb.RNGState(0).c()), C.byref(out)) == rb.RBX_INVALID_ARG
    D = rb.c_dense_dist(3, 5, b"Z", b"L")
    assert rb.lib().rbx_dense_skop_init(C.byref(D), C.byref(rb.RNGState(0).c()), C.byref(out)) == rb.RBX_INVALID_ARG
