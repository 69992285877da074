"""CPU: pin the oracle's Philox4x32-10, counter arithmetic and uniform maps.

Pins: Random123 KAT rows (r123_kat_vectors.txt:19-21 via tests/golden), the
counter-carry assertions of test_r123.cc:706-797, the key constructor of
test_r123.cc:675-703, and closed-form values of uneg11/u01.
"""
import json
import os

import numpy as np

from conftest import ROOT

KAT = json.load(open(os.path.join(ROOT, "tests", "golden", "philox4x32_10_kat.json")))


def test_philox_kat(orc):
    assert len(KAT["rows"]) == 3
    for row in KAT["rows"]:
        out = orc.philox(row["counter"], row["key"])
        assert [int(x) for x in out] == row["expected"]


def test_philox_stream_matches_single_calls_across_carry(orc):
    # counters straddling the 2^32 and 2^64 word boundaries
    for start in ([0xFFFFFFF0, 0, 0, 0], [0xFFFFFFFE, 0xFFFFFFFF, 7, 0], [5, 6, 7, 8]):
        words = orc.philox_stream(start, [42, 0], 40)
        for i in (0, 1, 15, 16, 17, 39):
            c = orc.ctr_incr(start, i)
            assert np.array_equal(words[i], orc.philox(c, [42, 0]))


def test_counter_incr_semantics(orc):
    # restates test_r123.cc:706-797
    i32max = 2**32 - 1
    c = orc.ctr_incr([0, 0, 0, 0], i32max)
    assert list(c) == [i32max, 0, 0, 0]
    c = orc.ctr_incr(c, 1)
    assert list(c) == [0, 1, 0, 0]
    c = orc.ctr_incr(c, 3)
    assert list(c) == [3, 1, 0, 0]
    assert list(orc.ctr_incr([0, 0, 0, 0], 2**32 - 1)) == [i32max, 0, 0, 0]
    assert list(orc.ctr_incr([0, 0, 0, 0], 2**32)) == [0, 1, 0, 0]
    c = orc.ctr_incr(orc.ctr_incr([0, 0, 0, 0], 2**63), 2**63 - 2**32)
    assert list(c) == [0, i32max, 0, 0]
    assert list(orc.ctr_incr(c, 2**32)) == [0, 0, 1, 0]
    assert list(orc.ctr_incr([i32max, i32max, i32max, 0], 1)) == [0, 0, 0, 1]


def test_counter_incr_is_128bit_addition(orc):
    rng = np.random.default_rng(3)
    for _ in range(200):
        c = rng.integers(0, 2**32, 4, dtype=np.uint64).astype(np.uint32)
        n = int(rng.integers(0, 2**63)) * 2 + int(rng.integers(0, 2))
        val = sum(int(w) << (32 * i) for i, w in enumerate(c))
        exp = (val + n) % 2**128
        got = orc.ctr_incr(c, n)
        assert sum(int(w) << (32 * i) for i, w in enumerate(got)) == exp


def test_key_from_uint64(orc):
    # RNGState(uint64 k): key = {k mod 2^32, k >> 32} (base.hh:119, test_r123.cc:675-703)
    k = orc._key(42)
    assert list(k) == [42, 0]
    k = orc._key(2**40 + 5)
    assert list(k) == [5, 2**8]


def test_uniform_maps_closed_form(orc):
    L = orc.lib()
    # uneg11(w) = (float)(int32)w * 2^-31 + 2^-32, u01(w) = (float)w * 2^-32 + 2^-33
    assert L.orc_uneg11(0) == np.float32(2.0**-32)
    assert L.orc_uneg11(0x7FFFFFFF) == np.float32(1.0)          # (float)(2^31-1) rounds to 2^31
    assert L.orc_uneg11(0x80000000) == np.float32(-1.0) + np.float32(2.0**-32)
    assert L.orc_uneg11(0xFFFFFFFF) == np.float32(-(2.0**-31) + 2.0**-32)
    assert L.orc_u01(0) == np.float32(2.0**-33)
    assert L.orc_u01(0xFFFFFFFF) == np.float32(1.0)
    rng = np.random.default_rng(5)
    for w in rng.integers(0, 2**32, 2000, dtype=np.uint64):
        w = int(w)
        s = w - 2**32 if w >= 2**31 else w
        exp_neg = np.float32(np.float32(s) * np.float32(2.0**-31) + np.float32(2.0**-32))
        exp_u = np.float32(np.float32(w) * np.float32(2.0**-32) + np.float32(2.0**-33))
        assert L.orc_uneg11(w) == exp_neg
        assert L.orc_u01(w) == exp_u
