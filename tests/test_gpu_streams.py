"""GPU: calls on different streams are independent (the reference keeps no
global state: RNGStates are never mutated, base.hh:92-95, and lskge3
allocates its temporaries per call, skge.hh:175-181).

Two different problems -- the BF16-split sketch (split planes and K-split
partials in scratch) and the SASO lskges apply (the operator's row CSR in
scratch) -- run concurrently on two streams of one device, from one host
thread and from two host threads, with sizes that change between calls so
the workspaces grow mid-sequence.  Every result must equal the serial run's
bit for bit.  A third case shares one stream between two host threads.
"""
import threading

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _bf16_problem(rb, d, m, n, key):
    A = torch.empty(m * n, dtype=torch.float32, device=DEV)
    rb.fill_dense(rb.DenseDist(m, n), A, rb.RNGState(key))
    S = rb.DenseSkOp(rb.DenseDist(d, m), rb.RNGState(key + 1))
    B = torch.empty(d * n, dtype=torch.float32, device=DEV)
    return lambda: rb.lskge3_bf16split("C", "N", "N", d, n, m, 1.0, S, 0, 0, A, m, 0.0, B, d, n_terms=3), B


def _saso_problem(rb, d, m, n, key):
    A = torch.empty(m * n, dtype=torch.float64, device=DEV)
    rb.fill_dense(rb.DenseDist(m, n, rb.ScalarDist.Uniform), A, rb.RNGState(key))
    S = rb.SparseSkOp(rb.SparseDist(d, m, 8, rb.Axis.Short), rb.RNGState(key + 1))
    rb.fill_sparse(S)
    B = torch.empty(d * n, dtype=torch.float64, device=DEV)
    return lambda: rb.lskges("C", "N", "N", d, n, m, 1.0, S, 0, 0, A, m, 0.0, B, d), B


def _problems(rb):
    # sizes alternate small / large so each stream's scratch grows mid-run
    return [_bf16_problem(rb, 256, 8192, 1024, 3), _saso_problem(rb, 500, 20000, 300, 5),
            _bf16_problem(rb, 512, 65536, 1024, 7), _saso_problem(rb, 2000, 100000, 500, 9)]


def _serial(problems):
    ref = []
    for run, B in problems:
        run()
        torch.cuda.synchronize()
        ref.append(B.clone())
        B.fill_(float("nan"))
    torch.cuda.synchronize()
    return ref


def _fresh_streams(rb, n):
    return [torch.cuda.Stream() for _ in range(n)]


def test_two_streams_one_thread(rb):
    probs = _problems(rb)
    ref = _serial(probs)
    s1, s2 = _fresh_streams(rb, 2)
    torch.cuda.synchronize()
    for rep in range(3):
        for i in (0, 2):  # interleave: bf16 on s1, saso on s2
            with torch.cuda.stream(s1):
                probs[i][0]()
            with torch.cuda.stream(s2):
                probs[i + 1][0]()
    torch.cuda.synchronize()
    for (run, B), R in zip(probs, ref):
        assert torch.equal(B, R)
    for s in (s1, s2):
        rb.release_workspace(s)
    torch.cuda.synchronize()


def test_two_streams_two_threads(rb):
    probs = _problems(rb)
    ref = _serial(probs)
    streams = _fresh_streams(rb, 2)
    torch.cuda.synchronize()
    errors = []

    def worker(k):
        try:
            with torch.cuda.stream(streams[k]):
                for rep in range(3):
                    probs[k][0]()        # small problem of this kind
                    probs[k + 2][0]()    # large one: the workspace grows
                    probs[k][0]()
            streams[k].synchronize()
        except Exception as e:  # reported after join
            errors.append(e)

    th = [threading.Thread(target=worker, args=(k,)) for k in (0, 1)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    torch.cuda.synchronize()
    assert not errors, errors
    for (run, B), R in zip(probs, ref):
        assert torch.equal(B, R)
    for s in streams:
        rb.release_workspace(s)
    torch.cuda.synchronize()


def test_one_stream_two_threads(rb):
    probs = _problems(rb)
    ref = _serial(probs)
    (s,) = _fresh_streams(rb, 1)
    torch.cuda.synchronize()
    errors = []

    def worker(k):
        try:
            with torch.cuda.stream(s):
                for rep in range(3):
                    probs[k][0]()
                    probs[k + 2][0]()
        except Exception as e:
            errors.append(e)

    th = [threading.Thread(target=worker, args=(k,)) for k in (0, 1)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    s.synchronize()
    assert not errors, errors
    for (run, B), R in zip(probs, ref):
        assert torch.equal(B, R)
    rb.release_workspace(s)
    torch.cuda.synchronize()


def test_release_then_reuse(rb):
    run, B = _bf16_problem(rb, 256, 8192, 1024, 11)
    run()
    torch.cuda.synchronize()
    R = B.clone()
    rb.release_workspace()
    B.zero_()
    run()
    torch.cuda.synchronize()
    assert torch.equal(B, R)
    rb.release_all_workspaces()
    B.zero_()
    run()
    torch.cuda.synchronize()
    assert torch.equal(B, R)
