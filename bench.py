#!/usr/bin/env python3
"""Benchmark of the fused Gaussian sketch B = S * A on one MI355X.

Workload (the large fused Gaussian sketch): A is 262144 x 1024 FP32,
column-major, A = G * diag(10^(-6 j / 1023)) with G iid N(0,1) drawn by this
library's own generator (DenseDist(262144, 1024), key 5), so its condition
number is ~1e6; S = DenseDist(4096, 262144) Gaussian, key 42, counter 0,
never materialised.  One step is one call of the BF16-split sketch
(rbx_slskge3_bf16split: A split into bf16 terms, S rounded to bf16, FP32
MFMA accumulation, FP64 combination of the K-splits), inputs resident in HBM.

value = algorithmic 2*d*m*n / time of the whole step (TFLOP/s).  Beside it:
  * fp64 / fp32: the plain FP64 fused sketch (rbx_dlskge3 on A promoted to
    FP64) and the plain FP32 one (rbx_slskge3);
  * errors of the BF16 result against FP64 products (same bf16 operator, and
    the unrounded operator) on a block of rows, and the subspace-embedding
    distortion of (1/sqrt d) S Q for A = QR;
  * roofline of the dominant kernel from HIP events around its launches;
  * cpu_baseline: the CPU oracle (OpenMP counter-addressed fill, then an
    OpenBLAS GEMM) on a row block of the same workload.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

D, M, N = 4096, 262144, 1024
KEY_S, KEY_A = 42, 5
METRIC = "Gaussian sketch SA TFLOP/s & %roofline, BF16-split vs FP64, at stated rel err"
PEAK_BF16 = 2500.0   # TFLOP/s, MI355X dense BF16 MFMA
PEAK_FP64 = 78.6     # TFLOP/s, MI355X FP64 MFMA
PEAK_FP32 = 157.3    # TFLOP/s, MI355X FP32 MFMA


PROFILE_MD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                          "bf16_skge_kernel_262144x1024_d4096.md")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def profiled_traffic(kernel: str):
    """HBM bytes per launch of `kernel` at this workload (T = 3), as measured by
    rocprofv3 FETCH_SIZE / WRITE_SIZE and recorded in the profile summary."""
    try:
        with open(PROFILE_MD) as f:
            for line in f:
                if line.startswith(f"traffic_bytes_per_launch {kernel}:"):
                    return int(line.split(":")[1])
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--terms", type=int, default=3, help="bf16 split terms of A (1-3)")
    ap.add_argument("--fp64-steps", type=int, default=2)
    ap.add_argument("--cpu-rows", type=int, default=2048, help="rows of S in the CPU baseline sample")
    ap.add_argument("--err-rows", type=int, default=64, help="rows of B checked against the FP64 oracle")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip fp64 / error / distortion legs")
    ap.add_argument("--legs", default="all",
                    help="comma list of extra legs: fp64,sparse,saso,generator,vector,solve,formats,sampling "
                         "(default all)")
    ap.add_argument("--copy-probe-gib", type=float, default=0.0,
                    help="first run one plain device copy of this many GiB (calibrates traffic counters)")
    ap.add_argument("--fp64-only", action="store_true", help="time only the plain FP64 sketch (profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: measures one GPU only; a multi-GPU run would shard the sketch by row "
              "blocks of S (one counter range per GPU) and is not part of this benchmark", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import randblas_amd as rb

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    flops = 2.0 * D * M * N

    # ---- inputs (resident in HBM before timing) ----
    log("generating A")
    from randblas_amd.datagen import conditioned_gaussian
    A = conditioned_gaussian(M, N, 1e6, KEY_A, dev)   # natural ColMajor
    S = rb.DenseSkOp(rb.DenseDist(D, M), rb.RNGState(KEY_S))
    B = torch.empty(D * N, dtype=torch.float32, device=dev)
    if args.copy_probe_gib > 0:
        n_el = int(args.copy_probe_gib * (1 << 30)) // 4
        src = torch.ones(n_el, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: {n_el * 4} bytes read, {n_el * 4} bytes written")
        del src, dst
    if args.fp64_only:
        A64 = A.double()
        B64 = torch.empty(D * N, dtype=torch.float64, device=dev)
        for _ in range(args.warmup + args.steps):
            rb.lskge3("C", "N", "N", D, N, M, 1.0, S, 0, 0, A64, M, 0.0, B64, D)
        torch.cuda.synchronize()
        log("fp64-only done")
        return

    def step():
        rb.lskge3_bf16split("C", "N", "N", D, N, M, 1.0, S, 0, 0, A, M, 0.0, B, D, n_terms=args.terms)

    log("warmup")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    rb.profile_collect()
    rb.profile_enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    rb.profile_enable(False)
    prof = rb.profile_collect()
    ms = (t1 - t0) * 1e3 / args.steps
    value = flops / (ms * 1e-3) / 1e12
    log(f"bf16-split x{args.terms}: {ms:.3f} ms/step, {value:.1f} TFLOP/s; kernels {prof}")

    # the main product may run as several launches per step (one per K range,
    # each 1/launches of the flops): achieved = flops per launch / mean launch time
    cnt, tot = prof.get("bf16_skge_kernel", (1, float("nan")))
    launches = max(cnt // max(args.steps, 1), 1)
    k_ms = tot / max(cnt, 1)
    achieved = (flops / launches) / (k_ms * 1e-3) / 1e12
    roofline = {"bound": "mfma", "kernel": "bf16_skge_kernel", "achieved": round(achieved, 2),
                "peak": PEAK_BF16, "unit": "TFLOP/s", "frac": round(achieved / PEAK_BF16, 4),
                "executed_mfma_frac": round(achieved * args.terms / PEAK_BF16, 4),
                "kernel_ms": round(k_ms, 4), "launches_per_step": launches,
                "main_ms_per_step": round(tot / max(args.steps, 1), 4),
                "traffic": profiled_traffic("bf16_skge_kernel") if args.terms == 3 else None,
                "per_kernel_ms": {k: round(v[1] / max(v[0], 1), 4) for k, v in prof.items()}}

    out = {"metric": METRIC, "value": round(value, 2), "unit": "TFLOP/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16",
           "data": "synthetic (A and S from this library's Philox generator)",
           "config": {"workload": "fused Gaussian sketch S(4096x262144, key 42) * A(262144x1024 FP32, "
                                  "key 5, column scales to cond 1e6), BF16-split",
                      "d": D, "m": M, "n": N, "split_terms": args.terms, "layout": "ColMajor"},
           "roofline": roofline}

    B_h = B.view(N, D).t().cpu().numpy().astype(np.float64)
    B64_h = B64_rows = None
    all_legs = ("fp64", "sparse", "saso", "generator", "vector", "solve", "formats", "sampling")
    legs = set(all_legs) if args.legs == "all" else set(args.legs.split(","))
    unknown = legs - set(all_legs)
    if unknown:
        ap.error(f"unknown legs: {sorted(unknown)}")
    if not args.no_extras and "fp64" in legs:
        # ---- plain FP64 fused sketch ----
        log("fp64 leg")
        A64 = A.double()
        B64 = torch.empty(D * N, dtype=torch.float64, device=dev)
        rb.lskge3("C", "N", "N", D, N, M, 1.0, S, 0, 0, A64, M, 0.0, B64, D)
        torch.cuda.synchronize()
        rb.profile_collect()
        rb.profile_enable(True)
        t0 = time.perf_counter()
        for _ in range(args.fp64_steps):
            rb.lskge3("C", "N", "N", D, N, M, 1.0, S, 0, 0, A64, M, 0.0, B64, D)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        rb.profile_enable(False)
        p64 = rb.profile_collect()
        ms64 = (t1 - t0) * 1e3 / args.fp64_steps
        tf64 = flops / (ms64 * 1e-3) / 1e12
        main64 = "f64_wide_kernel" if "f64_wide_kernel" in p64 else "skge_kernel_f64"
        c64, k64 = p64.get(main64, (1, float("nan")))
        B64_h = B64.view(N, D).t().cpu().numpy()
        B64_rows = B64_h[:args.cpu_rows].copy()
        out["fp64"] = {"ms_per_step": round(ms64, 3), "tflops": round(tf64, 3), "kernel": main64,
                       "frac_of_fp64_peak": round(flops / (k64 / c64 * 1e-3) / 1e12 / PEAK_FP64, 4),
                       "bf16_speedup": round(ms64 / ms, 2),
                       "per_kernel_ms": {k: round(v[1] / max(v[0], 1), 4) for k, v in p64.items()}}
        log(f"fp64: {ms64:.2f} ms, {tf64:.2f} TFLOP/s, bf16 speedup {ms64 / ms:.1f}x")

        del A64, B64
        torch.cuda.synchronize()

        # ---- BF16 result against the FP64 product with the same (bf16) operator, all rows ----
        out["rel_err_vs_fp64_same_operator_all_rows"] = bf16_error_all_rows(rb, torch, dev, A, B)
        log(f"bf16 vs same-operator FP64, all {D} rows: {out['rel_err_vs_fp64_same_operator_all_rows']:.3e}")
        meas, assumed = rb.bf16_fastgen_bounds()
        out["bf16_fast_generator"] = {"max_abs_sin_cos_err": max(meas[0], meas[1]), "max_rel_radius_err": meas[2],
                                      "zero_mismatches": meas[3], "assumed_sin_cos": assumed[0],
                                      "assumed_radius": assumed[1], "words": 2 ** 32}

        # ---- plain FP32 fused sketch of the same product ----
        B32 = torch.empty(D * N, dtype=torch.float32, device=dev)
        rb.lskge3("C", "N", "N", D, N, M, 1.0, S, 0, 0, A, M, 0.0, B32, D)
        torch.cuda.synchronize()
        rb.profile_collect()
        rb.profile_enable(True)
        t0 = time.perf_counter()
        for _ in range(args.fp64_steps):
            rb.lskge3("C", "N", "N", D, N, M, 1.0, S, 0, 0, A, M, 0.0, B32, D)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        rb.profile_enable(False)
        p32 = rb.profile_collect()
        ms32 = (t1 - t0) * 1e3 / args.fp64_steps
        main32 = "f32_wide_kernel" if "f32_wide_kernel" in p32 else "skge_kernel_f32"
        c32, k32 = p32.get(main32, (1, float("nan")))
        B64v = torch.empty(D * N, dtype=torch.float64, device=dev)
        rb.lskge3("C", "N", "N", D, N, M, 1.0, S, 0, 0, A.double(), M, 0.0, B64v, D)
        err32 = float(torch.linalg.norm(B32.double() - B64v) / torch.linalg.norm(B64v))
        del B64v
        out["fp32"] = {"ms_per_step": round(ms32, 3), "tflops": round(flops / (ms32 * 1e-3) / 1e12, 3),
                       "kernel": main32, "frac_of_fp32_peak": round(flops / (k32 / c32 * 1e-3) / 1e12 / PEAK_FP32, 4),
                       "bf16_speedup": round(ms32 / ms, 2), "rel_err_vs_fp64_path": err32,
                       "per_kernel_ms": {k: round(v[1] / max(v[0], 1), 4) for k, v in p32.items()}}
        log(f"fp32: {ms32:.2f} ms, {flops / (ms32 * 1e-3) / 1e12:.2f} TFLOP/s, rel err vs FP64 path {err32:.2e}, "
            f"kernels {p32}")
        del B32
        torch.cuda.synchronize()

    if not args.no_extras:
        # ---- sparse data, sparse operators, generator, consumers ----
        for key, name, fn in (("sparse", "sparse_rsksp3", sparse_leg), ("saso", "saso_lskges", saso_leg),
                              ("generator", "generator", generator_leg), ("vector", "sketch_vector", vector_leg),
                              ("solve", "sketch_and_solve", sketch_solve_leg),
                              ("formats", "sparse_formats", formats_leg), ("sampling", "sampling", sampling_leg)):
            if key in legs:
                out[name] = fn(rb, torch, dev)

    # ---- CPU: FP64 oracle rows (also the timed CPU baseline), errors, distortion ----
    # All reference linear algebra runs on the host (numpy/OpenBLAS, scipy).
    if not args.no_cpu:
        base, ref_rows, A_h = cpu_baseline(A, args.cpu_rows, np, torch)
        out["cpu_baseline"] = base
        if not args.no_extras and B64_rows is not None:
            import oracle
            rows = min(args.err_rows, args.cpu_rows)
            Sblk = ref_rows["S"][:rows]
            got = B.view(N, D).t()[:rows].cpu().numpy().astype(np.float64)
            got64 = B64_rows[:rows]
            ref_op = ref_rows["B"][:rows]
            Sb16 = oracle.bf16_round(Sblk.astype(np.float32)).astype(np.float64)
            ref_same = Sb16 @ A_h
            rel = lambda x, r: float(np.linalg.norm(x - r) / np.linalg.norm(r))
            out["rel_err_vs_fp64_same_operator"] = rel(got, ref_same)
            out["rel_err_vs_fp64_reference_operator"] = rel(got, ref_op)
            out["fp64_rel_err_vs_oracle"] = rel(got64, ref_op)
            out["error_sample"] = (f"rel_err_*: rows 0..{rows - 1} of B against the CPU oracle in FP64; "
                                   f"_all_rows: all rows against an FP64 GPU GEMM of the bf16-rounded operator")
            out["gaussian_vs_host_libm"] = libm_distance(rb, torch, dev, np, oracle)
            log(f"rel err: bf16 vs same-operator FP64 {out['rel_err_vs_fp64_same_operator']:.3e}, vs reference "
                f"operator {out['rel_err_vs_fp64_reference_operator']:.3e}; fp64 path {out['fp64_rel_err_vs_oracle']:.3e}")
            # distortion of (1/sqrt d) S Q with A = QR: the eigenvalues of the
            # pencil (B^T B, A^T A) / d are the squared singular values of S Q / sqrt(d)
            import scipy.linalg as sla
            log("distortion (host)")
            AtA = A_h.T @ A_h
            for name, Bm in (("bf16", B_h), ("fp64", B64_h)):
                ev = sla.eigh(Bm.T @ Bm / D, AtA, eigvals_only=True)
                sv = np.sqrt(np.clip(ev, 0, None))
                out[f"distortion_{name}"] = round(max(float(sv.max()) - 1.0, 1.0 - float(sv.min())), 6)
            log(f"distortion bf16 {out['distortion_bf16']}, fp64 {out['distortion_fp64']}")

    print(json.dumps(out), flush=True)


def _bf16_round(torch, x):
    """FP32 -> bfloat16 (nearest, ties to even), returned as FP32."""
    u = x.view(torch.int32).to(torch.int64) & 0xFFFFFFFF
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    u = torch.where(u >= 2 ** 31, u - 2 ** 32, u)
    return u.to(torch.int32).view(torch.float32)


def bf16_error_all_rows(rb, torch, dev, A, B, rows=256):
    """Relative Frobenius error of the BF16-split result B over all D rows
    against the FP64 product of the same operator: S generated in FP32 by
    fill_dense, rounded to bfloat16, promoted, times A in FP64 (rocBLAS
    through torch), 256 rows at a time."""
    got = B.view(N, D).t().double()
    A64 = A.view(N, M).double()
    Sf = torch.empty(rows * M, dtype=torch.float32, device=dev)
    num = den = 0.0
    for r0 in range(0, D, rows):
        rb.fill_dense_unpacked("R", rb.DenseDist(D, M), rows, M, r0, 0, Sf, rb.RNGState(KEY_S))
        ref = _bf16_round(torch, Sf).view(rows, M).double() @ A64.t()
        num += float(torch.linalg.norm(got[r0:r0 + rows] - ref)) ** 2
        den += float(torch.linalg.norm(ref)) ** 2
        del ref
    del got, A64, Sf
    torch.cuda.synchronize()
    return (num / den) ** 0.5


def libm_distance(rb, torch, dev, np, oracle, rows=64):
    """Distance in float32 ulps of the device Gaussians (rows 0..63 of the
    headline operator, 16.8M entries) from the Box-Muller the reference
    computes over this host's C library (the oracle's mathmode 1)."""
    buf = torch.empty(rows * M, dtype=torch.float32, device=dev)
    rb.fill_dense_unpacked("R", rb.DenseDist(D, M), rows, M, 0, 0, buf, rb.RNGState(KEY_S))
    got = buf.cpu().numpy()
    ref, _ = oracle.fill_dense_unpacked("R", D, M, rows, M, 0, 0, [0, 0, 0, 0], [KEY_S, 0], dtype=np.float32,
                                        mathmode=1)

    def key(x):
        i = np.asarray(x, np.float32).view(np.int32).astype(np.int64)
        return np.where(i < 0, -(i & 0x7FFFFFFF), i)
    dist = np.abs(key(got) - key(ref.reshape(-1)))
    return {"entries": int(dist.size), "max_ulp": int(dist.max()), "frac_above_2_ulp": float((dist > 2).mean()),
            "frac_bit_identical": float((dist == 0).mean())}


def sparse_leg(rb, torch, dev, reps=20):
    """rsksp3 on a 200000 x 50000 CSR matrix with power-law column degrees
    (exponent 2.1, 4e6 draws, ~20 stored entries per row, N(0,1) FP32 values,
    generated on the device from key 17) times a
    generated 50000 x 256 Gaussian operator (key 42), column-major B; and the
    gradient with respect to A's stored values (SDDMM).  Both are bound by
    memory traffic: achieved GB/s of their algorithmic bytes against HBM."""
    from randblas_amd.datagen import power_law_csr
    m, n, d = 200000, 50000, 256
    log("sparse leg: generating the CSR matrix on the device")
    A = power_law_csr(m, n, 20, 2.1, 17)
    nnz = int(A.vals.numel())
    S = rb.DenseSkOp(rb.DenseDist(n, d), rb.RNGState(42))
    B = torch.empty(m * d, dtype=torch.float32, device=dev)
    Bbar = torch.empty(m * d, dtype=torch.float32, device=dev)
    rb.fill_dense(rb.DenseDist(m, d), Bbar, rb.RNGState(3))
    dv = torch.empty(nnz, dtype=torch.float32, device=dev)

    def timed(fn):
        return _time(rb, torch, fn, reps)

    fwd_ms, fwd_k = timed(lambda: rb.rsksp3("C", "N", "N", m, d, n, 1.0, A, 0, 0, S, 0, 0, 0.0, B, m))
    adj_ms, adj_k = timed(lambda: rb.rsksp3_adjoint("C", "N", m, d, n, 1.0, A, S, 0, 0, Bbar, m, dv))
    # algorithmic bytes: stored entries (4 B value + 8 B column index), row
    # pointers, the dense operand read once, the output written once
    fwd_bytes = nnz * 12 + (m + 1) * 8 + n * d * 4 + m * d * 4
    adj_bytes = nnz * 8 + (m + 1) * 8 + n * d * 4 + m * d * 4 + nnz * 4
    res = {"workload": "CSR 200000x50000 power-law (exp 2.1, ~20/row) FP32 * Gaussian 50000x256 (key 42)",
           "nnz": nnz, "flop": 2 * nnz * d,
           "fwd_ms": round(fwd_ms, 4), "fwd_GBps": round(fwd_bytes / (fwd_ms * 1e-3) / 1e9, 1),
           "fwd_frac_hbm": round(fwd_bytes / (fwd_ms * 1e-3) / 8e12, 4), "fwd_kernels_ms": fwd_k,
           "adj_ms": round(adj_ms, 4), "adj_GBps": round(adj_bytes / (adj_ms * 1e-3) / 1e9, 1),
           "adj_frac_hbm": round(adj_bytes / (adj_ms * 1e-3) / 8e12, 4), "adj_kernels_ms": adj_k}
    log(f"sparse: rsksp3 {fwd_ms:.3f} ms, adjoint {adj_ms:.3f} ms ({nnz} stored entries)")
    # BF16-split form: bf16 operator, A's values split into three bf16 terms
    bf_ms, bf_k = timed(lambda: rb.rsksp3_bf16split("C", "N", "N", m, d, n, 1.0, A, 0, 0, S, 0, 0, 0.0, B, m))
    bf_bytes = nnz * 12 + (m + 1) * 8 + n * d * 2 + m * d * 4
    # error against the FP64 product of the same (bf16) operator: the library's
    # FP64 rsksp3 on a materialised FP64 copy of bf16(S)
    Sf = torch.empty(n * d, dtype=torch.float32, device=dev)
    rb.fill_dense(rb.DenseDist(n, d), Sf, rb.RNGState(42))
    S64 = rb.DenseSkOp(rb.DenseDist(n, d), rb.RNGState(42), buff=_bf16_round(torch, Sf).double(), layout="C")
    A64 = rb.CSRMatrix(m, n, A.vals.double(), A.rowptr, A.colidxs)
    B64 = torch.empty(m * d, dtype=torch.float64, device=dev)
    rb.rsksp3("C", "N", "N", m, d, n, 1.0, A64, 0, 0, S64, 0, 0, 0.0, B64, m)
    bf_err = float(torch.linalg.norm(B.double() - B64) / torch.linalg.norm(B64))
    rb.rsksp3("C", "N", "N", m, d, n, 1.0, A, 0, 0, S, 0, 0, 0.0, B, m)
    fp32_err = float(torch.linalg.norm(B.double() - B64) / torch.linalg.norm(B64))
    # distortion of each operator on a fixed 64-dimensional subspace of R^n (the
    # rank of A exceeds d, so A's own row space cannot be embedded): the
    # singular values of (1/sqrt d) Q^T S, Q an orthonormal n x 64 basis
    # (torch's generator, seed 0), for S in FP32 and rounded to bf16
    g = torch.Generator(device=dev).manual_seed(0)
    Q, _ = torch.linalg.qr(torch.randn(n, 64, dtype=torch.float64, device=dev, generator=g))
    dist = {}
    for name, Sm in (("fp32", Sf), ("bf16", _bf16_round(torch, Sf))):
        X = (Q.t() @ Sm.double().view(d, n).t()) / math.sqrt(d)
        sv = torch.linalg.svdvals(X.cpu())
        dist[name] = round(max(float(sv.max()) - 1.0, 1.0 - float(sv.min())), 6)
    res["bf16split"] = {"ms": round(bf_ms, 4), "GBps": round(bf_bytes / (bf_ms * 1e-3) / 1e9, 1),
                        "frac_hbm": round(bf_bytes / (bf_ms * 1e-3) / 8e12, 4), "kernels_ms": bf_k, "terms": 3,
                        "rel_err_vs_fp64_same_operator": bf_err,
                        "fp32_path_rel_err_vs_fp64_bf16_operator": fp32_err,
                        "distortion_64dim_subspace": dist}
    log(f"sparse: rsksp3 bf16-split {bf_ms:.3f} ms, rel err vs same-operator FP64 {bf_err:.2e}")
    del A, B, Bbar, dv, Sf, S64, A64, B64
    torch.cuda.synchronize()
    return res


def saso_leg(rb, torch, dev, reps=20):
    """lskges with a SASO operator: S = SparseDist(2000, 100000, vec_nnz 8,
    Short), key 7, sampled once (fill_sparse, which also prepares the
    operator's row-sorted copy); A 100000 x 500 column-major
    Uniform(-1, 1) (key 3); B = S * A, 2000 x 500, in FP64 and FP32.  Bound by
    HBM: algorithmic bytes = A read once + B written once (+ the operator's
    800000 stored entries, 24 B each, read once)."""
    d, m, n = 2000, 100000, 500
    res = {"workload": "SASO 2000x100000 vec_nnz 8 (key 7) * A 100000x500 ColMajor Uniform(-1,1) (key 3)"}
    outs = {}
    for name, dt in (("fp64", torch.float64), ("fp32", torch.float32)):
        S = rb.SparseSkOp(rb.SparseDist(d, m, 8, rb.Axis.Short), rb.RNGState(7), dtype=dt)
        A = torch.empty(m * n, dtype=dt, device=dev)
        rb.fill_dense(rb.DenseDist(m, n, rb.ScalarDist.Uniform), A, rb.RNGState(3))
        B = torch.empty(d * n, dtype=dt, device=dev)
        rb.fill_sparse(S)
        call = lambda: rb.lskges("C", "N", "N", d, n, m, 1.0, S, 0, 0, A, m, 0.0, B, d)
        ms, prof = _time(rb, torch, call, reps)
        esz = A.element_size()
        nbytes = m * n * esz + d * n * esz + 8 * m * 24
        # the same calls without the prepared copy (each call sorts the
        # operator's entries by row, as the reference's COO path does), and the
        # one-off cost of preparing it
        S.release()
        ms_sort, _ = _time(rb, torch, call, reps)
        t0 = time.perf_counter()
        S.prepare()
        torch.cuda.synchronize()
        ms_prep = (time.perf_counter() - t0) * 1e3
        res[name] = {"ms": round(ms, 4), "GBps": round(nbytes / (ms * 1e-3) / 1e9, 1),
                     "frac_hbm": round(nbytes / (ms * 1e-3) / 8e12, 4),
                     "ms_sorting_per_call": round(ms_sort, 4), "prepare_ms": round(ms_prep, 4),
                     "kernels_ms": prof}
        log(f"saso lskges {name}: {ms:.3f} ms (prepared operator), {ms_sort:.3f} ms sorting per call, "
            f"prepare {ms_prep:.3f} ms")
        outs[name] = B.double()
        del A, B
        torch.cuda.synchronize()
    # the FP32 product against the FP64 one (same operator, same A up to its
    # FP32 rounding); each is bit-identical to the reference's in its own
    # precision (tests/test_gpu_sparse.py)
    res["fp32_rel_err_vs_fp64"] = float(torch.linalg.norm(outs["fp32"] - outs["fp64"]) /
                                        torch.linalg.norm(outs["fp64"]))
    del outs
    return res


def _time(rb, torch, fn, reps, warm=3, prof_reps=3):
    """Median over `reps` calls of the GPU time between HIP events recorded
    around each call (after `warm` untimed calls), then the per-kernel
    breakdown from a separate pass of `prof_reps` calls with the library's
    kernel events on (so they add nothing to the timed calls)."""
    for _ in range(warm):
        fn()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    for a, b in evs:
        a.record()
        fn()
        b.record()
    torch.cuda.synchronize()
    times = sorted(a.elapsed_time(b) for a, b in evs)
    ms = times[len(times) // 2]
    rb.profile_collect()
    rb.profile_enable(True)
    for _ in range(prof_reps):
        fn()
    torch.cuda.synchronize()
    rb.profile_enable(False)
    prof = rb.profile_collect()
    return ms, {k: round(v[1] / max(v[0], 1), 4) for k, v in prof.items()}


def generator_leg(rb, torch, dev, reps=20):
    """The generator alone: fill_dense of the headline operator, DenseDist(4096,
    262144) Gaussian, key 42, into FP32 and FP64 buffers (4 / 8 GiB).  Bound by
    the generator's VALU work, priced against the HBM write rate."""
    res = {"workload": "fill_dense DenseDist(4096, 262144) Gaussian, key 42"}
    n_el = D * M
    for name, dt in (("fp32", torch.float32), ("fp64", torch.float64)):
        buf = torch.empty(n_el, dtype=dt, device=dev)
        ms, k = _time(rb, torch, lambda: rb.fill_dense(rb.DenseDist(D, M), buf, rb.RNGState(KEY_S)), reps)
        nbytes = n_el * buf.element_size()
        res[name] = {"ms": round(ms, 4), "gaussians_per_s": round(n_el / (ms * 1e-3), 1),
                     "GBps": round(nbytes / (ms * 1e-3) / 1e9, 1), "frac_hbm": round(nbytes / (ms * 1e-3) / 8e12, 4),
                     "kernels_ms": k}
        log(f"generator {name}: {ms:.3f} ms")
        del buf
        torch.cuda.synchronize()
    return res


def vector_leg(rb, torch, dev, reps=20):
    """One-column sketches: sketch_vector with the headline operator,
    y = S x for S = DenseDist(4096, 262144) Gaussian (key 42) and x of 262144
    N(0,1) entries, FP32 and FP64.  Bound by generating the 1.07e9 operator
    entries (each used once), reported as operator entries per second."""
    res = {"workload": "sketch_vector S DenseDist(4096, 262144) Gaussian (key 42) * x (262144)"}
    S = rb.DenseSkOp(rb.DenseDist(D, M), rb.RNGState(KEY_S))
    for name, dt in (("fp32", torch.float32), ("fp64", torch.float64)):
        x = torch.empty(M, dtype=dt, device=dev)
        rb.fill_dense(rb.DenseDist(M, 1), x, rb.RNGState(9))
        y = torch.empty(D, dtype=dt, device=dev)
        ms, k = _time(rb, torch, lambda: rb.sketch_vector("N", 1.0, S, x, 1, 0.0, y, 1), reps)
        res[name] = {"ms": round(ms, 4), "entries_per_s": round(D * M / (ms * 1e-3), 1), "kernels_ms": k}
        log(f"vector {name}: {ms:.3f} ms")
    return res


def sketch_solve_leg(rb, torch, dev, reps=20):
    """The sketch-and-solve product and its adjoint: S = DenseDist(1024, 65536)
    (key 42) times [A b], A 65536 x 256 N(0,1) (key 11), FP32 and FP64; the
    adjoint S^T G for G 1024 x 257 regenerates the operator.  2*1024*65536*257
    flop each, against the FP32 / FP64 MFMA peaks (157.3 / 78.6 TFLOP/s).
    Then the whole device solve x = argmin ||S(Ax - b)|| (the two sketches,
    FP64 QR of the 1024 x 256 system, triangular solve) and its adjoint."""
    d, m, n = 1024, 65536, 257
    S = rb.DenseSkOp(rb.DenseDist(d, m), rb.RNGState(42))
    fl = 2.0 * d * m * n
    res = {"workload": "S 1024x65536 (key 42) * [A b] 65536x257 and S^T * G, G 1024x257", "flop": fl}
    for name, dt, peak in (("fp32", torch.float32, 157.3), ("fp64", torch.float64, PEAK_FP64)):
        Ab = torch.empty(m * n, dtype=dt, device=dev)
        rb.fill_dense(rb.DenseDist(m, n), Ab, rb.RNGState(11))
        SA = torch.empty(d * n, dtype=dt, device=dev)
        G = torch.empty(d * n, dtype=dt, device=dev)
        rb.fill_dense(rb.DenseDist(d, n), G, rb.RNGState(13))
        Abar = torch.empty(m * n, dtype=dt, device=dev)
        fwd_ms, fk = _time(rb, torch, lambda: rb.lskge3("C", "N", "N", d, n, m, 1.0, S, 0, 0, Ab, m, 0.0, SA, d), reps)
        adj_ms, ak = _time(rb, torch, lambda: rb.lskge3_adjoint("C", "N", d, n, m, 1.0, S, 0, 0, G, d, 0.0, Abar, m),
                           reps)
        nA = n - 1
        A, b = Ab[:m * nA], Ab[m * nA:]
        x = torch.empty(nA, dtype=dt, device=dev)
        xbar = torch.ones(nA, dtype=dt, device=dev)
        bbar = torch.empty(m, dtype=dt, device=dev)
        solve_ms, sk = _time(rb, torch, lambda: rb.sketch_solve("C", d, m, nA, S, 0, 0, A, m, b, x), reps)
        sadj_ms, sak = _time(rb, torch, lambda: rb.sketch_solve_adjoint("C", d, m, nA, S, 0, 0, A, m, b, xbar,
                                                                        Abar, m, bbar), reps)
        res[name] = {"fwd_ms": round(fwd_ms, 4), "fwd_tflops": round(fl / (fwd_ms * 1e-3) / 1e12, 2),
                     "fwd_frac_peak": round(fl / (fwd_ms * 1e-3) / 1e12 / peak, 4),
                     "adj_ms": round(adj_ms, 4), "adj_tflops": round(fl / (adj_ms * 1e-3) / 1e12, 2),
                     "adj_frac_peak": round(fl / (adj_ms * 1e-3) / 1e12 / peak, 4),
                     "fwd_kernels_ms": fk, "adj_kernels_ms": ak,
                     "solve_ms": round(solve_ms, 4), "solve_adjoint_ms": round(sadj_ms, 4),
                     "solve_kernels_ms": sk, "solve_adjoint_kernels_ms": sak}
        log(f"sketch-and-solve {name}: fwd {fwd_ms:.3f} ms, adjoint {adj_ms:.3f} ms, "
            f"solve {solve_ms:.3f} ms, solve adjoint {sadj_ms:.3f} ms")
        del Ab, SA, G, Abar, x, xbar, bbar
        torch.cuda.synchronize()
    return res


def formats_leg(rb, torch, dev, reps=20):
    """The sparse-data path around the sketches, on the 200000 x 50000
    power-law matrix of the sparse leg (FP32 values, int64 indices):
    conversions CSR -> COO, COO -> CSC and COO -> CSR (stable sorts on the
    device), and the left sketch lsksp3 of that tall sparse matrix,
    B = S * A with S = DenseDist(256, 200000) Gaussian (key 42), B 256 x 50000
    column-major, given as CSR, CSC and COO.  All bound by memory traffic:
    achieved GB/s of the algorithmic bytes (index and value arrays read and
    written once; lsksp3: A's arrays, the generated operator's rows it
    touches, B written once) against HBM."""
    from randblas_amd.datagen import power_law_csr
    m, n, d = 200000, 50000, 256
    csr = power_law_csr(m, n, 20, 2.1, 17)
    nnz = int(csr.vals.numel())
    i64 = dict(dtype=torch.int64, device=dev)
    coo = rb.COOMatrix(m, n, torch.empty(nnz, dtype=torch.float32, device=dev), torch.empty(nnz, **i64),
                       torch.empty(nnz, **i64))
    csc = rb.CSCMatrix(m, n, torch.empty(nnz, dtype=torch.float32, device=dev), torch.empty(nnz, **i64),
                       torch.empty(n + 1, **i64))
    csr2 = rb.CSRMatrix(m, n, torch.empty(nnz, dtype=torch.float32, device=dev), torch.empty(m + 1, **i64),
                        torch.empty(nnz, **i64))
    res = {"workload": "CSR 200000x50000 power-law (exp 2.1, ~20/row) FP32, int64 indices", "nnz": nnz}

    def rate(name, fn, nbytes):
        ms, k = _time(rb, torch, fn, reps)
        res[name] = {"ms": round(ms, 4), "GBps": round(nbytes / (ms * 1e-3) / 1e9, 1),
                     "frac_hbm": round(nbytes / (ms * 1e-3) / 8e12, 4), "kernels_ms": k}
        log(f"formats: {name} {ms:.3f} ms")

    entry = 8 + 8 + 4  # row, column, value of a COO entry
    rate("csr_to_coo", lambda: rb.csr_to_coo(csr, coo), (m + 1) * 8 + nnz * 12 + nnz * entry)
    rate("coo_to_csc", lambda: rb.coo_to_csc(coo, csc), nnz * entry + nnz * 12 + (n + 1) * 8)
    rate("coo_to_csr", lambda: rb.coo_to_csr(coo, csr2), nnz * entry + nnz * 12 + (m + 1) * 8)
    S = rb.DenseSkOp(rb.DenseDist(d, m), rb.RNGState(42))
    B = torch.empty(d * n, dtype=torch.float32, device=dev)
    # the operator is generated, not read: its bytes are not counted
    sk_bytes = {"CSR": (m + 1) * 8 + nnz * 12, "CSC": (n + 1) * 8 + nnz * 12, "COO": nnz * entry}
    for name, A in (("CSR", csr), ("CSC", csc), ("COO", coo)):
        rate(f"lsksp3_{name}", lambda A=A: rb.lsksp3("C", "N", "N", d, n, m, 1.0, S, 0, 0, A, 0, 0, 0.0, B, d),
             sk_bytes[name] + d * n * 4)
    res["lsksp3_workload"] = "S DenseDist(256, 200000) Gaussian (key 42) * A -> B 256x50000 ColMajor FP32"
    del csr, coo, csc, csr2, B
    torch.cuda.synchronize()
    return res


def sampling_leg(rb, torch, dev, reps=20):
    """Index sampling: 2^27 iid uniform indices in [0, 10^9) (int64 out), the
    same with FP32 Rademacher signs, and 2^26 draws from a CDF over 10^6
    indices (weights_to_cdf of power-law weights first, one sequential pass,
    timed on its own).  Bound by the HBM writes of the samples (8 B per index,
    +4 B per sign)."""
    k, kc, nc = 1 << 27, 1 << 26, 1000000
    res = {"workload": "2^27 uniform indices in [0, 1e9); 2^26 draws from a 1e6-entry CDF"}
    s = torch.empty(k, dtype=torch.int64, device=dev)
    r = torch.empty(k, dtype=torch.float32, device=dev)
    st = rb.RNGState(9)

    def rate(name, fn, nbytes):
        ms, kk = _time(rb, torch, fn, reps)
        res[name] = {"ms": round(ms, 4), "GBps": round(nbytes / (ms * 1e-3) / 1e9, 1),
                     "frac_hbm": round(nbytes / (ms * 1e-3) / 8e12, 4), "kernels_ms": kk}
        log(f"sampling: {name} {ms:.3f} ms")

    rate("iid_uniform", lambda: rb.sample_indices_iid_uniform(10 ** 9, k, s, st), k * 8)
    rate("iid_uniform_signed", lambda: rb.sample_indices_iid_uniform(10 ** 9, k, s, st, rademachers=r), k * 12)
    w = 1.0 / torch.arange(1, nc + 1, dtype=torch.float64, device=dev) ** 1.1
    cdf = w.clone()
    rb.weights_to_cdf(nc, cdf)
    ms_cdf, _ = _time(rb, torch, lambda: (cdf.copy_(w), rb.weights_to_cdf(nc, cdf)), 2)
    res["weights_to_cdf_ms"] = round(ms_cdf, 4)
    rate("iid_from_cdf", lambda: rb.sample_indices_iid(nc, cdf, kc, s[:kc], st), kc * 8)
    del s, r, w, cdf
    torch.cuda.synchronize()
    return res


def cpu_baseline(A, rows, np, torch):
    """The CPU oracle on a row block of the same workload: counter-addressed
    OpenMP fill of `rows` rows of S (dense_skops.hh:96-170 restated), then the
    FP64 GEMM with OpenBLAS through numpy (blas::gemm at skge.hh:200).
    Returns (baseline record, {"S", "B"} of that block, A on the host)."""
    import oracle
    log(f"cpu baseline: {rows} rows of S")
    cores = os.cpu_count() or 1
    A_h = A.view(N, M).cpu().numpy().astype(np.float64).T  # M x N
    oracle.lib()
    t0 = time.perf_counter()
    Sblk, _ = oracle.fill_dense_unpacked("R", D, M, rows, M, 0, 0, [0, 0, 0, 0], [KEY_S, 0],
                                         dtype=np.float64, nthreads=cores)
    t1 = time.perf_counter()
    Bblk = Sblk @ A_h
    t2 = time.perf_counter()
    fl = 2.0 * rows * M * N
    rec = {"value": round(fl / (t2 - t0) / 1e12, 5), "unit": "TFLOP/s", "cores": cores, "kind": "port",
           "sample": f"rows 0..{rows - 1} of S (of 4096) times the full A: fill {t1 - t0:.2f} s, "
                     f"FP64 GEMM (OpenBLAS) {t2 - t1:.2f} s",
           "fill_s": round(t1 - t0, 3), "gemm_s": round(t2 - t1, 3)}
    log(f"cpu baseline {rec['value']} TFLOP/s ({rec['sample']})")
    return rec, {"S": Sblk, "B": Bblk}, A_h


if __name__ == "__main__":
    main()
