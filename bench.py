"""Headline benchmark: Ozaki DGEMM effective TFLOPS (2mnk/t) on one MI355X.

Workload (one "step"): one Ozaki-scheme DGEMM C = A*B, m = n = k = 8192, binary64 inputs from the
reference harness's phi-mode generator (seeds 1 and 2), split degree auto (nSplitMax = 0),
NearSum (sumMode = 1), alpha = 1, beta = 0, computed with INT8 digit slices on INT32-accumulating
MFMA.  The same product in the plain FP64-slice form (FP64 MFMA) is timed beside it and the two
outputs are compared bit for bit.  Inputs are resident in HBM before the timed region.

Prints one JSON line (see the field list at the bottom).  Extra fields: the FP64-slice time and
the speedup, the slice counts, the phi = 4 variant, the dominant kernel's roofline (MFMA), and the
CPU baseline: the oracle restatement of the reference CPU OzBLAS timed on this host's cores on a
bounded sample.  Under "results" two more workloads are reported beside the headline:
  cg_laplace100: reproducible CG (Ozaki DOT/NRM2/CSRMV, NearSum) on the 3-D 7-point Laplacian of
    a 100^3 grid (n = 10^6), 500 iterations, residual history checked bit for bit against the
    committed golden history;
  batched256: 256 uneven Ozaki DGEMMs (m, n, k uniform in [64, 1024] from srand48(4), phi = 2,
    seeds 1000 + 2p / 1001 + 2p) in the grouped INT8 launch replayed from its hipGraph, against
    the same problems one ozblasDgemm call at a time.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

INT8_PEAK_TOPS = 5000.0   # MI355X dense INT8 MFMA, TOPS
HALF_PEAK_TFLOPS = 2500.0  # MI355X dense FP16 / BF16 MFMA, TFLOPS
FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 MFMA


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--phi", type=float, default=0.5)
    ap.add_argument("--phi2", type=float, default=4.0, help="second phi value (0 = skip)")
    ap.add_argument("--form", default="int8", choices=["int8", "fp16", "bf16"])
    ap.add_argument("--scheme", default="auto", choices=["auto", "residues", "digits"],
                    help="slices of the fast form: residues modulo small moduli or radix-256 digits")
    ap.add_argument("--all-forms", type=int, default=1,
                    help="also time INT8/FP16/BF16 x residues/digits at the first phi (0 = skip)")
    ap.add_argument("--form-steps", type=int, default=2)
    ap.add_argument("--fp64-steps", type=int, default=2)
    ap.add_argument("--cpu-n", type=int, default=4096, help="CPU baseline sample size (0 = skip)")
    ap.add_argument("--cpu-reps", type=int, default=2, help="CPU baseline runs (the minimum is reported)")
    ap.add_argument("--cg-iters", type=int, default=500, help="CG iterations on 100^3 (0 = skip)")
    ap.add_argument("--batch", type=int, default=256, help="batched problems (0 = skip)")
    ap.add_argument("--level12", type=int, default=1, help="DGEMV / DDOT / DCSRMV rates (0 = skip)")
    ap.add_argument("--work-gb", type=float, default=64.0)
    ap.add_argument("--copy-gib", type=float, default=0.0,
                    help="only run a plain device copy of this many GiB (HBM counter calibration)")
    return ap.parse_args()


TRAFFIC_PROFILES = [os.path.join(ROOT, "profiles", f) for f in
                    ("k_digit_gemm_i8_residues_traffic.md", "k_digit_gemm_i8_traffic.md")]


def measured_traffic(n, phi, form, scheme):
    """HBM bytes per launch of the dominant digit-kernel launch, as the committed PMC summary states
    them (FETCH_SIZE + WRITE_SIZE of that launch, each scaled by the same counter's reading for a
    plain copy of known size).  Counters cannot be read inside this timed run, so the figure comes
    from that summary; None when it is absent or was measured on another workload."""
    for path in TRAFFIC_PROFILES:
        try:
            text = open(path).read()
        except OSError:
            continue
        for m in re.finditer(r"^traffic: n=(\d+) phi=([0-9.]+) form=(\w+) scheme=(\w+) bytes=(\d+)$", text, re.M):
            if int(m.group(1)) == n and float(m.group(2)) == phi and m.group(3) == form and m.group(4) == scheme:
                return int(m.group(5))
    return None


def copy_calibration(gib, torch):
    """One elementwise pass b = -a over `gib` GiB of binary64: reads and writes exactly gib GiB."""
    n = int(gib * (1 << 30)) // 8
    a = torch.empty(n, dtype=torch.float64, device="cuda:0")
    b = torch.empty_like(a)
    torch.neg(a, out=b)
    torch.cuda.synchronize()
    print(json.dumps({"copy_read_bytes": 8 * n, "copy_write_bytes": 8 * n}), flush=True)


def timed_gemm(H, n, dA, dB, dC, steps, warmup, torch):
    for _ in range(warmup):
        H.dgemm("N", "N", n, n, n, 1.0, dA, n, dB, n, 0.0, dC, n)
    torch.cuda.synchronize()
    tk = []
    t0 = time.perf_counter()
    for _ in range(steps):
        H.dgemm("N", "N", n, n, n, 1.0, dA, n, dB, n, 0.0, dC, n)
        tk.append(H.t_comp)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    return (t1 - t0) / steps, sum(tk) / len(tk), dict(nSplitA=H.nSplitA, nSplitB=H.nSplitB,
                                                      nSplitC=H.nSplitC, t_split=H.t_SplitA + H.t_SplitB,
                                                      t_sum=H.t_sum)


def cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(n_cpu, phi, reps=2):
    """Oracle restatement of the reference CPU OzBLAS (plain FP64 slices, NearSum) on this host's
    cores, its slice products through OpenBLAS's cblas_dgemm as in the reference
    (src/cpu/ozblas_blas_wrapper.cpp:158-168); the split and the sums use OpenMP over columns /
    elements as the reference does.  The minimum over `reps` runs."""
    import numpy as np
    from oracle import oracle as O
    blas = O.use_openblas(True)
    A = O.init_mat(n_cpu, n_cpu, n_cpu, phi, 2, 1)
    B = O.init_mat(n_cpu, n_cpu, n_cpu, phi, 2, 2)
    C = np.zeros(n_cpu * n_cpu)
    h = O.Handle(1 << 35, sumMode=1, reproMode=1)
    ts = []
    for _ in range(max(1, reps)):
        t0 = time.perf_counter()
        h.gemm("N", "N", n_cpu, n_cpu, n_cpu, 1.0, A, n_cpu, B, n_cpu, 0.0, C, n_cpu)
        ts.append(time.perf_counter() - t0)
    O.use_openblas(False)
    t = min(ts)
    cores = len(os.sched_getaffinity(0))
    return {"value": round(2.0 * n_cpu ** 3 / t / 1e12, 5), "unit": "TFLOPS (2mnk/t)", "cores": cores,
            "kind": "port",
            "sample": f"plain-form Ozaki DGEMM m=n=k={n_cpu}, phi={phi}, NearSum, degree auto "
                      f"({h.nSplitA:.0f}x{h.nSplitB:.0f} FP64 slices, {h.nSplitC:.0f} products): the oracle "
                      f"C restatement, slice products through "
                      f"{'OpenBLAS cblas_dgemm (' + os.path.basename(blas) + ')' if blas else 'its own loop'}, "
                      f"OpenMP on {cores} threads of {cpu_model()}; min of {len(ts)} runs: {t:.2f} s"}


def drand48_stream(seed):
    """POSIX drand48 after srand48(seed): X <- (0x5DEECE66D X + 0xB) mod 2^48, value X / 2^48."""
    x = ((seed & 0xFFFFFFFF) << 16) | 0x330E
    while True:
        x = (0x5DEECE66D * x + 0xB) % (1 << 48)
        yield x / float(1 << 48)


def laplace3d(g, shift):
    """3-D 7-point Laplacian on a g^3 grid (lexicographic, i fastest), diagonal 6 + shift, CSR
    with int32 indices and sorted columns."""
    import numpy as np
    n = g ** 3
    idx = np.arange(n, dtype=np.int64)
    i, j, k = idx % g, (idx // g) % g, idx // (g * g)
    cols, vals = [], []
    for (di, dj, dk, v) in ((0, 0, -1, -1.0), (0, -1, 0, -1.0), (-1, 0, 0, -1.0), (0, 0, 0, 6.0 + shift),
                            (1, 0, 0, -1.0), (0, 1, 0, -1.0), (0, 0, 1, -1.0)):
        ok = (i + di >= 0) & (i + di < g) & (j + dj >= 0) & (j + dj < g) & (k + dk >= 0) & (k + dk < g)
        cols.append(np.where(ok, idx + di + dj * g + dk * g * g, -1))
        vals.append(np.full(n, v))
    C = np.stack(cols, 1)
    V = np.stack(vals, 1)
    keep = C >= 0
    rowptr = np.zeros(n + 1, dtype=np.int32)
    rowptr[1:] = np.cumsum(keep.sum(1))
    return rowptr, C[keep].astype(np.int32), V[keep]


def bench_cg(iters, torch, oz):
    """Config: reproducible CG on the 100^3 Laplacian (n = 10^6), NearSum, `iters` iterations."""
    import hashlib
    import numpy as np
    g = 100
    rowptr, colind, val = laplace3d(g, 1e-3)
    n, nnz = g ** 3, len(val)
    r = drand48_stream(3)
    b = np.array([2.0 * next(r) - 1.0 for _ in range(n)])
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")
    dval, dcol, drow, db = d(val), d(colind), d(rowptr), d(b)
    H = oz.Handle(4 << 30, sumMode=1)
    dx = torch.zeros(n, dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hist = H.dcg(n, nnz, dval, dcol, drow, db, dx, iters, 0.0)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    out = {"iterations": iters, "seconds": round(t, 3), "ms_per_iteration": round(t / iters * 1e3, 3),
           "final_residual": float(hist[-1]), "nSplitA_": int(H.nSplitA_), "splitShift": int(H.splitShift),
           # reference accounting of one iteration (testing/testing_solver.cpp:146-149)
           "ref_bytes_per_iteration": 12 * nnz + 140 * n,
           "ref_GBps": round((12 * nnz + 140 * n) / (t / iters) / 1e9, 1)}
    fix = os.path.join(ROOT, "tests", "golden", "cg_laplace100.json")
    if os.path.exists(fix) and iters == 500:
        case = json.load(open(fix))["cases"]["1"]
        out["history_bitwise_equal_to_cpu"] = [float(v).hex() for v in hist] == \
            [float.fromhex(v).hex() for v in case["history_hex"]]
        out["x_bitwise_equal_to_cpu"] = hashlib.sha256(
            dx.cpu().numpy().tobytes()).hexdigest() == case["x_sha256"]
        out["cpu_oracle_seconds_500_iterations"] = case["cpu_seconds"]
    H.close()
    return out


def bench_level12(torch, oz):
    """Level 1/2 routines at the sizes of the CG and GEMM configs, NearSum, with the reference
    harness's byte counts (testing/testing_level1b.cpp:2, testing_level2.cpp:2,
    testing_level2s.cpp:2): DGEMV (2m + n + mn) 8, DDOT 2n 8, DCSRMV (nnz + 2m + n) 8 + (nnz + m + 1) 4."""
    import numpy as np
    out = {}

    def timeit(fn, reps):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps

    m = n = 8192
    A = torch.from_numpy(oz.gen_phi(m, n, m, 0.5, 1)).to("cuda:0")
    x = torch.from_numpy(oz.gen_phi(n, 1, n, 0.5, 2)).to("cuda:0")
    y = torch.zeros(m, dtype=torch.float64, device="cuda:0")
    H = oz.Handle(8 << 30, sumMode=1)
    t = timeit(lambda: H.dgemv("N", m, n, 1.0, A, m, x, 0.0, y), 5)
    out["dgemv_8192"] = {"ms": round(t * 1e3, 3), "slices": [float(H.nSplitA), float(H.nSplitB)],
                         "ref_GBps": round((2.0 * m + n + m * n) * 8 / t / 1e9, 1)}
    H.close()
    del A
    nv = 10 ** 6
    xv = torch.from_numpy(oz.gen_phi(nv, 1, nv, 2.0, 21)).to("cuda:0")
    yv = torch.from_numpy(oz.gen_phi(nv, 1, nv, 2.0, 22)).to("cuda:0")
    for mode in (0, 2):
        H = oz.Handle(1 << 30, sumMode=1, splitEpsMode=mode)
        t = timeit(lambda: H.ddot(nv, xv, yv), 20)
        out[f"ddot_1e6{'_dot2' if mode == 2 else ''}"] = {
            "ms": round(t * 1e3, 3), "slices": [float(H.nSplitA), float(H.nSplitB)],
            "ref_GBps": round(2.0 * nv * 8 / t / 1e9, 1)}
        H.close()
    g = 100
    rowptr, colind, val = laplace3d(g, 1e-3)
    nr, nnz = g ** 3, len(val)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")
    dval, dcol, drow = d(val), d(colind), d(rowptr)
    xs = torch.from_numpy(oz.gen_phi(nr, 1, nr, 1.0, 5)).to("cuda:0")
    ys = torch.zeros(nr, dtype=torch.float64, device="cuda:0")
    H = oz.Handle(4 << 30, sumMode=1)
    P = H.dcsrmv_splitA(nr, nr, nnz, dval, drow)
    t = timeit(lambda: H.dcsrmv(nr, nr, nnz, 1.0, P, dcol, drow, xs, 0.0, ys), 20)
    out["dcsrmv_laplace100_presplit"] = {
        "ms": round(t * 1e3, 3), "slices": [float(H.nSplitA), float(H.nSplitB)],
        "ref_GBps": round(((nnz + 2.0 * nr + nr) * 8 + (nnz + nr + 1.0) * 4) / t / 1e9, 1)}
    H.close()
    return out


def bench_batched(count, torch, oz):
    """Config: `count` uneven Ozaki DGEMMs, m, n, k = 64 + floor(961 drand48()) (srand48(4)),
    phi = 2 inputs (seeds 1000 + 2p, 1001 + 2p), NearSum, INT8 digit form."""
    r = drand48_stream(4)
    sizes = [tuple(64 + int(961 * next(r)) for _ in range(3)) for _ in range(count)]
    As = [torch.from_numpy(oz.gen_phi(m, k, m, 2.0, 1000 + 2 * p)).to("cuda:0") for p, (m, n, k) in enumerate(sizes)]
    Bs = [torch.from_numpy(oz.gen_phi(k, n, k, 2.0, 1001 + 2 * p)).to("cuda:0") for p, (m, n, k) in enumerate(sizes)]
    Cs = [torch.zeros(m * n, dtype=torch.float64, device="cuda:0") for m, n, k in sizes]
    flops = sum(2.0 * m * n * k for m, n, k in sizes)
    ms, ns, ks = [s[0] for s in sizes], [s[1] for s in sizes], [s[2] for s in sizes]
    H = oz.Handle(1 << 30, sumMode=1, sliceType="int8")
    call = lambda: H.dgemm_batched(["N"] * count, ["N"] * count, ms, ns, ks, [1.0] * count, As, ms,
                                   Bs, ks, [0.0] * count, Cs, ms)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    call()  # plan, autotune, capture, run
    torch.cuda.synchronize()
    t_first = time.perf_counter() - t0
    ts = []
    for _ in range(5):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        call()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    tb = sorted(ts)[len(ts) // 2]
    out = {"problems": count, "sum_2mnk": flops, "tile": int(H.batchTile), "replayed": int(H.batchReplay),
           "grouped_problems": int(H.batchGrouped), "digit_pairs": int(H.nSplitC),
           "first_call_s": round(t_first, 3), "ms_per_batch": round(tb * 1e3, 3),
           "eff_tflops": round(flops / tb / 1e12, 3),
           "grouped_mfma_kernel_ms": round(H.t_comp * 1e3, 3),
           "grouped_mfma_tops": round(H.n_comp / H.t_comp / 1e12, 1) if H.t_comp > 0 else None}
    # the same problems one plain-form (FP64 slices, NearSum) ozblasDgemm at a time: the bits the
    # batch must reproduce
    Hs = oz.Handle(8 << 30, sumMode=1, sliceType="fp64")
    Cl = [torch.zeros_like(c) for c in Cs]
    for rep in range(2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for p, (m, n, k) in enumerate(sizes):
            Hs.dgemm("N", "N", m, n, k, 1.0, As[p], m, Bs[p], k, 0.0, Cl[p], m)
        torch.cuda.synchronize()
        tl = time.perf_counter() - t0
    out["plain_form_loop_ms"] = round(tl * 1e3, 3)
    out["speedup_vs_plain_form_loop"] = round(tl / tb, 2)
    out["bitwise_equal_to_plain_form_solves"] = all(
        torch.equal(a.view(torch.int64), c.view(torch.int64)) for a, c in zip(Cs, Cl))
    H.close()
    Hs.close()
    return out


def run_form(oz, work, form, scheme, n, dA, dB, steps, warmup, torch):
    """Times one fast form; returns (seconds/step, MFMA-kernel seconds, info, output tensor)."""
    dC = torch.zeros(n * n, dtype=torch.float64, device="cuda:0")
    H = oz.Handle(work, sumMode=1, reproMode=1, sliceType=form, digitScheme=scheme)
    t, tk, info = timed_gemm(H, n, dA, dB, dC, steps, warmup, torch)
    info["scheme"] = {0: "fp64", 1: "digits", 2: "residues"}[int(H.schemeUsed)]
    info["second_pass_tiles"] = int(H.fallbackTiles)
    H.close()
    return t, tk, info, dC


def form_entry(n, t, tk, info, same):
    flops = 2.0 * n ** 3
    ops = flops * info["nSplitC"]
    return {"scheme": info["scheme"], "eff_tflops": round(flops / t / 1e12, 3), "ms_per_step": round(t * 1e3, 2),
            "products": info["nSplitC"], "mfma_kernel_ms": round(tk * 1e3, 2),
            "mfma_kernel_rate": round(ops / tk / 1e12, 1), "split_ms": round(info["t_split"] * 1e3, 2),
            "reconstruct_ms": round(info["t_sum"] * 1e3, 2), "bitwise_equal_to_fp64_slice_form": same}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "the sharded design is described in DESIGN.md", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    if args.copy_gib > 0:
        copy_calibration(args.copy_gib, torch)
        return
    import ozblas_amd

    n = args.n
    work = int(args.work_gb * (1 << 30))
    res = {}
    forms = {}
    eq = lambda a, b: bool(torch.equal(a.view(torch.int64), b.view(torch.int64)))
    for phi in [args.phi] + ([args.phi2] if args.phi2 else []):
        A = ozblas_amd.gen_phi(n, n, n, phi, 1)
        B = ozblas_amd.gen_phi(n, n, n, phi, 2)
        dA = torch.from_numpy(A).to("cuda:0")
        dB = torch.from_numpy(B).to("cuda:0")
        del A, B
        steps = args.steps if phi == args.phi else max(1, args.steps // 2)
        t_fast, tk_fast, info_fast, C_fast = run_form(ozblas_amd, work, args.form, args.scheme, n, dA, dB,
                                                      steps, args.warmup, torch)
        if args.fp64_steps > 0:
            dC = torch.zeros(n * n, dtype=torch.float64, device="cuda:0")
            H64 = ozblas_amd.Handle(work, sumMode=1, reproMode=1, sliceType="fp64")
            t_64, tk_64, info_64 = timed_gemm(H64, n, dA, dB, dC, args.fp64_steps, 1, torch)
            H64.close()
            same = eq(C_fast, dC)
        else:
            dC = None
            t_64 = tk_64 = float("nan")
            info_64 = dict(nSplitA=0, nSplitB=0, nSplitC=0)
            same = None
        res[phi] = dict(t_fast=t_fast, tk_fast=tk_fast, info_fast=info_fast, t_64=t_64,
                        tk_64=tk_64, info_64=info_64, same=same, steps=steps)
        print(f"phi={phi}: {args.form} ({info_fast['scheme']}) {t_fast*1e3:.1f} ms/step (MFMA kernel "
              f"{tk_fast*1e3:.1f} ms, {info_fast['nSplitC']:.0f} products), fp64-slice "
              f"{t_64*1e3:.1f} ms ({info_64['nSplitA']:.0f}x{info_64['nSplitB']:.0f} slices), "
              f"bitwise equal: {same}", file=sys.stderr, flush=True)
        del C_fast
        if phi == args.phi and args.all_forms:
            for form in ("int8", "fp16", "bf16"):
                for scheme in ("residues", "digits"):
                    if form == args.form and scheme == info_fast["scheme"]:
                        continue
                    t, tk, info, Cf = run_form(ozblas_amd, work, form, scheme, n, dA, dB, args.form_steps,
                                               1, torch)
                    forms[f"{form}_{scheme}"] = form_entry(n, t, tk, info, eq(Cf, dC) if dC is not None else None)
                    print(f"  {form} {scheme}: {t*1e3:.1f} ms/step, {info['nSplitC']:.0f} products, "
                          f"bitwise equal: {forms[f'{form}_{scheme}']['bitwise_equal_to_fp64_slice_form']}",
                          file=sys.stderr, flush=True)
                    del Cf
        del dA, dB, dC
        torch.cuda.empty_cache()

    r = res[args.phi]
    flops = 2.0 * n ** 3
    value = flops / r["t_fast"] / 1e12
    # dominant kernel: the slice-product MFMA kernel (k_digit_gemm_i8 / k_digit_gemm_h, one launch);
    # algorithmic ops per launch = 2 m n k per product (nSplitC = moduli, or digit pairs of pass 1)
    ops = 2.0 * n * n * n * r["info_fast"]["nSplitC"]
    achieved = ops / r["tk_fast"] / 1e12
    peak = INT8_PEAK_TOPS if args.form == "int8" else HALF_PEAK_TFLOPS
    unit = "TOPS" if args.form == "int8" else "TFLOP/s"
    roof = {"bound": "mfma", "achieved": round(achieved, 1), "peak": peak,
            "unit": unit, "frac": round(achieved / peak, 4),
            "traffic": measured_traffic(n, args.phi, args.form, r["info_fast"]["scheme"]),
            "traffic_source": "profiles/ PMC summary of this workload (FETCH_SIZE + WRITE_SIZE, copy-calibrated); "
                              "not measured in this run",
            "kernel": "k_digit_gemm_i8" if args.form == "int8" else "k_digit_gemm_h",
            "note": (f"{args.form} MFMA ops; ops per launch = 2*n^3*products "
                     f"({r['info_fast']['scheme']}: {r['info_fast']['nSplitC']:.0f} products)")}
    cpu = cpu_baseline(args.cpu_n, args.phi, args.cpu_reps) if args.cpu_n > 0 else None
    extra = {}
    for phi, rr in res.items():
        e = {f"{args.form}_tflops": round(flops / rr["t_fast"] / 1e12, 3),
             "scheme": rr["info_fast"]["scheme"],
             "products": rr["info_fast"]["nSplitC"],
             f"{args.form}_mfma_kernel_ms": round(rr["tk_fast"] * 1e3, 2),
             "split_ms": round(rr["info_fast"]["t_split"] * 1e3, 2),
             "reconstruct_ms": round(rr["info_fast"]["t_sum"] * 1e3, 2)}
        if rr["same"] is not None:
            e.update({
                "fp64_slice_tflops": round(flops / rr["t_64"] / 1e12, 3),
                "speedup_vs_fp64_slice": round(rr["t_64"] / rr["t_fast"], 2),
                "bitwise_equal_to_fp64_slice_form": rr["same"],
                "fp64_slices": [rr["info_64"]["nSplitA"], rr["info_64"]["nSplitB"]],
                "fp64_mfma_kernel_ms": round(rr["tk_64"] * 1e3, 2),
                "fp64_mfma_frac": round(2.0 * n ** 3 * rr["info_64"]["nSplitC"] / rr["tk_64"] / 1e12
                                        / FP64_PEAK_TFLOPS, 4)})
        extra[f"phi{phi:g}"] = e
    if forms:
        t64 = res[args.phi]["t_64"]
        for k2, v in forms.items():
            v["speedup_vs_fp64_slice"] = round(t64 / (v["ms_per_step"] / 1e3), 2)
        extra[f"forms_phi{args.phi:g}"] = forms
    if args.cg_iters > 0:
        extra["cg_laplace100"] = bench_cg(args.cg_iters, torch, ozblas_amd)
    if args.batch > 0:
        extra[f"batched{args.batch}"] = bench_batched(args.batch, torch, ozblas_amd)
    if args.level12:
        extra["level12"] = bench_level12(torch, ozblas_amd)
    line = {
        "metric": "Ozaki DGEMM eff. TFLOPS (2mnk/t): FP16/INT8-slice vs FP64-slice, same bits",
        "value": round(value, 3), "unit": "TFLOPS", "n_gpus": 1, "steps": r["steps"],
        "warmup": args.warmup, "ms_per_step": round(r["t_fast"] * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": args.form, "data": "synthetic (reference phi-mode generator, seeds 1/2)",
        "config": {"workload": f"Ozaki DGEMM m=n=k={n}, phi={args.phi:g}, degree auto, NearSum, "
                               f"{args.form} slices ({r['info_fast']['scheme']})", "m": n, "n": n, "k": n,
                   "phi": args.phi, "sumMode": 1, "parallelism": "single GPU"},
        "roofline": roof, "cpu_baseline": cpu, "results": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
