"""The CG diagnostics of ozblasRcg (src/cpu/ozblas_cg.cpp:4-21,40-112) on the GPU: the second
history cg_verbose2 (the true relative residual with trueresFlag, the elapsed seconds otherwise)
and the time totals of the handle.  The true residual is checked bit for bit against an exact
rational evaluation of the same rule: t = b - A x by the plain CSRMV's fma chain, |t| as
sqrt(RN(sum t_i^2)), divided by |b|, at the x the oracle's solver reaches after j iterations."""
import math
import time
from fractions import Fraction as F

import numpy as np
import pytest

from gpu_util import bits_equal, dev, first_diff, say, torch_cuda

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def oz():
    torch_cuda()
    import ozblas_amd
    return ozblas_amd


def rn_f(q):
    """Round a Fraction to the nearest binary32, ties to even."""
    if q == 0:
        return np.float32(0.0)
    s = -1 if q < 0 else 1
    q = abs(q)
    e = q.numerator.bit_length() - q.denominator.bit_length()
    while F(2) ** e > q:
        e -= 1
    while F(2) ** (e + 1) <= q:
        e += 1
    e = max(e, -126)
    ulp = F(2) ** (e - 23)
    k = q / ulp
    fl = k.numerator // k.denominator
    rem = k - fl
    if rem > F(1, 2) or (rem == F(1, 2) and fl % 2 == 1):
        fl += 1
    return np.float32(s * float(fl * ulp))


def true_residual(val, colind, rowptr, x, b, nrmb, f32):
    """|b - A x| / |b| as the library defines it, exactly evaluated then rounded step by step."""
    n = len(b)
    acc = F(0)
    for r in range(n):
        tr = 0.0
        for i in range(rowptr[r], rowptr[r + 1]):
            e = F(float(val[i])) * F(float(x[colind[i]])) + F(float(tr))
            tr = float(np.float32(float(e))) if f32 else float(e)  # fma1: double fma (then float)
        e = F(-1) * F(tr) + F(float(b[r]))
        t = float(np.float32(float(e))) if f32 else float(e)
        acc += F(t) * F(t)
    if f32:
        nrm = np.float32(math.sqrt(float(rn_f(acc))))
        return np.float32(nrm / np.float32(nrmb))
    return math.sqrt(float(acc)) / nrmb


@pytest.mark.parametrize("prec", ["d", "s"])
def test_cg_true_residual_history(oz, oracle, prec):
    g, iters = 8, 12
    f32 = prec == "s"
    dt = np.float32 if f32 else np.float64
    rowptr, colind, val = oracle.laplace3d(g, 1e-3)
    val = val.astype(dt)
    n, nnz = g ** 3, len(val)
    oracle.srand48(3)
    b = np.array([2.0 * oracle.drand48() - 1.0 for _ in range(n)], dtype=dt)
    H = oz.Handle(1 << 28, sumMode=1)
    dx = dev(np.zeros(n, dtype=dt))
    say(f"{prec}cg trueres n={n} iters={iters}")
    run = H.scg if f32 else H.dcg
    hist, tr = run(n, nnz, dev(val), dev(colind), dev(rowptr), dev(b), dx, iters, 0.0, second="trueres")
    assert H.cgDevicePath == 0
    solve = oracle.Handle(1 << 30, sumMode=1)
    want_hist = (solve.scg if f32 else solve.cg)(n, nnz, val, colind, rowptr, b, np.zeros(n, dtype=dt),
                                                  iters, dt(0))
    assert bits_equal(hist.astype(np.float64), np.asarray(want_hist, dtype=dt).astype(np.float64))
    nb = F(0)
    for v in b:
        nb += F(float(v)) ** 2
    nrmb = np.float32(math.sqrt(float(rn_f(nb)))) if f32 else math.sqrt(float(nb))
    for j in (0, 1, 5, iters):
        xj = np.zeros(n, dtype=dt)
        if j:
            h = oracle.Handle(1 << 30, sumMode=1)
            (h.scg if f32 else h.cg)(n, nnz, val, colind, rowptr, b, xj, j, dt(0))
        want = true_residual(val, colind, rowptr, xj, b, nrmb, f32)
        assert np.asarray(tr[j], dtype=dt).tobytes() == np.asarray(want, dtype=dt).tobytes(), (j, tr[j], want)


@pytest.mark.parametrize("path", ["device", "host"])
def test_cg_time_history_and_totals(oz, oracle, path):
    g, iters = 10, 30
    rowptr, colind, val = oracle.laplace3d(g, 1e-3)
    n, nnz = g ** 3, len(val)
    oracle.srand48(5)
    b = np.array([2.0 * oracle.drand48() - 1.0 for _ in range(n)])
    H = oz.Handle(1 << 28, sumMode=1, cgDeviceSlices=0 if path == "device" else -1)
    dx = dev(np.zeros(n))
    say(f"dcg times path={path}")
    t0 = time.perf_counter()
    hist, secs = H.dcg(n, nnz, dev(val), dev(colind), dev(rowptr), dev(b), dx, iters, 0.0,
                       second="times")
    wall = time.perf_counter() - t0
    assert H.cgDevicePath == (1 if path == "device" else 0)
    want = oracle.Handle(1 << 30, sumMode=1).cg(n, nnz, val, colind, rowptr, b, np.zeros(n), iters, 0.0)
    assert bits_equal(hist, want), first_diff(hist, want)
    assert len(secs) == iters + 1
    assert secs[0] > 0 and np.all(np.diff(secs) >= 0) and secs[-1] <= wall
    assert H.t_SplitMat_total > 0
    # the category totals on both paths (the device path times its categories on the GPU clock)
    totals = (H.t_SpMV_SpMM_total, H.t_DOT_NRM2_total, H.t_AXPY_SCAL_total)
    assert all(t > 0 for t in totals), totals
    assert sum(totals) + H.t_SplitMat_total <= wall, (totals, wall)
