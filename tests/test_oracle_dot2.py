"""Pins the oracle's Dot2 mode (splitEpsMode = 2; src/cpu/eft.h, dgemm_tn_skinny_x2 and csrmm_n_x2
of src/cpu/ozblas_blas_wrapper.cpp) without running the reference:
  - the error-free transformations against exact rational arithmetic;
  - the chunked Dot2 dot product against an independent transcription (Python floats are
    binary64 with round-to-nearest; fma through exact rationals) and against the Ogita-Rump-Oishi
    error bound of Dot2, which a dropped error term would violate;
  - Dot2 DOT / NRM2 / CSRMV / CG through the Ozaki driver against the exact results.
"""
from fractions import Fraction as F

import numpy as np
import pytest


def fma(a, b, c):
    return float(F(a) * F(b) + F(c))


def two_sum(a, b):
    s = a + b
    v = s - a
    return s, (a - (s - v)) + (b - v)


def dot2_py(x, y, T):
    """Dot2 over T static chunks (libgomp partition: the first n % T chunks one longer), chunks
    combined by QuadAdd in order, then QuadMul(1, .) -- written from the algorithm's definition."""
    n = len(x)
    q, r = divmod(n, T)
    G = (0.0, 0.0)
    start = 0
    for t in range(T):
        size = q + (1 if t < r else 0)
        cx, cy = 0.0, 0.0
        for i in range(start, start + size):
            h = x[i] * y[i]
            e = fma(x[i], y[i], -h)
            cx, qq = two_sum(cx, h)
            cy = cy + (qq + e)
        start += size
        tx, ty = two_sum(G[0], cx)
        ty = ty + (G[1] + cy)
        s = tx + ty
        G = (s, ty - (s - tx))
    hx = 1.0 * G[0]
    hy = fma(1.0, G[0], -hx)
    hy = fma(1.0, G[1], hy)
    s = hx + hy
    return s, hy - (s - hx)


def test_eft_primitives_exact(oracle):
    rng = np.random.default_rng(5)
    for _ in range(300):
        a, b = rng.standard_normal(2) * 2.0 ** rng.integers(-40, 40, 2)
        hi, lo = oracle.dd_dot2i(a, b)
        assert F(hi) + F(lo) == F(a) * F(b)  # TwoProdFMA + TwoSum from zero: exact
        c = (float(rng.standard_normal()), 0.0)
        c = (c[0], float(rng.standard_normal()) * 2.0 ** -60 * abs(c[0]))
        x, y = oracle.dd_quad_add(c, (hi, lo))
        exact = F(c[0]) + F(c[1]) + F(hi) + F(lo)
        assert abs(F(x) + F(y) - exact) <= abs(exact) * F(2) ** -100 + F(2) ** -1070
        assert abs(y) <= abs(np.spacing(x))  # renormalised (QuickTwoSum)
        m = float(rng.standard_normal())
        x, y = oracle.dd_quad_mul(m, (hi, lo))
        exact = F(m) * (F(hi) + F(lo))
        assert abs(F(x) + F(y) - exact) <= abs(exact) * F(2) ** -100 + F(2) ** -1070


@pytest.mark.parametrize("n,T", [(1, 1), (7, 3), (50, 1), (50, 4), (101, 8), (33, 64)])
def test_dot2_chunked_matches_transcription(oracle, n, T):
    rng = np.random.default_rng(n * 100 + T)
    x = rng.standard_normal(n) * np.exp(3 * rng.standard_normal(n))
    y = rng.standard_normal(n)
    hi, lo = oracle.dot2_chunked(x, y, T)
    want = dot2_py(list(map(float, x)), list(map(float, y)), T)
    assert (hi.hex(), lo.hex()) == (want[0].hex(), want[1].hex())


def test_dot2_error_bound(oracle):
    """|hi + lo - x.y| <= 2 u^2 |x.y| + gamma_{n-1}^2 sum|x_i y_i| (Dot2 of Ogita, Rump and Oishi,
    2005, with the final renormalisation); ill-conditioned sums make the gap to the one-fold
    (u * sum|xy|) bound visible."""
    rng = np.random.default_rng(9)
    u = 2.0 ** -53
    for n, T in ((2000, 1), (2000, 16), (20000, 1024)):
        x = rng.standard_normal(n) * np.exp(10 * rng.standard_normal(n))
        y = rng.standard_normal(n)
        y[n // 2:] = -x[: n - n // 2] * y[: n - n // 2] / x[n // 2:]  # heavy cancellation
        hi, lo = oracle.dot2_chunked(x, y, T)
        exact = sum(F(a) * F(b) for a, b in zip(x.tolist(), y.tolist()))
        absum = sum(abs(F(a) * F(b)) for a, b in zip(x.tolist(), y.tolist()))
        g = n * u / (1 - n * u)
        err = abs(F(hi) + F(lo) - exact)
        assert err <= 2 * F(u) ** 2 * abs(exact) + F(g) ** 2 * absum
        assert err < F(u) * absum * F(2) ** -20  # far below a plain FP64 dot's bound


def test_dot2_driver_dot_and_csrmv(oracle):
    """Ozaki DOT / NRM2 / CSRMV in Dot2 mode: wider slices (rho = ceil(log2(dim) / 2)), hi and lo
    of every slice pair summed by NearSum; the results are within a few ulps of the exact ones."""
    rng = np.random.default_rng(3)
    n = 5000
    x = rng.standard_normal(n) * np.exp(2 * rng.standard_normal(n))
    y = rng.standard_normal(n)
    h = oracle.Handle(1 << 30, sumMode=1, splitEpsMode=2)
    got = h.dot(n, x, y)
    ex = float(sum(F(a) * F(b) for a, b in zip(x.tolist(), y.tolist())))
    assert abs(got - ex) <= 4 * abs(np.spacing(ex))
    assert h.nSplitA <= 3 and h.nSplitB <= 3  # wide slices
    nr = h.nrm2(n, x)
    assert abs(nr - float(np.sqrt(np.float64(sum(F(a) ** 2 for a in x.tolist()))))) <= 4 * np.spacing(nr)
    rowptr, colind, val = oracle.laplace3d(10, 1e-3)
    m = 1000
    xv = rng.standard_normal(m)
    for sumMode in (0, 1):
        hd = oracle.Handle(1 << 30, sumMode=sumMode, splitEpsMode=2)
        yd = np.zeros(m)
        hd.csrmv(m, m, len(val), 1.0, val, colind, rowptr, xv, 0.0, yd)
        he = oracle.Handle(1 << 30, sumMode=1)
        ye = np.zeros(m)
        he.csrmv(m, m, len(val), 1.0, val, colind, rowptr, xv, 0.0, ye)  # RN of the exact SpMV
        assert np.all(np.abs(yd - ye) <= 2 * np.abs(np.spacing(ye)))


def test_dot2_cg_tracks_plain_cg(oracle):
    rowptr, colind, val = oracle.laplace3d(12, 1e-3)
    n, nnz = 12 ** 3, len(val)
    oracle.srand48(3)
    b = np.array([2.0 * oracle.drand48() - 1.0 for _ in range(n)])
    h2 = oracle.Handle(1 << 30, sumMode=1, splitEpsMode=2)
    hist2 = h2.cg(n, nnz, val, colind, rowptr, b, np.zeros(n), 40, 0.0)
    h0 = oracle.Handle(1 << 30, sumMode=1)
    hist0 = h0.cg(n, nnz, val, colind, rowptr, b, np.zeros(n), 40, 0.0)
    assert np.max(np.abs(hist2 - hist0) / hist0) < 1e-10
