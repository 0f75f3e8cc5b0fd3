"""Pins the binary32 oracle restatement (oracle/ozoracle_s.c) beyond the reference's binary32 logs
(tests/test_oracle_kat.py replays 136 of them):
  - NearSum in binary32 returns the correctly rounded (ties-to-even) float of the exact sum;
  - the float split is error-free: the slices of every vector / column sum to the input exactly,
    and each slice sits on the grid of its sigma;
  - DOT / CSRMV under NearSum are the correctly rounded float of the exact result;
  - CG under NearSum equals an independent float CG built from correctly rounded DOT/SpMV and the
    reference's scalar operations (float division, sqrt and fma evaluated in double).
"""
from fractions import Fraction as F
import math

import numpy as np


def fma_d(a, b, c):
    """binary64 fused multiply-add (correctly rounded through exact rationals)"""
    return float(F(float(a)) * F(float(b)) + F(float(c)))


def rn_f(q):
    """Round a Fraction to the nearest binary32, ties to even (normal and subnormal range)."""
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


def test_s_nearsum_is_correctly_rounded(oracle):
    rng = np.random.default_rng(2)
    for t in range(600):
        n = int(rng.integers(1, 60))
        v = (rng.standard_normal(n) * np.exp2(rng.integers(-20, 20, n))).astype(np.float32)
        if t % 3 == 0:
            v = np.concatenate([v, -v[: n // 2]])
            v[0] = np.float32(v[0] + np.float32(2.0 ** -14))
        if t % 7 == 0:  # ties at the last place
            v = np.array([1.0, 2.0 ** -24, 2.0 ** -60], dtype=np.float32)[: 2 + (t % 2)]
        got = oracle.s_nearsum(v)
        want = rn_f(sum(F(float(a)) for a in v))
        assert got.tobytes() == want.tobytes(), (t, got, want)


def test_s_tau_rule(oracle):
    """(short)ceil((float)log2((double)x)): the float rounding of log2 maps the first floats above
    some powers of two to the power itself."""
    for e in range(-120, 120, 7):
        for off in (0, 1, 2, 5, 40):
            x = np.float32(np.ldexp(np.float32(1.0) + np.float32(off) * np.float32(2.0 ** -23), e))
            want = math.ceil(float(np.float32(math.log2(float(x)))))
            assert oracle.s_tau(x) == want, (e, off)
    assert oracle.s_tau(0.0) == 0


def test_s_split_is_error_free(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(3)
    for n, cols in ((1000, 1), (77, 9), (5, 3)):
        x = (rng.standard_normal(n * cols) * np.exp2(rng.integers(-12, 12, n * cols))).astype(np.float32)
        h = oracle.Handle(1 << 24, reproMode=1)
        cap = 31
        split = np.zeros(n * cols * (cap + 3), dtype=np.float32)
        out = np.zeros(n * cols, dtype=np.float32)
        mx = np.zeros(max(n, cols), dtype=np.float32)
        cnt = L.ozo_s_split(h._h, n, cols, x, n, out, n, split, n, mx)
        for j in range(cols):
            for i in range(n):
                tot = sum(F(float(split[s * n * cols + j * n + i])) for s in range(cnt))
                assert tot == F(float(x[j * n + i]))


def test_s_dot_and_csrmv_are_correctly_rounded(oracle):
    rng = np.random.default_rng(4)
    for n in (1, 13, 400, 3000):
        x = (rng.standard_normal(n) * np.exp2(rng.integers(-8, 8, n))).astype(np.float32)
        y = rng.standard_normal(n).astype(np.float32)
        h = oracle.Handle(1 << 28, sumMode=1)
        got = h.sdot(n, x, y)
        assert got.tobytes() == rn_f(sum(F(float(a)) * F(float(b)) for a, b in zip(x, y))).tobytes()
    rowptr, colind, val = oracle.laplace3d(6, 1e-3)
    m = len(rowptr) - 1
    v32 = val.astype(np.float32)
    xv = rng.uniform(-1, 1, m).astype(np.float32)
    yv = np.zeros(m, dtype=np.float32)
    oracle.Handle(1 << 28, sumMode=1).scsrmv(m, m, len(val), np.float32(1.0), v32, colind, rowptr, xv,
                                              np.float32(0.0), yv)
    for r in range(m):
        ex = sum(F(float(v32[i])) * F(float(xv[colind[i]])) for i in range(rowptr[r], rowptr[r + 1]))
        assert yv[r].tobytes() == rn_f(ex).tobytes(), r


def test_s_cg_matches_closed_form(oracle):
    """Under NearSum the DOT and the SpMV rows are RN(exact); the rest is the reference's float
    recurrence (src/cpu/ozblas_cg.cpp:61-112), with fma1 and sqrt evaluated in double."""
    rowptr, colind, val = oracle.laplace3d(5, 1e-3)
    n = len(rowptr) - 1
    v32 = val.astype(np.float32)
    oracle.srand48(3)
    b = np.array([2.0 * oracle.drand48() - 1.0 for _ in range(n)], dtype=np.float32)
    iters = 25
    h = oracle.Handle(1 << 28, sumMode=1)
    xo = np.zeros(n, dtype=np.float32)
    hist = h.scg(n, len(val), v32, colind, rowptr, b, xo, iters, np.float32(0.0))

    def dot(u, w):
        return rn_f(sum(F(float(a)) * F(float(c)) for a, c in zip(u, w)))

    def spmv(u, alpha, beta, yy):
        out = np.empty(n, dtype=np.float32)
        for r in range(n):
            t = rn_f(sum(F(float(v32[i])) * F(float(u[colind[i]])) for i in range(rowptr[r], rowptr[r + 1])))
            out[r] = np.float32(fma_d(float(alpha), float(t), float(np.float32(beta * yy[r]))))
        return out

    def fma1(a, x, y):  # (float)fma((double)a, (double)x, (double)y), elementwise
        return np.array([fma_d(float(a), float(p), float(q)) for p, q in zip(x, y)]).astype(np.float32)

    x = np.zeros(n, dtype=np.float32)
    nrmb = np.float32(math.sqrt(float(dot(b, b))))
    r = spmv(x, np.float32(-1.0), np.float32(1.0), b)
    dold = dot(r, r)
    want = [np.float32(np.float32(math.sqrt(float(dold))) / nrmb)]
    p = r.copy()
    for _ in range(iters):
        q = spmv(p, np.float32(1.0), np.float32(0.0), np.zeros(n, dtype=np.float32))
        alpha = np.float32(dold / dot(p, q))
        x = fma1(alpha, p, x)
        r = fma1(-alpha, q, r)
        dnew = dot(r, r)
        resi = np.float32(math.sqrt(float(dnew)))
        beta = np.float32(dnew / dold)
        dold = dnew
        p = fma1(np.float32(1.0), r, (beta * p).astype(np.float32))
        want.append(np.float32(resi / nrmb))
    assert np.array_equal(hist.view(np.uint32), np.array(want, dtype=np.float32).view(np.uint32))
    assert np.array_equal(xo.view(np.uint32), x.view(np.uint32))
