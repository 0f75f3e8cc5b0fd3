"""Pins of the oracle beyond the reference logs: closed forms, invariants and libc cross-checks.

- NearSum (src/cpu/ozblas_sum.cpp:82-112) == round-to-nearest-even of the exact sum, checked with
  Python fractions (independent of the oracle's own exact accumulator) incl. ties, cancellation,
  power-of-two boundaries and subnormals.
- the exact accumulator == fractions.
- generators == glibc drand48 / rand through ctypes (the functions the reference harness calls).
- the split: slices sum exactly to the input, each slice fits the bit budget implied by rho.
- CG under NearSum == the closed form (correctly rounded SpMV/DOT + IEEE fma/div/sqrt), computed
  here with an independent implementation.
"""
import ctypes
import math
import random
from fractions import Fraction

import numpy as np
import pytest


def rn_even(fr: Fraction) -> float:
    """Round a rational to the nearest binary64, ties to even (independent of the oracle)."""
    if fr == 0:
        return 0.0
    sign = -1.0 if fr < 0 else 1.0
    a = abs(fr)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    if Fraction(2) ** e > a:
        e -= 1
    q = max(e - 52, -1074)
    scaled = a / (Fraction(2) ** q)
    fl = scaled.numerator // scaled.denominator
    rem = scaled - fl
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and fl % 2 == 1):
        fl += 1
    return sign * math.ldexp(float(fl), q)


def fma(a, b, c):
    """IEEE fused multiply-add via exact rationals (math.fma needs Python 3.13)."""
    return rn_even(Fraction(a) * Fraction(b) + Fraction(c))


def exact(vals):
    return sum((Fraction(v) for v in vals), Fraction(0))


def _adversarial_vectors(rng):
    out = []
    for _ in range(300):
        n = rng.randint(1, 40)
        kind = rng.randint(0, 5)
        if kind == 0:
            v = [rng.uniform(-1, 1) * 2.0 ** rng.randint(-60, 60) for _ in range(n)]
        elif kind == 1:  # heavy cancellation
            base = [rng.uniform(-1, 1) * 2.0 ** rng.randint(-20, 20) for _ in range(n)]
            v = base + [-x for x in base] + [rng.uniform(-1, 1) * 2.0 ** -70]
        elif kind == 2:  # exact ties: 1 + 2^-53 and friends
            v = [1.0, 2.0 ** -53] + [0.0] * (n % 3)
            if rng.random() < 0.5:
                v.append(2.0 ** -200)
        elif kind == 3:  # near powers of two
            e = rng.randint(-30, 30)
            v = [2.0 ** e, -(2.0 ** (e - 54)), 2.0 ** (e - 110) * rng.choice([1, -1])]
        elif kind == 4:  # subnormal range
            v = [rng.uniform(-1, 1) * 2.0 ** rng.randint(-1074, -1000) for _ in range(n)]
        else:  # integers with huge spread
            v = [float(rng.randint(-2 ** 53, 2 ** 53)) * 2.0 ** rng.randint(-300, 300)
                 for _ in range(n)]
        out.append(v)
    return out


def test_nearsum_is_rn_of_exact(oracle):
    rng = random.Random(7)
    for v in _adversarial_vectors(rng):
        want = rn_even(exact(v))
        got = oracle.nearsum(v)
        assert got == want or (got == 0 and want == 0), (v, got, want)


def test_exact_accumulator_matches_fractions(oracle):
    rng = random.Random(11)
    for v in _adversarial_vectors(rng):
        want = rn_even(exact(v))
        assert oracle.exact_sum_rn(np.array(v)) == want
    for _ in range(100):
        n = rng.randint(1, 30)
        x = [rng.uniform(-1, 1) * 2.0 ** rng.randint(-500, 500) for _ in range(n)]
        y = [rng.uniform(-1, 1) * 2.0 ** rng.randint(-500, 500) for _ in range(n)]
        want = rn_even(sum((Fraction(a) * Fraction(b) for a, b in zip(x, y)), Fraction(0)))
        assert oracle.exact_dot_rn(np.array(x), np.array(y)) == want


def test_drand48_and_rand_match_glibc(oracle):
    libc = ctypes.CDLL("libc.so.6")
    libc.drand48.restype = ctypes.c_double
    for seed in (0, 1, 2, 3, 12345, 2 ** 31 + 5):
        libc.srand48(ctypes.c_long(seed))
        oracle.srand48(seed)
        for _ in range(1000):
            assert libc.drand48() == oracle.drand48()
    libc.srand(1)
    oracle.srand(1)
    for _ in range(5000):
        assert libc.rand() == oracle.rand()


def test_init_mat_phi_mode_matches_glibc_generator(oracle):
    """mublasInitMat mode 2 (testing/testing_common.cpp:703-717) with glibc's own drand48."""
    libc = ctypes.CDLL("libc.so.6")
    libc.drand48.restype = ctypes.c_double
    libm = ctypes.CDLL("libm.so.6")
    for f in ("log", "sin", "exp", "sqrt"):
        getattr(libm, f).restype = ctypes.c_double
        getattr(libm, f).argtypes = [ctypes.c_double]
    for phi, seed in ((1.0, 1), (0.5, 2), (4.0, 3)):
        got = oracle.init_mat(7, 5, 7, phi, 2, seed)
        libc.srand48(ctypes.c_long(seed))
        want = []
        for _ in range(35):
            r1 = libc.drand48()
            r2 = libc.drand48()
            x1 = 0.0 + (libm.sqrt(-2.0 * libm.log(r1)) * libm.sin(2.0 * math.pi * r2)) * 1.0
            want.append((r1 - 0.5) * libm.exp(phi * x1))
        assert got.tolist() == want


def test_rho_closed_form(oracle):
    # rho = ceil(53 - (53 - log2 dim)/2): k=256 -> 31, 1024/2048 -> 32, 4096/8192 -> 33, 1e6 -> 37
    for dim, want in ((1, 27), (256, 31), (1024, 32), (2048, 32), (4096, 33), (8192, 33),
                      (10 ** 6, 37), (4, 28), (7, 28)):
        assert oracle.rho(dim) == want


def test_split_invariants(oracle):
    rng = np.random.default_rng(3)
    for k, n, phi in ((37, 5, 1.0), (256, 9, 4.0), (1000, 3, 0.5), (64, 1, 2.0)):
        h = oracle.Handle(1 << 28, reproMode=1)
        src = oracle.init_mat(k, n, k, phi, 2, int(rng.integers(1, 1000)))
        lds = oracle.pitch(k)
        split = np.zeros(lds * n * 32, dtype=np.float64)
        work = src.copy()
        mx = np.zeros(max(k, n), dtype=np.float64)
        ns = oracle.lib().ozo_split(h._h, k, n, src, k, work, k, split, lds, mx)
        rho = oracle.rho(k)
        total = [[Fraction(0)] * k for _ in range(n)]
        for s in range(ns):
            for j in range(n):
                col = split[lds * n * s + lds * j: lds * n * s + lds * j + k]
                nz = col[col != 0]
                if len(nz):
                    # every slice element is a multiple of the slice unit, with <= 53-rho bits
                    e = max(math.frexp(float(abs(x)))[1] for x in nz)
                    unit_e = min(math.frexp(float(x))[1] - 53 + _tz(float(x)) for x in nz)
                    assert e - unit_e <= 53 - rho + 1
                for i in range(k):
                    total[j][i] += Fraction(float(col[i]))
        for j in range(n):
            for i in range(k):
                assert total[j][i] == Fraction(float(src[j * k + i]))


def _tz(x):
    m, e = math.frexp(x)
    mi = int(abs(m) * 2 ** 53)
    return (mi & -mi).bit_length() - 1


def _cg_closed_form(oracle, g, b, iters):
    """CG with correctly rounded SpMV/DOT and IEEE fma/div/sqrt (src/cpu/ozblas_cg.cpp order)."""
    rowptr, colind, val = oracle.laplace3d(g, 1e-3)
    n = g ** 3

    def spmv(x):
        out = np.empty(n)
        for r in range(n):
            s, e = rowptr[r], rowptr[r + 1]
            out[r] = oracle.exact_dot_rn(val[s:e], x[colind[s:e]])
        return out

    def dot(x, y):
        return oracle.exact_dot_rn(x, y)

    x = np.zeros(n)
    nrmb = math.sqrt(dot(b, b))
    t = spmv(x)
    r = np.array([fma(-1.0, t[i], 1.0 * b[i]) for i in range(n)])
    dold = dot(r, r)
    hist = [math.sqrt(dold) / nrmb]
    p = r.copy()
    for _ in range(iters):
        q = spmv(p)
        alpha = dold / dot(p, q)
        x = np.array([fma(alpha, p[i], x[i]) for i in range(n)])
        r = np.array([fma(-alpha, q[i], r[i]) for i in range(n)])
        dnew = dot(r, r)
        beta = dnew / dold
        dold = dnew
        p = np.array([fma(1.0, r[i], beta * p[i]) for i in range(n)])
        hist.append(math.sqrt(dnew) / nrmb)
    return hist, x


def test_cg_nearsum_closed_form(oracle):
    g, iters = 5, 12
    rowptr, colind, val = oracle.laplace3d(g, 1e-3)
    n = g ** 3
    oracle.srand48(3)
    b = np.array([2.0 * oracle.drand48() - 1.0 for _ in range(n)])
    h = oracle.Handle(1 << 28, sumMode=1, reproMode=1)
    x = np.zeros(n)
    hist = h.cg(n, len(val), val, colind, rowptr, b, x, iters, 0.0)
    want, xw = _cg_closed_form(oracle, g, b, iters)
    assert hist.tolist() == want
    assert x.tolist() == xw.tolist()


def test_csrmv_nearsum_is_correctly_rounded(oracle):
    g = 6
    rowptr, colind, val = oracle.laplace3d(g, 1e-3)
    n = g ** 3
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, n) * np.exp2(rng.integers(-30, 30, n))
    for shift_tuned in (False, True):
        h = oracle.Handle(1 << 28, sumMode=1, reproMode=1)
        if shift_tuned:
            ptr = h.csrmv_splitA(n, n, len(val), val, rowptr)
            y = np.zeros(n)
            h.csrmv(n, n, len(val), 1.0, ptr, colind, rowptr, x, 0.0, y)
        else:
            y = np.zeros(n)
            h.csrmv(n, n, len(val), 1.0, val, colind, rowptr, x, 0.0, y)
        for r in range(n):
            s, e = rowptr[r], rowptr[r + 1]
            assert y[r] == oracle.exact_dot_rn(val[s:e], x[colind[s:e]])


def test_laplacian_structure(oracle):
    g = 4
    rowptr, colind, val = oracle.laplace3d(g, 1e-3)
    assert len(val) == 7 * g ** 3 - 6 * g ** 2
    for r in range(g ** 3):
        cols = colind[rowptr[r]:rowptr[r + 1]]
        assert list(cols) == sorted(cols)
        assert val[rowptr[r]:rowptr[r + 1]].sum() == pytest.approx(1e-3 + 6.0 - (len(cols) - 1))


def test_openblas_slice_products_same_bits(oracle):
    """The CPU baseline runs the oracle's slice products through OpenBLAS's cblas_dgemm, as the
    reference does: the products are exact, so every sum mode and transposition keeps its bits."""
    m, n, k = 150, 90, 300
    A = oracle.init_mat(k, m, k, 2.0, 2, 1)
    B = oracle.init_mat(k, n, k, 2.0, 2, 2)
    outs = []
    for use in (False, True):
        if use and oracle.use_openblas(True) is None:
            pytest.skip("scipy's OpenBLAS is not present")
        try:
            for sumMode in (0, 1):
                for ta in "NT":
                    h = oracle.Handle(1 << 30, sumMode=sumMode)
                    C = np.zeros(m * n)
                    if ta == "T":
                        h.gemm("T", "N", m, n, k, 1.0, A, k, B, k, 0.0, C, m)
                    else:
                        At = A.reshape(m, k).T.copy().ravel()
                        h.gemm("N", "N", m, n, k, 1.0, At, m, B, k, 0.0, C, m)
                    outs.append((use, sumMode, ta, C))
        finally:
            oracle.use_openblas(False)
    half = len(outs) // 2
    for a, b in zip(outs[:half], outs[half:]):
        assert a[3].view(np.uint64).tolist() == b[3].view(np.uint64).tolist(), a[1:3]
