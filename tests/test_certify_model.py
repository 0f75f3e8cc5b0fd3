"""CPU model of the binary64 fast certification of the INT8 form's early termination
(recon_certify_fast in ozblas_amd/csrc/oz_gemm_digits.hip), restated operation for operation in
IEEE binary64 (numpy), against the exact rule it must agree with: when it accepts, every N in
[N1 - R, N1 + R] rounds (to nearest, ties to even) to the value it returns.  Inputs are drawn
near rounding boundaries of binary64, with ties, powers of two, cancellation and large R."""
from fractions import Fraction

import numpy as np


def fast(tv, levels, Lmax, r_up, e0):
    hi = lo = asum = np.float64(0.0)
    for T, L in zip(tv, levels):
        tb = np.ldexp(np.float64(T), 8 * (Lmax - L))
        s = hi + tb
        bb = s - hi
        err = (hi - (s - bb)) + (tb - bb)
        hi = s
        lo = lo + err
        asum = asum + abs(tb)
    h2 = hi + lo
    bb = h2 - hi
    l2 = (hi - (h2 - bb)) + (lo - bb)
    if h2 == 0.0:
        return None
    f, ex = np.frexp(h2)
    ex = int(ex)
    if ex + e0 < -1020 or ex + e0 > 1022:
        return None
    nb = np.float64(len(tv))
    delta = np.float64(4.0) * nb * nb * np.float64(2.0 ** -106) * asum * np.float64(1.0 + 2.0 ** -40)
    half = np.ldexp(np.float64(1.0), ex - (55 if abs(f) == 0.5 else 54))
    if abs(l2) + delta + r_up >= half * np.float64(1.0 - 2.0 ** -20):
        return None
    return float(np.ldexp(h2, e0))


def exact_rn(N, e0):
    return float(Fraction(N) * Fraction(2) ** e0)


def test_fast_certification_agrees_with_exact_rounding():
    rng = np.random.default_rng(2024)
    accepted = 0
    for trial in range(20000):
        Lmax = int(rng.integers(3, 20))
        nb = int(rng.integers(1, 14))
        levels = sorted(int(x) for x in rng.integers(0, Lmax + 1, nb))
        mode = trial % 4
        tv = [int(x) for x in rng.integers(-2 ** 31, 2 ** 31, nb)]
        if mode == 1:  # cancellation: a pair of opposite terms at the top
            tv[0] = 2 ** 30
            if nb > 1:
                tv[1] = -2 ** 30 if levels[1] == levels[0] else tv[1]
        N1 = sum(T * 2 ** (8 * (Lmax - L)) for T, L in zip(tv, levels))
        if mode == 2 and N1 != 0:  # N1 steered onto a binary64 tie or boundary
            q = abs(N1).bit_length() - 53
            if q > 0:
                base = (N1 >> q) << q
                N1b = base + (1 << (q - 1)) * int(rng.integers(0, 3))
                tv[-1] += (N1b - N1) // 2 ** (8 * (Lmax - levels[-1]))
                tv[-1] = max(min(tv[-1], 2 ** 31 - 1), -2 ** 31)
                N1 = sum(T * 2 ** (8 * (Lmax - L)) for T, L in zip(tv, levels))
        if mode == 3 and N1 != 0:  # a power of two
            N1 = 2 ** (abs(N1).bit_length() - 1) * (1 if N1 > 0 else -1)
            tv = [0] * (nb - 1) + [N1 // 2 ** (8 * (Lmax - levels[-1]))]
            if abs(tv[-1]) >= 2 ** 31:
                continue
            N1 = sum(T * 2 ** (8 * (Lmax - L)) for T, L in zip(tv, levels))
        R = int(rng.integers(0, 2 ** 20)) * 2 ** int(rng.integers(0, 8 * Lmax + 1))
        r_up = float(R) * (1.0 + 2.0 ** -40)
        e0 = int(rng.integers(-1100, 900))
        got = fast(tv, levels, Lmax, np.float64(r_up), e0)
        if got is None:
            continue
        accepted += 1
        lo, hi = exact_rn(N1 - R, e0), exact_rn(N1 + R, e0)
        assert lo == hi == got, (tv, levels, Lmax, R, e0, got, lo, hi)
    assert accepted > 2000
