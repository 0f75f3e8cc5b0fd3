"""The DOT completion sums the pair dots with NearSum spread over the lanes of a wave
(nearsum_wave in ozblas_amd/csrc/oz_nearsum.hpp): the extraction sums tau and the maxima mu reduce
in butterfly order, the remainder sum stays in index order.  This restates the reference's
ozblasTransform / ozblasTransformK / ozblasNearsum (src/cpu/ozblas_sum.cpp:10-112) in binary64 with
(a) the reference's sequential order and (b) the butterfly order over 64 lanes, and checks that
every intermediate and the result agree bit for bit on finite inputs of up to 64 terms, including
wide exponent ranges and heavy cancellation.  A butterfly for the remainder sum, or an extraction
that is not exact (sigma below 2^m max|v|), would fail it."""
import math

import numpy as np
import pytest

D = np.float64


def ceil_log2(x):
    m, e = math.frexp(float(x))
    return e - 1 if m == 0.5 else e


def transform(v, n, rho, butterfly):
    v = [D(x) for x in v]
    if n == 0:
        return v, D(rho), D(0.0), []
    if butterfly:
        mu = max(abs(x) for x in v)  # max is exact in any order
    else:
        mu = abs(v[0])
        for i in range(1, n):
            mu = mu if mu > abs(v[i]) else abs(v[i])
    if mu == 0:
        return v, D(rho), D(0.0), []
    m = math.ceil(math.log2(n + 2))
    sigma = D(math.ldexp(1.0, m + ceil_log2(mu)))
    t = D(rho)
    tiny = D(2.0 ** -1022)
    fac = D(math.ldexp(1.0, 2 * m + 1 - 53))
    shrink = D(math.ldexp(1.0, m - 53))
    taus = []
    while True:
        tmp = [(sigma + x) - sigma for x in v]
        v = [x - q for x, q in zip(v, tmp)]
        if butterfly:
            lanes = tmp + [D(0.0)] * (64 - len(tmp))
            o = 32
            while o:
                lanes = [lanes[i] + lanes[i ^ o] for i in range(64)]
                o //= 2
            tau = lanes[0]
            assert all(l == tau for l in lanes)
        else:
            tau = D(0.0)
            for q in tmp[:n]:
                tau = tau + q
        taus.append(tau)
        tau1 = t + tau
        if sigma <= tiny or abs(tau1) >= fac * sigma:
            return v, tau1, tau - (tau1 - t), taus
        sigma = shrink * sigma
        t = tau1


def transform_k(v, n, rho, butterfly):
    v, tau1, tau2, taus = transform(v, n, rho, butterfly)
    tmp = D(0.0)
    for x in v[:n]:  # index order in both forms
        tmp = tmp + x
    res = tau1 + (tau2 + tmp)
    return v, res, tau2 - (res - tau1), taus


def nearsum(vals, butterfly):
    n = len(vals)
    eps = D(2.0 ** -53)
    sign = lambda x: D(-1.0) if x < 0 else D(1.0)
    v, res, r, t1 = transform_k(vals, n, 0.0, butterfly)
    v, delta, r2, t2 = transform_k(v, n, r, butterfly)
    trace = [res, r, delta, r2] + t1 + t2
    if delta == 0:
        return res, trace
    res2 = res + sign(delta) * eps * abs(res)
    if res2 == res:
        mu = sign(delta) * eps * abs(res)
        res2 = res + D(2.0) * sign(delta) * eps * abs(res)
    else:
        mu = (res2 - res) / D(2.0)
    if abs(delta) < abs(mu):
        return res, trace
    if abs(delta) > abs(mu):
        return res2, trace
    v, delta2, tmp, t3 = transform_k(v, n, r2, butterfly)
    trace += [delta2] + t3
    if delta2 == 0:
        return res + mu, trace
    return (res2 if sign(delta2) == sign(mu) else res), trace


def cases():
    rng = np.random.default_rng(9)
    for n in (1, 2, 3, 7, 25, 36, 49, 64):
        for _ in range(30):
            yield rng.standard_normal(n) * np.exp2(rng.integers(-80, 80, n).astype(float))
        x = rng.standard_normal(n) * np.exp2(rng.integers(-30, 30, n).astype(float))
        yield np.concatenate([x, -x])[:n]  # cancellation
        big = rng.standard_normal(n)
        big[: n // 2] *= 2.0 ** 60
        yield big
        yield np.ldexp(rng.integers(-2 ** 52, 2 ** 52, n).astype(float), -1000)  # near underflow
    # pair-dot-like: integer multiples of shrinking units
    for n in (25, 36, 64):
        units = np.exp2(-17.0 * (np.arange(n) % 11))
        yield np.round(rng.standard_normal(n) * 2 ** 40) * units


def test_butterfly_nearsum_matches_sequential():
    count = 0
    for vals in cases():
        a, ta = nearsum(list(vals), butterfly=False)
        b, tb = nearsum(list(vals), butterfly=True)
        assert a.tobytes() == b.tobytes(), (vals, a, b)
        assert [x.tobytes() for x in ta] == [x.tobytes() for x in tb]
        count += 1
    assert count > 200


def test_remainder_sum_is_order_dependent():
    """The remainder sum must stay in index order: a butterfly changes its bits here."""
    vals = [D(1.0), D(2.0 ** -53), D(2.0 ** -53), D(-1.0)]
    seq = D(0.0)
    for x in vals:
        seq = seq + x
    bfly = (vals[0] + vals[1]) + (vals[2] + vals[3])
    assert seq != bfly


# ---- the wave reduction's lane network (wave_sum63_f64 / wave_max_nonneg_f64 in oz_internal.hpp) --
def dpp(vals, ctrl, row_mask=0xF):
    """One DPP move of 64 lanes (GFX9 semantics): quad_perm 0x00-0xFF, row_mirror 0x140,
    row_half_mirror 0x141, row_bcast:15 0x142, row_bcast:31 0x143; lanes of rows outside row_mask,
    or whose source does not exist, take the old value 0."""
    out = [0.0] * 64
    for i in range(64):
        row = i // 16
        if not (row_mask >> row) & 1:
            continue
        if ctrl < 0x100:
            src = (i & ~3) + ((ctrl >> (2 * (i & 3))) & 3)
        elif ctrl == 0x140:
            src = (i & ~15) + (15 - (i & 15))
        elif ctrl == 0x141:
            src = (i & ~7) + (7 - (i & 7))
        elif ctrl == 0x142:
            src = 16 * (row - 1) + 15 if row > 0 else None
        elif ctrl == 0x143:
            src = 31 if row >= 2 else None
        out[i] = vals[src] if src is not None else 0.0
    return out


STEPS = [(0xB1, 0xF), (0x4E, 0xF), (0x141, 0xF), (0x140, 0xF), (0x142, 0xA), (0x143, 0xC)]


def test_wave_reduction_network_reaches_lane_63():
    rng = np.random.default_rng(3)
    for _ in range(50):
        ints = [int(v) for v in rng.integers(-2 ** 40, 2 ** 40, 64)]
        v = [float(x) for x in ints]  # exact sums in any order
        for ctrl, mask in STEPS:
            d = dpp(v, ctrl, mask)
            v = [a + b for a, b in zip(v, d)]
        assert v[63] == float(sum(ints))
        m = [abs(float(x)) for x in rng.integers(0, 2 ** 40, 64)]
        mx = list(m)
        for ctrl, mask in STEPS:
            d = dpp(mx, ctrl, mask)
            mx = [max(a, b) for a, b in zip(mx, d)]
        assert mx[63] == max(m)
    # without the row_bcast:31 step, or with the row masks swapped, lane 63 misses rows
    for steps in (STEPS[:5], STEPS[:4] + [(0x142, 0xC), (0x143, 0xA)]):
        v = [1.0] * 64
        for ctrl, mask in steps:
            d = dpp(v, ctrl, mask)
            v = [a + b for a, b in zip(v, d)]
        assert v[63] != 64.0
