"""The level model behind the barrier-free vector split kernels (k_vsplit_scan / k_vsplit_write).

The model: with sigma_s = 0.75 * 2^(rho+tau_s) (rho >= 3), the remainder the reference split
(src/cpu/ozblas_split.cpp:75-104,254-289) keeps after slice s is the centred residue
cmod(x, u_s) = x - u_s * rne(x / u_s) of the input itself, u_s = 2^(rho+tau_s-53); so the level
maxima are values of F(k) = max_i |cmod(x_i, 2^k)|, M_{s+1} = F(rho + tau(M_s) - 53), and
slice s = cmod(x, u_{s-1}) - cmod(x, u_s).  This restates the model in numpy and checks every
slice bit for bit, and the slice count, against the oracle's restatement of the reference loop.
"""
import numpy as np
import pytest


def cmod(x, k):
    n = np.rint(np.ldexp(x, -k))  # x * 2^-k is exact for the magnitudes used here
    return x - np.ldexp(n, k)


def model_split(oracle, x, rho, cap, reproMode):
    M = float(np.max(np.abs(x))) if len(x) else 0.0
    ks = []
    zero = False
    s = 0
    for s in range(cap):
        tau = oracle.tau(M)
        k = rho + tau - 53
        ks.append(k)
        M = float(np.max(np.abs(cmod(x, k)))) if len(x) else 0.0
        if M == 0.0:
            zero = True
            break
    else:
        s = cap
    if cap <= 1:
        count = 1
    elif zero:
        count = max(s, 1) + 1
    else:
        count = cap
    if reproMode == 0:
        count += 1
    last = s if zero else cap - 1
    slots = []
    r = x.copy()
    for q in range(last + 1):
        c = cmod(x, ks[q])
        slots.append(r - c)
        r = c
    slots.append(np.zeros_like(x) if zero else r)
    while len(slots) < count:
        slots.append(np.zeros_like(x))
    return count, slots


def oracle_split(oracle, x, nSplitMax, reproMode):
    L = oracle.lib()
    n = len(x)
    h = oracle.Handle(1 << 24, nSplitMax=nSplitMax, reproMode=reproMode)
    cap = 31
    split = np.zeros(n * (cap + 2))
    out = np.zeros(n)
    mx = np.zeros(max(n, 1))
    xin = np.ascontiguousarray(x, dtype=np.float64)
    cnt = L.ozo_split(h._h, n, 1, xin, n, out, n, split, n, mx)
    return cnt, [split[s * n:(s + 1) * n] for s in range(cnt)]


def vectors():
    rng = np.random.default_rng(11)
    yield rng.uniform(-1, 1, 1000)
    yield rng.standard_normal(777) * np.exp(8 * rng.standard_normal(777))
    yield np.ldexp(rng.integers(-2**20, 2**20, 500).astype(float), -30)  # few bits: exact early
    v = np.ldexp(1.0, rng.integers(-40, 40, 300)).astype(float)             # powers of two
    v[::3] *= -1
    yield v
    w = rng.uniform(-1, 1, 400)
    w[::7] = np.nextafter(np.ldexp(1.0, -3), 1.0)                             # just above 2^e
    w[1::11] = 0.0
    yield w
    t = np.ldexp(rng.integers(0, 2**52, 256).astype(float) + 0.5, -40)        # ties at a level
    yield t
    yield np.zeros(64)
    yield rng.uniform(-1, 1, 1) * 1e-300
    yield np.concatenate([rng.uniform(-1, 1, 100), rng.uniform(-1, 1, 100) * 1e-200])


@pytest.mark.parametrize("reproMode,nSplitMax", [(1, 0), (0, 0), (1, 3), (1, 1), (0, 2)])
def test_level_model_matches_reference_loop(oracle, reproMode, nSplitMax):
    for x in vectors():
        n = len(x)
        rho = oracle.rho(n, 0)
        cap = (nSplitMax if nSplitMax > 0 else 30) - (1 if reproMode == 0 else 0)
        want_cnt, want = oracle_split(oracle, x, nSplitMax, reproMode)
        got_cnt, got = model_split(oracle, x, rho, cap, reproMode)
        assert got_cnt == want_cnt, (n, got_cnt, want_cnt)
        for s in range(want_cnt):
            assert np.array_equal(got[s].view(np.uint64), want[s].view(np.uint64)), (n, s)


def test_model_needs_rho_at_least_3(oracle):
    """With rho < 3 the sum x + sigma can leave the binade of sigma (the residue model fails), which
    is why the kernels hand such splits (Dot2 on short vectors) to the fallback loop."""
    x = np.array([0.99, -0.75, 0.5001])
    rho = 1
    M = float(np.max(np.abs(x)))
    tau = oracle.tau(M)
    sigma = 0.75 * 2.0 ** (rho + tau)
    ref = (x + sigma) - sigma
    model = x - cmod(x, rho + tau - 53)
    assert not np.array_equal(ref, model)


def residue_key(x, k):
    """The scan kernel's integer form: x = m 2^eb, r = m mod 2^s centred (ties to even), s = k - eb,
    key = |r| 2^(64-s) = |cmod(x, 2^k)| 2^(64-k)."""
    m, e = np.frexp(abs(x))
    mi = int(np.ldexp(m, 53))
    eb = int(e) - 53
    while mi and mi % 2 == 0:  # the kernel uses the stored mantissa; the key is the same
        mi //= 2
        eb += 1
    s = k - eb
    if s <= 0:
        return 0
    r = mi & ((1 << s) - 1)
    half = 1 << (s - 1)
    if r > half or (r == half and (mi >> s) & 1):
        r = (1 << s) - r
    return r << (64 - s) if s <= 64 else None


def test_integer_residue_key_matches_cmod():
    rng = np.random.default_rng(4)
    xs = list(rng.standard_normal(200) * np.exp2(rng.integers(-60, 60, 200)))
    xs += [np.ldexp(np.float64(2**52 + 2**51 + 1), -70), 0.75, 2.0 ** -30, 3 * 2.0 ** -40]
    for x in xs:
        m, e = np.frexp(abs(x))
        for k in range(int(e) - 60, int(e) + 3):
            want = abs(float(cmod(np.array([x]), k)[0]))
            key = residue_key(x, k)
            got = np.ldexp(float(key), k - 64)
            assert got == want, (x, k, got, want)


# ---- binary32: the same model with 24 for 53 (k_vsplit_scan<.., float>) ------------------------
def model_split_f(oracle, x, rho, cap, reproMode):
    """Returns (count, slots) of the model, or None where the kernels hand the split to the
    fallback loop (rho outside [3, 24], sigma outside the binary32 normal range)."""
    x = x.astype(np.float64)
    if rho < 3 or rho > 24:
        return None
    M = float(np.max(np.abs(x))) if len(x) else 0.0
    ks = []
    zero = False
    s = 0
    for s in range(cap):
        tau = oracle.s_tau(np.float32(M))
        if rho + tau > 126 or rho + tau - 1 < -126:
            return None
        k = rho + tau - 24
        ks.append(k)
        M = float(np.max(np.abs(cmod(x, k)))) if len(x) else 0.0
        if M == 0.0:
            zero = True
            break
    else:
        s = cap
    count = 1 if cap <= 1 else (max(s, 1) + 1 if zero else cap)
    if reproMode == 0:
        count += 1
    last = s if zero else cap - 1
    slots, r = [], x.copy()
    for q in range(last + 1):
        c = cmod(x, ks[q])
        slots.append(r - c)
        r = c
    slots.append(np.zeros_like(x) if zero else r)
    while len(slots) < count:
        slots.append(np.zeros_like(x))
    return count, slots


def oracle_split_f(oracle, x, nSplitMax, reproMode):
    import ctypes as C
    L = oracle.lib()
    n = len(x)
    h = oracle.Handle(1 << 24, nSplitMax=nSplitMax, reproMode=reproMode)
    cap = 31
    split = np.zeros(n * (cap + 2), dtype=np.float32)
    out = np.zeros(n, dtype=np.float32)
    mx = np.zeros(max(n, 1), dtype=np.float32)
    xin = np.ascontiguousarray(x, dtype=np.float32)
    cnt = L.ozo_s_split(h._h, n, 1, xin, n, out, n, split, n, mx)
    return cnt, [split[s * n:(s + 1) * n] for s in range(cnt)]


def vectors_f():
    rng = np.random.default_rng(12)
    yield rng.uniform(-1, 1, 1000).astype(np.float32)
    yield (rng.standard_normal(777) * np.exp(6 * rng.standard_normal(777))).astype(np.float32)
    yield np.ldexp(rng.integers(-2**12, 2**12, 500).astype(np.float32), -20)  # few bits
    v = np.ldexp(np.float32(1), rng.integers(-30, 30, 300)).astype(np.float32)
    v[::3] *= -1
    yield v
    w = rng.uniform(-1, 1, 400).astype(np.float32)
    w[::7] = np.nextafter(np.float32(0.125), np.float32(1))
    w[1::11] = 0
    yield w
    yield np.ldexp(rng.integers(0, 2**23, 256).astype(np.float32) + np.float32(0.5), -20)  # ties
    yield np.zeros(64, dtype=np.float32)
    yield np.concatenate([rng.uniform(-1, 1, 100), rng.uniform(-1, 1, 100) * 1e-30]).astype(np.float32)
    yield np.array([1e-40, -3e-42, 1e-45, 0.0], dtype=np.float32)       # subnormal: fallback
    yield np.array([3e38, -1.0, 2.0 ** 100], dtype=np.float32)         # huge: fallback


@pytest.mark.parametrize("reproMode,nSplitMax", [(1, 0), (0, 0), (1, 3), (1, 1), (0, 2)])
def test_binary32_level_model_matches_reference_loop(oracle, reproMode, nSplitMax):
    used = 0
    for x in vectors_f():
        n = len(x)
        rho = oracle.s_rho(n, 0)
        cap = (nSplitMax if nSplitMax > 0 else 30) - (1 if reproMode == 0 else 0)
        want_cnt, want = oracle_split_f(oracle, x, nSplitMax, reproMode)
        m = model_split_f(oracle, x, rho, cap, reproMode)
        if m is None:
            continue
        used += 1
        got_cnt, got = m
        assert got_cnt == want_cnt, (n, got_cnt, want_cnt)
        for s in range(want_cnt):
            g = got[s].astype(np.float32)
            assert np.array_equal(g.astype(np.float64), got[s]), (n, s)  # slices are floats
            assert np.array_equal(g.view(np.uint32), want[s].view(np.uint32)), (n, s)
    assert used >= 7


def window_keys(x, w0, width):
    """The scan kernel's register form: keys of one element for k = w0 .. w0 + width - 1 from the
    recurrence t_s = (t_{s-1} >> 1) | bit_{s-1} << 63 (t_s = (m mod 2^s) << (64 - s), s = k - eb),
    key = t > 2^63 ? 2^64 - t : t; valid while s <= 64 (keys past the element's msb + 2 may lose
    bits below 2^-64 and then stay at or below |x|)."""
    m, e = np.frexp(abs(x))
    mi = int(np.ldexp(m, 53))
    eb = int(e) - 53
    M64 = (1 << 64) - 1
    s0 = w0 - eb
    t = (mi << (64 - s0)) & M64 if s0 >= 1 else 0
    out = []
    for kk in range(width):
        s = s0 + kk
        if kk > 0:
            bi = s - 1
            bit = (mi >> bi) & 1 if 0 <= bi < 64 else 0
            t = (t >> 1) | (bit << 63)
        out.append(((1 << 64) - t) & M64 if t > (1 << 63) else t)
    return out


def test_window_key_recurrence_matches_cmod():
    """Inside [lowest set bit + 1, msb + 2] the recurrence gives the exact scaled residue; below it
    gives 0 and above it at most |x| 2^(64-k): what the table's prefix max tolerates."""
    rng = np.random.default_rng(7)
    xs = list(rng.standard_normal(300) * np.exp2(rng.integers(-80, 80, 300)))
    xs += [0.75, 2.0 ** -30, 3 * 2.0 ** -40, np.ldexp(np.float64(2**52 + 2**51 + 1), -70),
           np.ldexp(1.0, -1074) * 12345, 2.0 ** 1000 * 1.5]
    for x in xs:
        m, e = np.frexp(abs(x))
        mi = int(np.ldexp(m, 53))
        eb = int(e) - 53
        tz = (mi & -mi).bit_length() - 1
        lo, hi = eb + tz + 1, eb + mi.bit_length() + 1
        for w0 in (lo - 5, lo, lo + 3):
            if w0 > lo:
                continue  # the kernel's window always starts at or below every element's lo
            keys = window_keys(x, w0, 64)
            for kk, key in enumerate(keys):
                k = w0 + kk
                exact = abs(float(cmod(np.array([x]), k)[0]))
                got = np.ldexp(float(key), k - 64) if key else 0.0
                if k < lo:
                    assert key == 0, (x, k)
                elif k <= hi:
                    assert got == exact, (x, k, got, exact)
                else:
                    assert got <= abs(x), (x, k)


# ---- predicted chain (k_vsplit_pre / k_vsplit_levels / k_vsplit_verify) ---------------------------
PRED = 12  # kVsPred: levels the prediction evaluates at most


def lowest_unit(x):
    """min over non-zero x of lo = the lowest k with cmod(x, 2^k) != 0 (one above the lowest set bit)."""
    los = []
    for v in x:
        if v == 0:
            continue
        m, e = np.frexp(abs(float(v)))
        mi, eb = int(np.ldexp(m, 53)), int(e) - 53
        while mi % 2 == 0:
            mi //= 2
            eb += 1
        los.append(eb + 1)
    return min(los) if los else None


def predicted_plan(oracle, x, rho, cap, prec=53, tau=None):
    """The unit exponents k_s of the split if the predicted chain k_s = k_0 - d s (d = prec - rho + 1)
    holds for every level the walk visits, else None (the kernels then run the table scan)."""
    tau = tau or oracle.tau
    x = x.astype(np.float64)
    M0 = float(np.max(np.abs(x))) if len(x) else 0.0
    lo = lowest_unit(x)
    if M0 == 0.0 or lo is None:
        return None
    k0 = rho + tau(M0) - prec
    d = prec - rho + 1
    L = (k0 - lo) // d + 2 if (d > 0 and k0 >= lo) else 1
    L = min(L, cap, PRED)
    F = [float(np.max(np.abs(cmod(x, k0 - d * s)))) for s in range(L)]
    M, ks = M0, []
    for s in range(cap):
        k = rho + tau(M) - prec
        if s >= L or k != k0 - d * s:
            return None
        ks.append(k)
        M = F[s]
        if M == 0.0:
            break
    return ks


def model_ks(oracle, x, rho, cap, prec=53, tau=None):
    tau = tau or oracle.tau
    x = x.astype(np.float64)
    M = float(np.max(np.abs(x)))
    ks = []
    for _ in range(cap):
        k = rho + tau(M) - prec
        ks.append(k)
        M = float(np.max(np.abs(cmod(x, k))))
        if M == 0.0:
            break
    return ks


def test_predicted_chain_agrees_with_the_level_walk(oracle):
    """Where the prediction is accepted its units are the walk's; the vectors of varied mantissas (the
    CG and DOT inputs) are accepted, and vectors with gaps in their bits go to the table scan."""
    rng = np.random.default_rng(5)
    cases = {
        "uniform": rng.uniform(-1, 1, 5000),
        "normal 2^20": rng.standard_normal(1 << 20),
        "wide": rng.standard_normal(3000) * np.exp(8 * rng.standard_normal(3000)),
        "one": np.array([1.0 / 3.0]),
        "powers-of-two": np.ldexp(1.0, rng.integers(-40, 40, 1500)).astype(float),
        "two-ranges": np.concatenate([rng.uniform(-1, 1, 500), rng.uniform(-1, 1, 500) * 1e-200]),
        "sparse bits": np.array([1.0 + 2.0 ** -30, -3.0]),
    }
    accepted = set()
    for name, x in cases.items():
        n = len(x)
        rho = oracle.rho(n, 0)
        for cap in (30, 3, 1):
            ks = predicted_plan(oracle, x, rho, cap)
            if ks is not None:
                assert ks == model_ks(oracle, x, rho, cap), (name, cap)
                if cap == 30:
                    accepted.add(name)
    assert {"uniform", "normal 2^20", "one"} <= accepted
    assert "sparse bits" not in accepted and "two-ranges" not in accepted
    # binary32: prec 24 and the binary32 tau rule
    x = rng.uniform(-1, 1, 4000).astype(np.float32)
    rho = oracle.s_rho(len(x), 0)
    tf = lambda M: oracle.s_tau(np.float32(M))
    ks = predicted_plan(oracle, x, rho, 30, prec=24, tau=tf)
    assert ks is not None and ks == model_ks(oracle, x, rho, 30, prec=24, tau=tf)
