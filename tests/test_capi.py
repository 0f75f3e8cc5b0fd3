"""CPU checks of the C-ABI library: it loads, exports every entry point of include/*.h, and the
host-side rules the kernels share (tau = (short)ceil(log2(max)) via the libm-measured table, rho
of a CSR row without a log) agree with the C library the reference calls.  No GPU call is made."""
import ctypes
import glob
import math
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "ozblas_amd", "libozblas_mi355x.so")


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail("libozblas_mi355x.so is not built (run __graft_entry__.build())")
    return ctypes.CDLL(LIB)


def declared_functions():
    names = []
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*[\w\s\*]+?\b(ozblas\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_reference_entry_points():
    names = declared_functions()
    for p in "DS":
        for r in ("dot", "nrm2", "axpy", "gemv", "gemm", "csrmv", "csrmvSplitA", "cg"):
            assert f"ozblas{p}{r}" in names
    for r in ("dot", "nrm2", "gemv", "gemm", "csrmv", "csrmvSplitA", "cg"):
        assert f"ozblasDS{r}" in names
    assert "ozblasCreate" in names and "ozblasDestroy" in names


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert missing == []


def test_tau_rule_matches_libm(lib, oracle):
    """(short)ceil(log2(x)) through the table equals glibc on the first 420 values above every
    power of two (where log2 may round to the integer), on powers of two, just below them,
    on subnormals, zero, inf and NaN."""
    f = lib.ozblas_internal_tau_batch
    f.argtypes = [ctypes.c_int64, np.ctypeslib.ndpointer(np.float64), np.ctypeslib.ndpointer(np.int32)]
    xs = []
    for e in range(-1074, 1024):
        base = math.ldexp(1.0, e)
        u = math.ldexp(1.0, max(e - 52, -1074))
        nv = (1 << 52) - 1 if e - 52 >= -1074 else (1 << (e + 1074)) - 1
        xs += [base + i * u for i in range(0, min(420, nv + 1))]
        if e > -1074:
            xs.append(math.nextafter(base, 0.0))
    rng = np.random.default_rng(1)
    xs += list(np.ldexp(rng.uniform(1, 2, 20000), rng.integers(-1074, 1023, 20000)))
    xs += [0.0, -0.0, float("inf"), float("nan"), -3.0, 5e-324, 1.7976931348623157e308]
    x = np.array(xs, dtype=np.float64)
    got = np.zeros(len(x), dtype=np.int32)
    f(len(x), x, got)
    want = oracle.tau_batch(x)
    bad = np.nonzero(got != want)[0]
    assert len(bad) == 0, [(x[i], got[i], want[i]) for i in bad[:10]]


def test_rho_rules_match_libm(lib, oracle):
    lib.ozblas_internal_rho_row.restype = ctypes.c_int32
    lib.ozblas_internal_rho.restype = ctypes.c_int32
    for d in list(range(1, 5000)) + [10 ** 6, 2 ** 20, 2 ** 20 + 1, 2 ** 31 - 1, 123456789]:
        want = oracle.rho(d)
        assert lib.ozblas_internal_rho(d, 0) == want
        assert lib.ozblas_internal_rho_row(d) == want, d


def test_rho_modes_match_libm(lib, oracle):
    """getRho for splitEpsMode 2 (Dot2) and 9: the host form (glibc log2/sqrt) and the integer
    form the device uses for CSR rows agree with the oracle's libm evaluation."""
    lib.ozblas_internal_rho.restype = ctypes.c_int32
    lib.ozblas_internal_rho_row_mode.restype = ctypes.c_int32
    dims = list(range(0, 5000)) + [2 ** p + o for p in range(12, 31) for o in (-1, 0, 1)]
    dims += [10 ** 6, 123456789, 2 ** 31 - 1]
    for mode in (0, 2, 9):
        for d in dims:
            want = oracle.rho(d, mode)
            assert lib.ozblas_internal_rho(d, mode) == want, (mode, d)
            assert lib.ozblas_internal_rho_row_mode(d, mode) == want, (mode, d)


def _binary32_edge_floats():
    """The first 300 floats above every binary32 power of two (where a log2 rounded to float may
    equal the integer), the float just below, random floats, and the special values."""
    xs = []
    for e in range(-149, 128):
        base = np.ldexp(np.float32(1), e)
        u = np.ldexp(np.float32(1), max(e - 23, -149))
        nv = (1 << 23) - 1 if e - 23 >= -149 else (1 << (e + 149)) - 1
        xs += [base + np.float32(i) * u for i in range(0, min(300, nv + 1))]
        if e > -149:
            xs.append(np.nextafter(base, np.float32(0)))
    rng = np.random.default_rng(2)
    xs += list(np.ldexp(rng.uniform(1, 2, 20000), rng.integers(-149, 127, 20000)).astype(np.float32))
    xs += [0.0, -0.0, np.inf, np.nan, -3.0, 1e-45, 3.4028235e38]
    return np.array(xs, dtype=np.float32)


def test_binary32_tau_rules_match_libm(lib, oracle):
    """The binary32 split's tau = (short)ceil((float)log2((double)max)) and NearSum's
    ceil(log2f(mu)) through their tables equal glibc on every binade's critical values."""
    import ozblas_amd
    libm = ctypes.CDLL("libm.so.6")
    libm.log2f.restype = ctypes.c_float
    libm.log2f.argtypes = [ctypes.c_float]
    for x in _binary32_edge_floats():
        xf = float(x)
        assert ozblas_amd.internal_tau_f(xf) == oracle.s_tau(xf), xf
        if np.isfinite(x) and x > 0:
            want = math.ceil(libm.log2f(xf))
            assert ozblas_amd.internal_nearsum_exp_f(xf) == want, xf


def test_binary32_rho_rules_match_libm(lib, oracle):
    """getRho<float,float> for splitEpsMode 0, 2 and 9: the host form and the integer form the
    device uses for CSR rows agree with the oracle's libm evaluation."""
    import ozblas_amd
    dims = list(range(0, 5000)) + [2 ** p + o for p in range(12, 31) for o in (-1, 0, 1)]
    dims += [10 ** 6, 123456789, 2 ** 31 - 1]
    for mode in (0, 2, 9):
        for d in dims:
            want = oracle.s_rho(d, mode)
            assert ozblas_amd.internal_rho_f(d, mode) == want, (mode, d)
            assert ozblas_amd.internal_rho_row_f(d, mode) == want, (mode, d)


def test_mixed_rho_rules_match_libm(lib, oracle):
    """getRho<double,float> (binary64 data, binary32 slices) for splitEpsMode 0, 2 and 9: the host
    form and the integer form of the CSR rows agree with the oracle's libm evaluation."""
    import ozblas_amd
    dims = list(range(0, 5000)) + [2 ** p + o for p in range(12, 31) for o in (-1, 0, 1)]
    dims += [10 ** 6, 123456789, 2 ** 31 - 1]
    for mode in (0, 2, 9):
        for d in dims:
            want = oracle.ds_rho(d, mode)
            assert ozblas_amd.internal_rho_ds(d, mode) == want, (mode, d)
            assert ozblas_amd.internal_rho_row_ds(d, mode) == want, (mode, d)


def test_f32_term_is_the_scaled_double_rounding(lib):
    """f32_term(P) of an unscaled product equals the reference's route through the scaled product:
    c = (float)(P * 2^-e), then (float)scalbn((double)c, e) -- a rounding to 24 bits, then onto the
    binary32 subnormal grid below 2^-126.  Checked against numpy's float32 conversions with the
    scale e chosen to bring P into the normal range, including ties of either rounding."""
    import ozblas_amd
    rng = np.random.default_rng(11)
    mant = rng.integers(2 ** 52, 2 ** 53, 3000)
    # ties and near-ties of the 24-bit rounding and of the subnormal grid
    mant[:200] = (mant[:200] >> 29 << 29) | (1 << 28)
    mant[200:400] = (mant[200:400] >> 30 << 30) | (1 << 29)
    mant[400:600] = (mant[400:600] >> 29 << 29) | (1 << 28) | 1
    exps = rng.integers(-160, 40, mant.size)
    exps[:600] = rng.integers(-150, -124, 600)
    # the double-rounding case: half an ulp of the subnormal grid plus less than half an ulp of
    # the 24-bit one (a single rounding goes up, the reference's two go to even)
    for q in range(600, 800):
        e = int(rng.integers(-148, -127))
        drop = 53 - (e + 150)  # bits below the subnormal grid
        top = int(rng.integers(2 ** 52, 2 ** 53)) >> (drop + 1) << (drop + 1)
        mant[q] = top | (1 << (drop - 1)) | int(rng.integers(1, 2 ** 27))
        exps[q] = e
    sign = np.where(rng.random(mant.size) < 0.5, -1.0, 1.0)
    p = sign * np.ldexp(mant.astype(np.float64), exps - 52)
    # the constructed cases do differ from a single rounding
    assert sum(np.float32(v) != np.float32(np.ldexp(np.float64(np.float32(np.ldexp(v, 60))), -60))
               for v in p[600:800]) > 50
    for v in list(p) + [0.0, -0.0, 2.0 ** -126, -(2.0 ** -149), 2.0 ** -150, 3 * 2.0 ** -151, np.inf]:
        e = 0 if v == 0 or not np.isfinite(v) else int(np.floor(np.log2(abs(v))))
        c = np.float32(np.ldexp(v, -e)) if np.isfinite(v) else np.float32(v)
        want = np.float32(np.ldexp(np.float64(c), e))
        got = np.float32(ozblas_amd.internal_f32_term(float(v)))
        assert got.tobytes() == want.tobytes(), (float(v).hex(), float(got).hex(), float(want).hex())


def test_mixed_sd_rho_rules_match_libm(lib, oracle):
    """getRho<float,double> (binary32 data, binary64-held slices) for splitEpsMode 0, 2 and 9: the
    host form and the integer form of the CSR rows agree with the oracle's libm evaluation."""
    import ozblas_amd
    dims = list(range(0, 5000)) + [2 ** p + o for p in range(12, 31) for o in (-1, 0, 1)]
    dims += [10 ** 6, 123456789, 2 ** 31 - 1]
    for mode in (0, 2, 9):
        for d in dims:
            want = oracle.sd_rho(d, mode)
            assert ozblas_amd.internal_rho_sd(d, mode) == want, (mode, d)
            assert ozblas_amd.internal_rho_row_sd(d, mode) == want, (mode, d)


def test_gen_phi_matches_oracle_generator(oracle):
    import ozblas_amd
    for phi, seed in ((1.0, 1), (0.5, 2), (4.0, 7)):
        a = ozblas_amd.gen_phi(33, 9, 40, phi, seed)
        b = oracle.init_mat(33, 9, 40, phi, 2, seed)
        assert a.tobytes() == b.tobytes()


def test_exact_reconstruction_matches_big_integers(lib):
    """recon_elem (the exact reconstruction of the fast forms, host build through a hook):
    fma(alpha, RN(N 2^e0), beta c) with N = sum_b T_b 256^(Lmax - L_b), against Python integers and
    the correctly rounded float(Fraction) -- random level structures (repeated levels, more than
    128 bins of one level), extreme int32 values, subnormal and huge results."""
    from fractions import Fraction as F
    import math
    f = lib.ozblas_internal_recon
    f.restype = ctypes.c_double
    f.argtypes = [ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                  ctypes.c_double, ctypes.c_double, ctypes.c_double]
    rng = np.random.default_rng(12)
    for trial in range(3000):
        Lmax = int(rng.integers(0, 66))
        levels = []
        for L in range(Lmax + 1):
            reps = int(rng.integers(0, 3)) if trial % 50 else int(rng.integers(0, 300))
            levels += [L] * reps
        if not levels:
            levels = [0]
        T = rng.integers(-2 ** 31, 2 ** 31, len(levels)).astype(np.int64)
        if trial % 7 == 0:
            T[:] = rng.choice([-2 ** 31, 2 ** 31 - 1, 0, 1, -1], len(levels))
        T = T.astype(np.int32)
        lv = np.array(levels, dtype=np.int32)
        N = sum(int(t) * 256 ** (Lmax - int(L)) for t, L in zip(T, lv))
        bits = max(1, abs(N).bit_length())
        e0 = int(rng.integers(-1074 - bits - 5, 1000 - bits))
        alpha, beta, c = ((1.0, 0.0, 0.0) if trial % 3 else
                          (float(rng.standard_normal()), float(rng.standard_normal()), float(rng.standard_normal())))
        got = f(len(lv), T.ctypes.data, lv.ctypes.data, Lmax, e0, alpha, beta, c)
        t = float(F(N) * F(2) ** e0)
        want = float(F(alpha) * F(t) + F(beta * c))
        if want == 0.0:
            want = alpha * t + beta * c  # signed zero as fma gives it
        assert got == want and math.copysign(1, got) == math.copysign(1, want), (trial, Lmax, e0, got, want)



def test_certified_rounding_from_leading_levels(lib):
    """recon_certify: when it accepts, every N in [N1 - R, N1 + R] rounds to the returned value;
    when it refuses, the two ends round differently -- checked with Python integers and exact
    rounding on random N1 / R straddling and avoiding rounding boundaries."""
    import random
    from fractions import Fraction as F
    f = lib.ozblas_internal_recon_certify
    f.restype = ctypes.c_int32
    f.argtypes = [ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                  ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
    rng = np.random.default_rng(21)
    accepted = refused = 0
    for trial in range(2000):
        Lmax = int(rng.integers(2, 40))
        nb1 = int(rng.integers(1, Lmax))
        lv = np.arange(nb1, dtype=np.int32)
        T = rng.integers(-2 ** 31, 2 ** 31, nb1).astype(np.int32)
        N1 = sum(int(t) * 256 ** (Lmax - int(L)) for t, L in zip(T, lv))
        Rint = int(rng.integers(0, 2 ** 62)) << int(rng.integers(0, max(1, 8 * (Lmax - nb1))))
        words = np.array([(Rint >> (64 * q)) & (2 ** 64 - 1) for q in range(9)], dtype=np.uint64)
        bits = max(1, abs(N1).bit_length())
        e0 = int(rng.integers(-1074 - bits, 900 - bits))
        out = ctypes.c_double()
        ok = f(nb1, T.ctypes.data, lv.ctypes.data, Lmax, e0, words.ctypes.data, ctypes.byref(out))
        lo = float(F(N1 - Rint) * F(2) ** e0)
        hi = float(F(N1 + Rint) * F(2) ** e0)
        if ok:
            accepted += 1
            assert lo == hi == out.value
            pick = random.Random(trial).randint(-Rint, Rint)
            for N in (N1 - Rint, N1, N1 + Rint, N1 + pick):
                assert float(F(N) * F(2) ** e0) == out.value
        else:
            refused += 1
            assert lo != hi or (lo == hi == 0.0 and np.signbit(lo) != np.signbit(hi))
    assert accepted > 100 and refused > 100


def test_cpp_shim_binds_every_entry_point():
    """The cuOzBLAS C++ templates of include/cuozblas_mi355x.hpp compile (the check program built
    by build()) and resolve to the library's D, S, DS and SD entry points."""
    import subprocess
    exe = os.path.join(ROOT, "ozblas_amd", "_build", "shim_check")
    if not os.path.exists(exe):
        pytest.fail("ozblas_amd/_build/shim_check is not built (run __graft_entry__.build())")
    out = subprocess.run(["nm", "-D", "--undefined-only", exe], capture_output=True, text=True,
                         check=True).stdout
    used = set(re.findall(r"\b(ozblas\w+)\b", out))
    for p in ("D", "S", "DS", "SD"):
        for r in ("dot", "nrm2", "gemv", "gemm", "csrmv", "csrmvSplitA", "cg"):
            assert f"ozblas{p}{r}" in used
