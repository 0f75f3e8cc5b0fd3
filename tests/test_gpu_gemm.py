"""GPU parity of DGEMM/DGEMV through the C ABI against the oracle (bitwise).

Plain form (FP64 slices): bitwise identical to the oracle restatement of the reference CPU OzBLAS
for every transposition, summation mode (Fsum, NearSum), reproMode, fastMode, degree and shape
below.  Fast forms (INT8/FP16/BF16 digit slices): bitwise identical to the plain form under
NearSum, i.e. RN of the exact product, checked against the oracle's exact accumulator too.
"""
import types

import numpy as np
import pytest

from gpu_util import bits_equal, dev, first_diff, host, say, torch_cuda

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def oz():
    torch_cuda()
    import ozblas_amd
    return ozblas_amd


def _shape(trans, rows, cols):
    return (rows, cols) if trans in "Nn" else (cols, rows)


_ORACLE_CACHE = {}
_DEVICE_ONLY = ("digitLevels", "digitScheme")  # handle fields the oracle has no counterpart for


def run_pair(oz, oracle, ta, tb, m, n, k, phi=1.0, seeds=(1, 2), alpha=1.0, beta=0.0, cinit=None,
             slice_type="fp64", ws=1 << 30, **flags):
    ar, ac = _shape(ta, m, k)
    br, bc = _shape(tb, k, n)
    lda, ldb, ldc = ar + 3, br + 1, m + 2
    A = oracle.init_mat(ar, ac, lda, phi, 2, seeds[0])
    B = oracle.init_mat(br, bc, ldb, phi, 2, seeds[1])
    C0 = np.zeros(ldc * n) if cinit is None else cinit(ldc * n)
    oflags = {f: v for f, v in flags.items() if f not in _DEVICE_ONLY}
    # the oracle's result of a case is computed once (the fast forms compare against the same one)
    key = (ta, tb, m, n, k, phi, seeds, alpha, beta, tuple(sorted(oflags.items())),
           None if cinit is None else C0.tobytes())
    if key not in _ORACLE_CACHE:
        ho = oracle.Handle(1 << 31, **oflags)
        Cw = C0.copy()
        ho.gemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, Cw, ldc)
        _ORACLE_CACHE[key] = (Cw, types.SimpleNamespace(nSplitA=ho.nSplitA, nSplitB=ho.nSplitB))
        ho.close()
    Cw, h = _ORACLE_CACHE[key]
    Cw = Cw.copy()
    H = oz.Handle(ws, sliceType=slice_type, **flags)
    dA, dB, dC = dev(A), dev(B), dev(C0)
    H.dgemm(ta, tb, m, n, k, alpha, dA, lda, dB, ldb, beta, dC, ldc)
    got = host(dC)
    return got, Cw, H, h


CASES = [
    # (m, n, k, phi)
    (256, 256, 256, 1.0),
    (77, 133, 199, 4.0),
    (130, 129, 300, 0.5),
    (1, 1, 37, 1.0),
    (64, 1, 300, 2.0),
    (33, 17, 1, 1.0),
    (5, 300, 64, 3.0),
]


@pytest.mark.parametrize("m,n,k,phi", CASES)
@pytest.mark.parametrize("trans", ["NN", "NT", "TN", "TT"])
@pytest.mark.parametrize("sumMode", [0, 1])
def test_plain_gemm_bitwise(oz, oracle, m, n, k, phi, trans, sumMode):
    say(f"plain dgemm {trans} m={m} n={n} k={k} phi={phi} sumMode={sumMode}")
    got, want, H, h = run_pair(oz, oracle, trans[0], trans[1], m, n, k, phi=phi, sumMode=sumMode,
                               reproMode=1)
    assert H.nSplitA == h.nSplitA and H.nSplitB == h.nSplitB
    assert bits_equal(got, want), first_diff(got, want)


@pytest.mark.parametrize("flags", [dict(reproMode=0), dict(fastMode=1), dict(nSplitMax=2),
                                   dict(nSplitMax=3), dict(fastMode=1, sumMode=1),
                                   dict(nSplitMax=2, sumMode=1, fastMode=1)])
def test_plain_gemm_modes(oz, oracle, flags):
    f = dict(sumMode=0, reproMode=1)
    f.update(flags)
    for (m, n, k) in ((96, 80, 64), (40, 1, 50)):
        say(f"plain dgemm modes {f} m={m} n={n} k={k}")
        got, want, H, h = run_pair(oz, oracle, "N", "N", m, n, k, phi=2.0, **f)
        assert bits_equal(got, want), first_diff(got, want)


def test_truncated_degree_with_remainder_slice_is_close(oz, oracle):
    """nSplitMax > 0 with reproMode = 0 appends the unsplit remainder as a last slice
    (src/cpu/ozblas_split.cpp:276-287): its products are not exact, so their bits depend on the
    summation order inside the slice GEMM (OpenBLAS, cuBLAS and MFMA all differ) -- the reference
    itself is not reproducible there.  The result must still agree to a few ulps."""
    got, want, _, _ = run_pair(oz, oracle, "N", "N", 96, 80, 64, phi=2.0, sumMode=0, reproMode=0,
                               nSplitMax=3)
    rel = np.abs(got - want) / np.maximum(np.abs(want), 1e-300)
    assert np.nanmax(rel[want != 0]) < 1e-14


def test_plain_gemm_alpha_beta(oz, oracle):
    rng = np.random.default_rng(4)
    for sumMode in (0, 1):
        got, want, _, _ = run_pair(oz, oracle, "T", "N", 70, 90, 110, alpha=-1.25, beta=0.5,
                                   cinit=lambda s: rng.standard_normal(s), sumMode=sumMode)
        assert bits_equal(got, want), first_diff(got, want)


def test_plain_gemm_blocked_workspace(oz, oracle):
    """A workspace too small for the whole problem forces m/n blocking; NearSum bits are
    independent of it (the reference blocks too, src/cpu/ozblas_gemm.cpp:59-97)."""
    got, want, H, _ = run_pair(oz, oracle, "N", "N", 300, 280, 256, phi=1.0, sumMode=1,
                               ws=10 << 20)
    assert H.mbk < 300 or H.nbk < 280
    assert bits_equal(got, want), first_diff(got, want)


@pytest.mark.parametrize("scheme", ["digits", "residues"])
@pytest.mark.parametrize("slice_type", ["int8", "fp16", "bf16"])
@pytest.mark.parametrize("m,n,k,phi,trans", [(256, 256, 256, 1.0, "NN"), (77, 133, 199, 4.0, "TN"),
                                             (130, 129, 300, 0.5, "NT"), (200, 64, 1000, 2.0, "TT"),
                                             (1, 1, 37, 1.0, "NN"), (64, 1, 300, 1.0, "NN"),
                                             (600, 520, 700, 8.0, "NT")])
def test_fast_forms_match_plain_nearsum(oz, oracle, slice_type, scheme, m, n, k, phi, trans):
    say(f"{slice_type} {scheme} dgemm {trans} m={m} n={n} k={k} phi={phi}")
    got, want, H, _ = run_pair(oz, oracle, trans[0], trans[1], m, n, k, phi=phi, sumMode=1,
                               slice_type=slice_type, digitScheme=scheme)
    assert H.schemeUsed == (1 if scheme == "digits" else 2)
    assert bits_equal(got, want), first_diff(got, want)


@pytest.mark.parametrize("slice_type", ["int8", "fp16", "bf16"])
def test_residue_form_edges(oz, oracle, slice_type):
    """Residue form: k over one int32 chunk (two k chunks summed modulo p), the most moduli, and
    operands whose span the moduli cannot cover (the digit form takes over) -- each RN(exact)."""
    rng = np.random.default_rng(31)
    cases = []
    m, n, k = 24, 40, 70000  # two k chunks of 65536
    cases.append(("two chunks", m, n, k, rng.standard_normal(m * k), rng.standard_normal(k * n), 2))
    m, n, k = 50, 30, 300  # spans near 130 bits per operand: about 36 moduli
    A = rng.standard_normal(m * k) * np.exp2(rng.integers(-35, 35, m * k).astype(float))
    B = rng.standard_normal(k * n) * np.exp2(rng.integers(-35, 35, k * n).astype(float))
    cases.append(("wide", m, n, k, A, B, 2))
    A2 = A.copy()
    A2[::7] *= 2.0 ** 100  # spans beyond what the 49 moduli cover: digit planes
    cases.append(("beyond the moduli", m, n, k, A2, B, 1))
    # |X| attains the Cauchy-Schwarz bound ||a'|| ||b'||: constant 53-bit entries of one sign
    m, n, k = 16, 12, 4096
    v = 1.0 + 2.0 ** -52
    A3 = np.full(m * k, v)
    B3 = np.full(k * n, -v)
    B3[k * 5:k * 6] = v  # one column of the other sign
    cases.append(("bound attained", m, n, k, A3, B3, 2))
    for name, m, n, k, A, B, used in cases:
        want = oracle.exact_gemm_rn("N", "N", m, n, k, A, m, B, k)
        H = oz.Handle(1 << 30, sumMode=1, sliceType=slice_type, digitScheme="residues")
        dC = dev(np.zeros(m * n))
        say(f"residue edge {name} {slice_type} m={m} n={n} k={k}")
        H.dgemm("N", "N", m, n, k, 1.0, dev(A), m, dev(B), k, 0.0, dC, m)
        got = host(dC).reshape(n, m).T
        assert H.schemeUsed == used, (name, H.schemeUsed)
        assert bits_equal(got, want), (name, first_diff(got, want))


def test_fast_form_alpha_beta_and_exact(oz, oracle):
    rng = np.random.default_rng(9)
    m, n, k = 150, 70, 513
    got, want, H, _ = run_pair(oz, oracle, "N", "T", m, n, k, phi=1.0, alpha=0.75, beta=-2.0,
                               cinit=lambda s: rng.standard_normal(s), sumMode=1,
                               slice_type="int8")
    assert bits_equal(got, want), first_diff(got, want)
    # and against the oracle's exact accumulator directly (alpha=1, beta=0)
    got, _, _, _ = run_pair(oz, oracle, "N", "N", m, n, k, phi=4.0, sumMode=1, slice_type="int8")
    A = oracle.init_mat(m, k, m + 3, 4.0, 2, 1)
    B = oracle.init_mat(k, n, k + 1, 4.0, 2, 2)
    ex = oracle.exact_gemm_rn("N", "N", m, n, k, A, m + 3, B, k + 1)
    gotm = got.reshape(n, m + 2).T[:m, :]
    assert bits_equal(gotm, ex), first_diff(gotm, ex)


def test_fast_form_edge_values(oz, oracle):
    """Zeros, a zero row, powers of two, values at the digit boundaries, mixed exponents."""
    m, n, k = 40, 30, 96
    rng = np.random.default_rng(3)
    A = rng.standard_normal((m, k)) * np.exp2(rng.integers(-40, 40, (m, k)))
    A[3, :] = 0.0
    A[5, :7] = [1.0, -1.0, 2.0 ** 30, -(2.0 ** -30), 127.0, -128.0, 0.9921875]
    B = rng.standard_normal((k, n))
    B[:, 4] = 0.0
    B[10, :] = 2.0 ** -60
    Af, Bf = np.asfortranarray(A).T.ravel().copy(), np.asfortranarray(B).T.ravel().copy()
    want = oracle.exact_gemm_rn("N", "N", m, n, k, Af, m, Bf, k)
    for st, scheme in (("int8", "digits"), ("fp16", "digits"), ("bf16", "digits"), ("int8", "residues"),
                       ("fp16", "residues"), ("bf16", "residues"), ("fp64", "auto")):
        H = oz.Handle(1 << 28, sumMode=1, sliceType=st, digitScheme=scheme)
        dC = dev(np.zeros(m * n))
        H.dgemm("N", "N", m, n, k, 1.0, dev(Af), m, dev(Bf), k, 0.0, dC, m)
        got = host(dC).reshape(n, m).T
        assert bits_equal(got, want), (st, scheme, first_diff(got, want))


@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("sumMode", [0, 1])
def test_gemv_bitwise(oz, oracle, trans, sumMode):
    for (m, n) in ((100, 80), (7, 300), (300, 7)):
        say(f"dgemv {trans} m={m} n={n} sumMode={sumMode}")
        A = oracle.init_mat(m, n, m, 1.0, 2, 11)
        lx, ly = (n, m) if trans == "N" else (m, n)
        x = oracle.init_mat(lx, 1, lx, 1.0, 2, 12)
        y0 = np.zeros(ly)
        h = oracle.Handle(1 << 28, sumMode=sumMode)
        yw = y0.copy()
        h.gemv(trans, m, n, 1.0, A, m, x, 0.0, yw)
        H = oz.Handle(1 << 28, sumMode=sumMode)
        dy = dev(y0)
        H.dgemv(trans, m, n, 1.0, dev(A), m, dev(x), 0.0, dy)
        assert bits_equal(host(dy), yw), first_diff(host(dy), yw)


def _gemv_edge_cases():
    """(name, m, n, A, x, flags): inner lengths off the 512-lane row, rows of zeros and of one value,
    rows and x spanning many binades (up to 8 slices), the split cap reached (remainder slot)."""
    rng = np.random.default_rng(90)
    out = []
    for (m, n) in ((5, 1), (33, 7), (64, 513), (40, 8192)):
        A = rng.standard_normal(m * n)
        A[: n * 2 : 2] = 0.0                               # zeros spread over rows
        A.reshape(n, m)[:, 1] = 0.0                        # row 1 of A: a zero row of op(A) for 'N'
        x = rng.standard_normal(n)
        out.append((f"plain {m}x{n}", m, n, A, x, {}))
    m, n = 48, 1500
    A = rng.standard_normal(m * n) * np.exp2(rng.integers(-40, 40, m * n).astype(float))
    x = rng.standard_normal(n) * np.exp2(rng.integers(-10, 10, n).astype(float))
    out.append(("wide", m, n, A, x, {}))
    out.append(("wide cap 2", m, n, A, x, {"nSplitMax": 2}))
    out.append(("wide cap 3 repro 0", m, n, A, x, {"nSplitMax": 3, "reproMode": 0}))
    out.append(("wide repro 0", m, n, A, x, {"reproMode": 0}))
    A1 = np.tile(np.array([2.0 ** -3, 3.0, -1.0]), m * n // 3 + 1)[: m * n]
    out.append(("exact early", m, n, A1, np.ones(n), {}))
    return out


@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("sumMode", [0, 1])
def test_gemv_split_edge_cases_bitwise(oz, oracle, trans, sumMode):
    """The binary64 GEMV whose split of op(A) is fused into the slice products, on the edge cases
    of the split: y and the slice counts equal the oracle's (y to rounding-error bounds where the
    reproMode 0 remainder slice makes the products inexact)."""
    for name, m, n, A, x0, flags in _gemv_edge_cases():
        lx, ly = (n, m) if trans == "N" else (m, n)
        x = x0[:lx] if len(x0) >= lx else np.resize(x0, lx)
        y0 = np.linspace(-1.0, 1.0, ly)
        h = oracle.Handle(1 << 30, sumMode=sumMode, **flags)
        yw = y0.copy()
        h.gemv(trans, m, n, 0.75, A, m, x, -0.5, yw)
        H = oz.Handle(1 << 28, sumMode=sumMode, **flags)
        dy = dev(y0)
        say(f"dgemv edge {name} {trans} sumMode={sumMode}")
        H.dgemv(trans, m, n, 0.75, dev(A), m, dev(x), -0.5, dy)
        assert (H.nSplitA, H.nSplitB) == (h.nSplitA, h.nSplitB), name
        if flags.get("reproMode") == 0 and flags.get("nSplitMax", 0) > 0:
            # the unsplit remainder is the last slice: its products round, in an order of each
            # implementation's own (as test_truncated_degree_with_remainder_slice_is_close)
            Am = A.reshape(n, m).T
            Aop = Am if trans == "N" else Am.T
            # each of the nA x nB pair dots is within K u sum|A||x| of its exact value in either
            # implementation (u = 2^-53), hence the bound on the difference
            scale = 0.75 * (np.abs(Aop) @ np.abs(x)) + 0.5 * np.abs(y0)
            npairs = h.nSplitA * h.nSplitB
            bound = 2 * npairs * (Aop.shape[1] + 2) * 2.0 ** -53 * scale
            assert np.all(np.abs(host(dy) - yw) <= bound), name
        else:
            assert bits_equal(host(dy), yw), (name, first_diff(host(dy), yw))


def test_plain_gemm_k0_and_empty(oz, oracle):
    """k = 0: no slices, C = fma(alpha, +0, beta * C) as the reference computes it; m = 0 / n = 0:
    C untouched."""
    for sumMode in (0, 1):
        for (alpha, beta) in ((1.0, 0.0), (0.5, -2.0)):
            got, want, _, _ = run_pair(oz, oracle, "N", "N", 9, 11, 0, alpha=alpha, beta=beta,
                                       cinit=lambda n: np.linspace(-3, 3, n), sumMode=sumMode)
            assert bits_equal(got, want), first_diff(got, want)
    H = oz.Handle(1 << 24, sumMode=1)
    C = dev(np.full(12, 7.0))
    H.dgemm("N", "N", 0, 3, 4, 1.0, dev(np.ones(12)), 1, dev(np.ones(12)), 4, 0.0, C, 1)
    assert bits_equal(host(C), np.full(12, 7.0))


@pytest.mark.parametrize("slice_type", ["int8", "fp16", "bf16"])
@pytest.mark.parametrize("levels", [1, 2, 4, 0, -1])
@pytest.mark.parametrize("m,n,k,ta,tb", [(600, 520, 700, "N", "N"), (200, 130, 3000, "T", "N")])
def test_digit_certified_levels_bitwise(oz, oracle, slice_type, levels, m, n, k, ta, tb):
    """Digit forms with certified early termination: the first `levels` digit levels, then the
    rounding check and the remaining levels for the tiles that need them (levels 0: automatic,
    -1: every level in one pass) -- bitwise equal to the plain form in every case."""
    say(f"{slice_type} certified levels={levels} {ta}{tb} m={m} n={n} k={k}")
    want, _, _, _ = run_pair(oz, oracle, ta, tb, m, n, k, phi=1.0, sumMode=1, slice_type="fp64")
    got, _, H, _ = run_pair(oz, oracle, ta, tb, m, n, k, phi=1.0, sumMode=1, slice_type=slice_type,
                            digitLevels=levels, digitScheme="digits")
    assert H.schemeUsed == 1
    assert bits_equal(got, want), first_diff(got, want)
    if levels in (1, 2):
        assert H.fallbackTiles > 0  # the check cannot settle most elements from so few levels


@pytest.mark.parametrize("slice_type", ["int8", "fp16", "bf16"])
def test_digit_certified_cancellation(oz, oracle, slice_type):
    """Exact cancellation (every C entry is 0 or tiny): the check cannot settle the sign from the
    leading levels, so every tile takes the second pass; the result is still RN(exact).  The
    residue form is checked on the same product."""
    m, n, k = 300, 260, 512
    A = oracle.init_mat(m, k, m, 1.0, 2, 41).reshape(k, m).T.copy()
    B = oracle.init_mat(k, n, k, 1.0, 2, 42).reshape(n, k).T.copy()
    A[:, k // 2:] = A[:, : k // 2]
    B[k // 2:, :] = -B[: k // 2, :]
    B[k - 1, :] += np.ldexp(1.0, -70)  # a tiny exact remainder in part of the products
    Af = np.asfortranarray(A).T.ravel().copy()
    Bf = np.asfortranarray(B).T.ravel().copy()
    ex = np.ascontiguousarray(np.asarray(oracle.exact_gemm_rn("N", "N", m, n, k, Af, m, Bf, k)).T).ravel()
    for levels, scheme in ((0, "digits"), (6, "digits"), (0, "residues")):
        H = oz.Handle(1 << 30, sumMode=1, sliceType=slice_type, digitLevels=levels, digitScheme=scheme)
        dC = dev(np.zeros(m * n))
        H.dgemm("N", "N", m, n, k, 1.0, dev(Af), m, dev(Bf), k, 0.0, dC, m)
        assert bits_equal(host(dC), ex), (levels, scheme, first_diff(host(dC), ex))
        if levels == 6:
            assert H.fallbackTiles == ((m + 127) // 128) * ((n + 127) // 128)  # every tile
