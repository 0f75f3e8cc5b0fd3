"""GPU parity of DOT, NRM2, AXPY, CSRMV (with and without the pre-split of ozblasDcsrmvSplitA) and
the CG solver through the C ABI, bitwise against the oracle restatement of the reference."""
import math

import numpy as np
import pytest

from gpu_util import bits_equal, dev, first_diff, host, say, torch_cuda

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def oz():
    torch_cuda()
    import ozblas_amd
    return ozblas_amd


@pytest.mark.parametrize("n", [1, 2, 7, 1000, 65536, 100003])
@pytest.mark.parametrize("sumMode", [0, 1])
@pytest.mark.parametrize("reproMode", [0, 1])
def test_dot_nrm2_bitwise(oz, oracle, n, sumMode, reproMode):
    say(f"ddot n={n} sumMode={sumMode} reproMode={reproMode}")
    x = oracle.init_mat(n, 1, n, 2.0, 2, 21)
    y = oracle.init_mat(n, 1, n, 2.0, 2, 22)
    h = oracle.Handle(1 << 30, sumMode=sumMode, reproMode=reproMode)
    H = oz.Handle(1 << 28, sumMode=sumMode, reproMode=reproMode)
    dx, dy = dev(x), dev(y)
    got = H.ddot(n, dx, dy)
    want = oracle.Handle(1 << 30, sumMode=sumMode, reproMode=reproMode).dot(n, x, y)
    assert np.float64(got).tobytes() == np.float64(want).tobytes(), (got.hex(), want.hex())
    assert H.dnrm2(n, dx) == h.nrm2(n, x)


def _split_edge_vectors():
    rng = np.random.default_rng(31)
    yield "uniform", rng.uniform(-1, 1, 5000)
    yield "wide", rng.standard_normal(3000) * np.exp(8 * rng.standard_normal(3000))
    yield "few-bits", np.ldexp(rng.integers(-2**20, 2**20, 2000).astype(float), -30)
    v = np.ldexp(1.0, rng.integers(-40, 40, 1500)).astype(float)
    v[::3] *= -1
    yield "powers-of-two", v
    w = rng.uniform(-1, 1, 1200)
    w[::7] = np.nextafter(np.ldexp(1.0, -3), 1.0)
    w[1::11] = 0.0
    yield "just-above-2^e", w
    yield "ties", np.ldexp(rng.integers(0, 2**52, 700).astype(float) + 0.5, -40)
    yield "zeros", np.zeros(300)
    yield "two-ranges", np.concatenate([rng.uniform(-1, 1, 500), rng.uniform(-1, 1, 500) * 1e-200])
    # outside the residue model: the one-workgroup reference loop splits these
    yield "huge", np.array([1.5 * 2.0 ** 1022, -3.0, 1.0, 2.0 ** 1000])
    yield "subnormal", np.array([5e-324, -1e-310, 3e-320, 0.0])


@pytest.mark.parametrize("sumMode", [0, 1])
def test_dot_split_edge_cases_bitwise(oz, oracle, sumMode):
    """DOT through the vector split on vectors with ties, powers of two, exact-early splits, zeros,
    and inputs handled by the fallback loop: result and slice counts equal the oracle's."""
    for name, x in _split_edge_vectors():
        n = len(x)
        y = np.random.default_rng(n).uniform(-1, 1, n)
        for a, b in ((x, x), (x, y)):
            h = oracle.Handle(1 << 30, sumMode=sumMode)
            want = h.dot(n, a, b)
            H = oz.Handle(1 << 28, sumMode=sumMode)
            got = H.ddot(n, dev(a), dev(b))
            say(f"ddot edge {name} n={n} sumMode={sumMode}")
            same = np.float64(got).tobytes() == np.float64(want).tobytes()
            assert same or (np.isnan(got) and np.isnan(want)), (name, got, want)
            assert (H.nSplitA, H.nSplitB) == (h.nSplitA, h.nSplitB), name


def test_dot_sequence_after_inputs_outside_model_same_handle(oz, oracle):
    """One handle splits vectors outside the residue model (an element near 2^1023, alone or among
    thousands, so that one workgroup of several is outside it) between ordinary ones: every DOT and
    its slice counts equal the oracle's, so no workgroup table of an earlier split leaks into a later
    one."""
    rng = np.random.default_rng(41)
    big = rng.uniform(-1, 1, 5000)
    big[2500] = 1.25 * 2.0 ** 1022
    wide = rng.standard_normal(4000) * np.exp2(rng.integers(-200, 200, 4000).astype(float))
    seq = [np.array([1.5 * 2.0 ** 1022, -3.0, 1.0, 2.0 ** 1000]), rng.uniform(-1, 1, 3000), big, wide,
           big, rng.uniform(-1, 1, 70000)]
    for sumMode in (0, 1):
        H = oz.Handle(1 << 28, sumMode=sumMode)
        for q, x in enumerate(seq):
            n = len(x)
            y = np.random.default_rng(n + q).uniform(-1, 1, n)
            h = oracle.Handle(1 << 30, sumMode=sumMode)
            want = h.dot(n, x, y)
            say(f"ddot sequence item {q} n={n} sumMode={sumMode}")
            got = H.ddot(n, dev(x), dev(y))
            same = np.float64(got).tobytes() == np.float64(want).tobytes()
            assert same or (np.isnan(got) and np.isnan(want)), (q, got, want)  # the split overflows
            assert (H.nSplitA, H.nSplitB) == (h.nSplitA, h.nSplitB), q


@pytest.mark.parametrize("nSplitMax", [0, 3])
def test_dot_split_range_over_512_binades_bitwise(oz, oracle, nSplitMax):
    """A vector whose exponents span about 600 binades: the split's residue table then holds more
    entries than the level walk has threads, so its prefix maximum spans several entries per
    thread (written by other threads)."""
    rng = np.random.default_rng(77)
    n = 600
    x = rng.standard_normal(n) * np.exp2(rng.integers(-300, 300, n).astype(float))
    y = rng.uniform(-1, 1, n)
    for a, b in ((x, x), (x, y)):
        h = oracle.Handle(1 << 30, sumMode=1, nSplitMax=nSplitMax)
        want = h.dot(n, a, b)
        H = oz.Handle(1 << 28, sumMode=1, nSplitMax=nSplitMax)
        say(f"ddot range 600 binades nSplitMax={nSplitMax}")
        got = H.ddot(n, dev(a), dev(b))
        assert np.float64(got).tobytes() == np.float64(want).tobytes(), (got, want)
        assert (H.nSplitA, H.nSplitB) == (h.nSplitA, h.nSplitB)


def test_dot_is_correctly_rounded_under_nearsum(oz, oracle):
    rng = np.random.default_rng(8)
    for n in (3, 999, 40000):
        x = rng.standard_normal(n) * np.exp2(rng.integers(-30, 30, n))
        y = rng.standard_normal(n)
        H = oz.Handle(1 << 28, sumMode=1)
        assert H.ddot(n, dev(x), dev(y)) == oracle.exact_dot_rn(x, y)


def test_axpy(oz, oracle):
    rng = np.random.default_rng(1)
    x, y = rng.standard_normal(1001), rng.standard_normal(1001)
    for reproMode in (0, 1):
        H = oz.Handle(1 << 24, reproMode=reproMode)
        dy = dev(y)
        H.daxpy(1001, -0.3, dev(x), dy)
        want = oracle.Handle(1 << 24, reproMode=reproMode).axpy(1001, -0.3, x, y.copy())
        assert bits_equal(host(dy), want)


def _laplace(oracle, g):
    rowptr, colind, val = oracle.laplace3d(g, 1e-3)
    return rowptr, colind, val, g ** 3, len(val)


@pytest.mark.parametrize("sumMode", [0, 1])
@pytest.mark.parametrize("presplit", [False, True])
def test_csrmv_bitwise(oz, oracle, sumMode, presplit):
    rowptr, colind, val, n, nnz = _laplace(oracle, 12)
    rng = np.random.default_rng(2)
    x = rng.uniform(-1, 1, n) * np.exp2(rng.integers(-20, 20, n))
    y0 = rng.standard_normal(n)
    say(f"dcsrmv n={n} sumMode={sumMode} presplit={presplit}")
    h = oracle.Handle(1 << 30, sumMode=sumMode)
    H = oz.Handle(1 << 28, sumMode=sumMode)
    dval, dcol, drow, dx = dev(val), dev(colind), dev(rowptr), dev(x)
    yw = y0.copy()
    dy = dev(y0)
    if presplit:
        p = h.csrmv_splitA(n, n, nnz, val, rowptr)
        P = H.dcsrmv_splitA(n, n, nnz, dval, drow)
        assert (H.splitShift, H.nSplitA_) == (h.splitShift, h.nSplitA_)
        h.csrmv(n, n, nnz, 0.5, p, colind, rowptr, x, -1.5, yw)
        H.dcsrmv(n, n, nnz, 0.5, P, dcol, drow, dx, -1.5, dy)
    else:
        h.csrmv(n, n, nnz, 0.5, val, colind, rowptr, x, -1.5, yw)
        H.dcsrmv(n, n, nnz, 0.5, dval, dcol, drow, dx, -1.5, dy)
    assert bits_equal(host(dy), yw), first_diff(host(dy), yw)


@pytest.mark.parametrize("sumMode", [0, 1])
def test_cg_residual_history_bitwise(oz, oracle, sumMode):
    g, iters = 10, 40
    rowptr, colind, val, n, nnz = _laplace(oracle, g)
    oracle.srand48(3)
    b = np.array([2.0 * oracle.drand48() - 1.0 for _ in range(n)])
    say(f"dcg n={n} iters={iters} sumMode={sumMode}")
    h = oracle.Handle(1 << 30, sumMode=sumMode)
    xw = np.zeros(n)
    want = h.cg(n, nnz, val, colind, rowptr, b, xw, iters, 0.0)
    H = oz.Handle(1 << 28, sumMode=sumMode)
    dx = dev(np.zeros(n))
    got = H.dcg(n, nnz, dev(val), dev(colind), dev(rowptr), dev(b), dx, iters, 0.0)
    assert len(got) == iters + 1
    assert bits_equal(got, want), first_diff(got, want)
    assert bits_equal(host(dx), xw), first_diff(host(dx), xw)


@pytest.mark.parametrize("sumMode", [0, 1])
@pytest.mark.parametrize("path", ["device", "host", "overflow"])
def test_cg_paths_bitwise(oz, oracle, sumMode, path):
    """The device-resident iterations, the synchronous ones, and the fallback from the device path
    (forced by a slice bound of 1, so the first CSR product gives up) give the same bits."""
    g, iters = 12, 60
    rowptr, colind, val, n, nnz = _laplace(oracle, g)
    oracle.srand48(5)
    b = np.array([2.0 * oracle.drand48() - 1.0 for _ in range(n)])
    say(f"dcg paths n={n} iters={iters} sumMode={sumMode} path={path}")
    h = oracle.Handle(1 << 30, sumMode=sumMode)
    xw = np.zeros(n)
    want = h.cg(n, nnz, val, colind, rowptr, b, xw, iters, 0.0)
    bound = {"device": 0, "host": -1, "overflow": 1}[path]
    H = oz.Handle(1 << 28, sumMode=sumMode, cgDeviceSlices=bound)
    dx = dev(np.zeros(n))
    got = H.dcg(n, nnz, dev(val), dev(colind), dev(rowptr), dev(b), dx, iters, 0.0)
    assert H.cgDevicePath == {"device": 1, "host": 0, "overflow": 2}[path]
    assert len(got) == iters + 1
    assert bits_equal(got, want), first_diff(got, want)
    assert bits_equal(host(dx), xw), first_diff(host(dx), xw)


@pytest.mark.parametrize("path", ["device", "host"])
def test_cg_tolerance_stop_bitwise(oz, oracle, path):
    """A tolerance reached mid-way stops both paths after the same iteration (the p update of that
    iteration still runs, as in the reference), with the same history and x."""
    g, iters = 10, 80
    rowptr, colind, val, n, nnz = _laplace(oracle, g)
    oracle.srand48(7)
    b = np.array([2.0 * oracle.drand48() - 1.0 for _ in range(n)])
    full = oracle.Handle(1 << 30, sumMode=1).cg(n, nnz, val, colind, rowptr, b, np.zeros(n), iters, 0.0)
    tol = float(full[25]) * (1 + 2.0 ** -40)
    h = oracle.Handle(1 << 30, sumMode=1)
    xw = np.zeros(n)
    want = h.cg(n, nnz, val, colind, rowptr, b, xw, iters, tol)
    assert 1 < len(want) - 1 < iters
    H = oz.Handle(1 << 28, sumMode=1, cgDeviceSlices=0 if path == "device" else -1)
    dx = dev(np.zeros(n))
    got = H.dcg(n, nnz, dev(val), dev(colind), dev(rowptr), dev(b), dx, iters, tol)
    assert H.cgDevicePath == (1 if path == "device" else 0)
    assert len(got) == len(want)
    assert bits_equal(got, want), first_diff(got, want)
    assert bits_equal(host(dx), xw), first_diff(host(dx), xw)


# ---- Dot2 mode (splitEpsMode = 2): wide slices, double-double slice products ------------------
@pytest.mark.parametrize("n", [1, 7, 1000, 65536, 100003])
@pytest.mark.parametrize("sumMode", [0, 1])
def test_dot2_dot_nrm2_bitwise(oz, oracle, n, sumMode):
    """The DOT slice dots are the reference's OpenMP Dot2 partial sums for 1024 threads combined
    in thread order (oracle dot2Chunks = 1024, the GPU's fixed chunking)."""
    say(f"ddot Dot2 n={n} sumMode={sumMode}")
    x = oracle.init_mat(n, 1, n, 2.0, 2, 31)
    y = oracle.init_mat(n, 1, n, 2.0, 2, 32)
    h = oracle.Handle(1 << 30, sumMode=sumMode, splitEpsMode=2)
    assert h.dot2Chunks == 1024
    H = oz.Handle(1 << 28, sumMode=sumMode, splitEpsMode=2)
    dx, dy = dev(x), dev(y)
    got = H.ddot(n, dx, dy)
    want = h.dot(n, x, y)
    assert np.float64(got).tobytes() == np.float64(want).tobytes(), (got.hex(), want.hex())
    assert (H.nSplitA, H.nSplitB) == (h.nSplitA, h.nSplitB)
    assert H.dnrm2(n, dx) == oracle.Handle(1 << 30, sumMode=sumMode, splitEpsMode=2).nrm2(n, x)


@pytest.mark.parametrize("sumMode", [0, 1])
@pytest.mark.parametrize("presplit", [False, True])
def test_dot2_csrmv_bitwise(oz, oracle, sumMode, presplit):
    rowptr, colind, val, n, nnz = _laplace(oracle, 12)
    rng = np.random.default_rng(4)
    x = rng.uniform(-1, 1, n) * np.exp2(rng.integers(-20, 20, n))
    y0 = rng.standard_normal(n)
    say(f"dcsrmv Dot2 n={n} sumMode={sumMode} presplit={presplit}")
    h = oracle.Handle(1 << 30, sumMode=sumMode, splitEpsMode=2)
    H = oz.Handle(1 << 28, sumMode=sumMode, splitEpsMode=2)
    dval, dcol, drow, dx = dev(val), dev(colind), dev(rowptr), dev(x)
    yw = y0.copy()
    dy = dev(y0)
    if presplit:
        p = h.csrmv_splitA(n, n, nnz, val, rowptr)
        P = H.dcsrmv_splitA(n, n, nnz, dval, drow)
        assert (H.splitShift, H.nSplitA_) == (h.splitShift, h.nSplitA_)
        h.csrmv(n, n, nnz, 0.5, p, colind, rowptr, x, -1.5, yw)
        H.dcsrmv(n, n, nnz, 0.5, P, dcol, drow, dx, -1.5, dy)
    else:
        h.csrmv(n, n, nnz, 0.5, val, colind, rowptr, x, -1.5, yw)
        H.dcsrmv(n, n, nnz, 0.5, dval, dcol, drow, dx, -1.5, dy)
    assert bits_equal(host(dy), yw), first_diff(host(dy), yw)


@pytest.mark.parametrize("sumMode", [0, 1])
def test_dot2_cg_residual_history_bitwise(oz, oracle, sumMode):
    g, iters = 10, 40
    rowptr, colind, val, n, nnz = _laplace(oracle, g)
    oracle.srand48(3)
    b = np.array([2.0 * oracle.drand48() - 1.0 for _ in range(n)])
    say(f"dcg Dot2 n={n} iters={iters} sumMode={sumMode}")
    h = oracle.Handle(1 << 30, sumMode=sumMode, splitEpsMode=2)
    xw = np.zeros(n)
    want = h.cg(n, nnz, val, colind, rowptr, b, xw, iters, 0.0)
    H = oz.Handle(1 << 28, sumMode=sumMode, splitEpsMode=2)
    dx = dev(np.zeros(n))
    got = H.dcg(n, nnz, dev(val), dev(colind), dev(rowptr), dev(b), dx, iters, 0.0)
    assert bits_equal(got, want), first_diff(got, want)
    assert bits_equal(host(dx), xw), first_diff(host(dx), xw)


def test_dot2_gemm_not_supported(oz):
    """GEMM / GEMV with Dot2 slices would need inexact slice products (the reference's result then
    depends on the BLAS's summation order): refused with NOT_SUPPORTED."""
    H = oz.Handle(1 << 24, sumMode=1, splitEpsMode=2)
    A, B, C = dev(np.ones(16)), dev(np.ones(16)), dev(np.zeros(16))
    with pytest.raises(oz.OzblasError):
        H.dgemm("N", "N", 4, 4, 4, 1.0, A, 4, B, 4, 0.0, C, 4)
