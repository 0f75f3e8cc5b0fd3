"""GPU parity of the mixed routines with binary32 data and binary64-held slices (ozblasSD{gemm,gemv,
dot,nrm2,csrmv,csrmvSplitA,cg}) through the C ABI, bitwise against the oracle's restatement of the
reference's <float,double,double> / <float,double> instantiations (oracle/ozoracle_sd.c).

Inner dimensions are above 128 for the dense routines: below that getRho<float,double> is at most 1
and the reference's binary64 slice GEMM may round, in its BLAS's order (CSR rows are exempt: their
fma chain is the reference's own loop, and the device runs the same chain)."""
import numpy as np
import pytest

from gpu_util import dev, host, say, torch_cuda

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def oz():
    torch_cuda()
    import ozblas_amd
    return ozblas_amd


def f32_equal(a, b):
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def f32_diff(a, b):
    a = np.ascontiguousarray(a, dtype=np.float32).ravel()
    b = np.ascontiguousarray(b, dtype=np.float32).ravel()
    d = np.nonzero(a.view(np.uint32) != b.view(np.uint32))[0]
    if len(d) == 0:
        return None
    i = d[0]
    return dict(index=int(i), count=int(len(d)), got=float(a[i]).hex(), want=float(b[i]).hex())


def _mat(oracle, rows, cols, phi, seed):
    return oracle.s_init_mat(rows, cols, rows, phi, 2, seed)


GEMM_CASES = [
    (64, 48, 200, "n", "n"),
    (257, 130, 300, "t", "n"),
    (130, 257, 129, "n", "t"),
    (33, 65, 1000, "t", "t"),
    (200, 1, 300, "n", "n"),    # the GEMV branch of the reference GEMM
    (1, 77, 640, "t", "n"),
    (1, 1, 500, "t", "n"),      # the DOT branch
]


@pytest.mark.parametrize("case", GEMM_CASES, ids=lambda c: "x".join(map(str, c[:3])) + c[3] + c[4])
@pytest.mark.parametrize("phi", [0.5, 4.0])
def test_sdgemm_bitwise(oz, oracle, case, phi):
    m, n, k, ta, tb = case
    say(f"sdgemm {m}x{n}x{k} {ta}{tb} phi={phi}")
    ra, ca = (m, k) if ta == "n" else (k, m)
    rb, cb = (k, n) if tb == "n" else (n, k)
    A = _mat(oracle, ra, ca, phi, 1)
    B = _mat(oracle, rb, cb, phi, 2)
    C0 = _mat(oracle, m, n, 1.0, 3)
    want = C0.copy()
    h = oracle.Handle(1 << 30, sumMode=0)
    h.sdgemm(ta, tb, m, n, k, np.float32(1.25), A, ra, B, rb, np.float32(-0.5), want, m)
    H = oz.Handle(1 << 28, sumMode=0)
    dC = dev(C0)
    H.sdgemm(ta, tb, m, n, k, 1.25, dev(A), ra, dev(B), rb, -0.5, dC, m)
    assert f32_equal(host(dC), want), f32_diff(host(dC), want)
    assert (H.nSplitA, H.nSplitB) == (h.nSplitA, h.nSplitB)


@pytest.mark.parametrize("flags", [dict(reproMode=0), dict(nSplitMax=3), dict(splitEpsMode=9),
                                   dict(fastMode=1), dict(nSplitMax=2, reproMode=0)],
                         ids=lambda f: ",".join(f"{k}={v}" for k, v in f.items()))
def test_sdgemm_options_bitwise(oz, oracle, flags):
    m, n, k = 96, 80, 300
    say(f"sdgemm options {flags}")
    A = _mat(oracle, m, k, 2.0, 4)
    B = _mat(oracle, k, n, 2.0, 5)
    if flags.get("splitEpsMode") == 9:
        # the overflow mode's rho (ceil(log2(k) / 4) - 2) leaves the slice products inexact for
        # general data; integers of 7 bits keep them exact
        rng = np.random.default_rng(9)
        A = rng.integers(-64, 65, m * k).astype(np.float32)
        B = rng.integers(-64, 65, k * n).astype(np.float32)
    want = np.zeros(m * n, dtype=np.float32)
    h = oracle.Handle(1 << 30, sumMode=0, **flags)
    h.sdgemm("n", "n", m, n, k, np.float32(1.0), A, m, B, k, np.float32(0.0), want, m)
    H = oz.Handle(1 << 28, sumMode=0, **flags)
    dC = dev(np.zeros(m * n, dtype=np.float32))
    H.sdgemm("n", "n", m, n, k, 1.0, dev(A), m, dev(B), k, 0.0, dC, m)
    assert f32_equal(host(dC), want), f32_diff(host(dC), want)
    assert (H.nSplitA, H.nSplitB, H.nSplitC) == (h.nSplitA, h.nSplitB, h.nSplitC)


def test_sdgemm_blocked_workspace_bitwise(oz, oracle):
    m, n, k = 300, 260, 200
    A = _mat(oracle, m, k, 1.0, 7)
    B = _mat(oracle, k, n, 1.0, 8)
    want = np.zeros(m * n, dtype=np.float32)
    oracle.Handle(1 << 30, sumMode=0).sdgemm("n", "n", m, n, k, np.float32(1.0), A, m, B, k,
                                             np.float32(0.0), want, m)
    H = oz.Handle(3 << 20, sumMode=0)  # binary32 operands: half the binary64 test's need
    dC = dev(np.zeros(m * n, dtype=np.float32))
    say("sdgemm blocked")
    H.sdgemm("n", "n", m, n, k, 1.0, dev(A), m, dev(B), k, 0.0, dC, m)
    assert (H.mbk, H.nbk) != (m, n)
    assert f32_equal(host(dC), want), f32_diff(host(dC), want)


@pytest.mark.parametrize("trans", ["n", "t"])
def test_sdgemv_bitwise(oz, oracle, trans):
    m, n = 301, 177
    A = _mat(oracle, m, n, 2.0, 9)
    lx, ly = (n, m) if trans == "n" else (m, n)
    x = _mat(oracle, lx, 1, 2.0, 10)
    y0 = _mat(oracle, ly, 1, 1.0, 11)
    want = y0.copy()
    oracle.Handle(1 << 30, sumMode=0).sdgemv(trans, m, n, np.float32(0.75), A, m, x, np.float32(2.0), want)
    H = oz.Handle(1 << 28, sumMode=0)
    dy = dev(y0)
    say(f"sdgemv {trans}")
    H.sdgemv(trans, m, n, 0.75, dev(A), m, dev(x), 2.0, dy)
    assert f32_equal(host(dy), want), f32_diff(host(dy), want)


@pytest.mark.parametrize("n", [1, 129, 1000, 65536, 100003])
@pytest.mark.parametrize("reproMode", [0, 1])
def test_sddot_sdnrm2_bitwise(oz, oracle, n, reproMode):
    say(f"sddot n={n} reproMode={reproMode}")
    x = _mat(oracle, n, 1, 2.0, 21)
    y = _mat(oracle, n, 1, 2.0, 22)
    H = oz.Handle(1 << 28, sumMode=0, reproMode=reproMode)
    h = oracle.Handle(1 << 30, sumMode=0, reproMode=reproMode)
    dx = dev(x)
    got = np.float32(H.sddot(n, dx, dev(y)))
    want = h.sddot(n, x, y)
    assert got.tobytes() == want.tobytes(), (float(got).hex(), float(want).hex())
    assert (H.nSplitA, H.nSplitB) == (h.nSplitA, h.nSplitB)
    got2 = np.float32(H.sdnrm2(n, dx))
    want2 = oracle.Handle(1 << 30, sumMode=0, reproMode=reproMode).sdnrm2(n, x)
    assert got2.tobytes() == want2.tobytes()


def test_sddot_edge_cases_bitwise(oz, oracle):
    """Wide exponent ranges, zeros, and data whose slice products fall below 2^-126: there the
    reference rounds each scaled product to binary32 first and then onto the subnormal grid when
    it scales it back, which the device reproduces on the unscaled product (f32_term)."""
    rng = np.random.default_rng(51)
    vecs = [rng.uniform(-1, 1, 5000), rng.standard_normal(3000) * np.exp(4 * rng.standard_normal(3000)),
            np.ldexp(rng.integers(-2**20, 2**20, 2000).astype(float), -30), np.zeros(300),
            rng.standard_normal(4000) * 2.0 ** -68, rng.standard_normal(700) * 2.0 ** -75,
            np.array([1.5 * 2.0 ** 120, -3.0, 1.0, 2.0 ** 100] + [0.5] * 200)]
    for i, x in enumerate(vecs):
        x = x.astype(np.float32)
        n = len(x)
        y = np.random.default_rng(n).uniform(-1, 1, n).astype(np.float32)
        if i in (4, 5):
            y = (y * np.float32(2.0 ** -64)).astype(np.float32)
        h = oracle.Handle(1 << 30, sumMode=0)
        want = h.sddot(n, x, y)
        H = oz.Handle(1 << 28, sumMode=0)
        say(f"sddot edge {i} n={n}")
        got = np.float32(H.sddot(n, dev(x), dev(y)))
        same = got.tobytes() == want.tobytes()
        assert same or (np.isnan(got) and np.isnan(want)), (i, n, float(got).hex(), float(want).hex())


def _laplace_f(oracle, g):
    rowptr, colind, val = oracle.laplace3d(g, 1e-3)
    return rowptr, colind, val.astype(np.float32), g ** 3, len(val)


@pytest.mark.parametrize("presplit", [False, True])
@pytest.mark.parametrize("kind", ["laplace", "wide-x", "tiny"])
def test_sdcsrmv_bitwise(oz, oracle, presplit, kind):
    rowptr, colind, val, n, nnz = _laplace_f(oracle, 12)
    rng = np.random.default_rng(3)
    e = {"laplace": 20, "wide-x": 60, "tiny": 4}[kind]
    x = (rng.uniform(-1, 1, n) * np.exp2(rng.integers(-e, e, n))).astype(np.float32)
    if kind == "tiny":  # products below 2^-126
        x = (x * np.float32(2.0 ** -128)).astype(np.float32)
    y0 = rng.standard_normal(n).astype(np.float32)
    say(f"sdcsrmv {kind} presplit={presplit}")
    h = oracle.Handle(1 << 30, sumMode=0)
    H = oz.Handle(1 << 28, sumMode=0)
    dval, dcol, drow, dx = dev(val), dev(colind), dev(rowptr), dev(x)
    yw = y0.copy()
    dy = dev(y0)
    if presplit:
        p = h.sdcsrmv_splitA(n, n, nnz, val, rowptr)
        P = H.sdcsrmv_splitA(n, n, nnz, dval, drow)
        assert H.nSplitA_ == h.nSplitA_ and H.splitShift == h.splitShift
        h.sdcsrmv(n, n, nnz, np.float32(1.5), p, colind, rowptr, x, np.float32(-0.25), yw)
        H.sdcsrmv(n, n, nnz, 1.5, P, dcol, drow, dx, -0.25, dy)
    else:
        h.sdcsrmv(n, n, nnz, np.float32(1.5), val, colind, rowptr, x, np.float32(-0.25), yw)
        H.sdcsrmv(n, n, nnz, 1.5, dval, dcol, drow, dx, -0.25, dy)
    assert f32_equal(host(dy), yw), f32_diff(host(dy), yw)
    assert H.nSplitB == h.nSplitB


@pytest.mark.parametrize("path", ["device", "host", "overflow"])
def test_sdcg_history_bitwise(oz, oracle, path):
    g, iters = 10, 40
    rowptr, colind, val, n, nnz = _laplace_f(oracle, g)
    oracle.srand48(3)
    b = np.array([2.0 * oracle.drand48() - 1.0 for _ in range(n)], dtype=np.float32)
    h = oracle.Handle(1 << 30, sumMode=0)
    xw = np.zeros(n, dtype=np.float32)
    want = h.sdcg(n, nnz, val, colind, rowptr, b, xw, iters, np.float32(0))
    bound = {"device": 0, "host": -1, "overflow": 1}[path]
    H = oz.Handle(1 << 28, sumMode=0, cgDeviceSlices=bound)
    dx = dev(np.zeros(n, dtype=np.float32))
    say(f"sdcg n={n} iters={iters} path={path}")
    got = H.sdcg(n, nnz, dev(val), dev(colind), dev(rowptr), dev(b), dx, iters, 0.0)
    assert H.cgDevicePath == {"device": 1, "host": 0, "overflow": 2}[path]
    assert len(got) == iters + 1
    assert f32_equal(got.astype(np.float32), want), f32_diff(got, want)
    assert f32_equal(host(dx), xw), f32_diff(host(dx), xw)


def test_sd_nearsum_and_dot2_not_supported(oz):
    """The reference's mixed NearSum exits; Dot2 of the <float,double> instantiation is not
    provided."""
    x = dev(np.ones(200, dtype=np.float32))
    for flags in (dict(sumMode=1), dict(sumMode=0, splitEpsMode=2)):
        H = oz.Handle(1 << 24, **flags)
        with pytest.raises(oz.OzblasError) as e:
            H.sddot(200, x, x)
        assert e.value.status == oz.NOT_SUPPORTED
        C = dev(np.zeros(16, dtype=np.float32))
        with pytest.raises(oz.OzblasError) as e:
            H.sdgemm("n", "n", 4, 4, 4, 1.0, x, 4, x, 4, 0.0, C, 4)
        assert e.value.status == oz.NOT_SUPPORTED
