"""GPU parity of the mixed routines (binary64 data split into binary32 slices: ozblasDS{gemm,gemv,
dot,nrm2,csrmv,csrmvSplitA,cg}) through the C ABI, bitwise against the oracle's restatement of the
reference's <double,float,float> instantiations (oracle/ozoracle_ds.c), which reproduces the
reference's 90 DSGEMM/DSGEMV/DSDOT accuracy logs."""
import numpy as np
import pytest

from gpu_util import bits_equal, dev, first_diff, host, say, torch_cuda

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def oz():
    torch_cuda()
    import ozblas_amd
    return ozblas_amd


GEMM_CASES = [
    (64, 48, 100, "n", "n"),
    (257, 130, 300, "t", "n"),
    (130, 257, 129, "n", "t"),
    (33, 65, 1000, "t", "t"),
    (200, 1, 300, "n", "n"),
    (1, 1, 500, "t", "n"),
]


@pytest.mark.parametrize("case", GEMM_CASES, ids=lambda c: "x".join(map(str, c[:3])) + c[3] + c[4])
@pytest.mark.parametrize("phi", [0.5, 4.0])
def test_dsgemm_bitwise(oz, oracle, case, phi):
    m, n, k, ta, tb = case
    say(f"dsgemm {m}x{n}x{k} {ta}{tb} phi={phi}")
    ra, ca = (m, k) if ta == "n" else (k, m)
    rb, cb = (k, n) if tb == "n" else (n, k)
    A = oracle.init_mat(ra, ca, ra, phi, 2, 1)
    B = oracle.init_mat(rb, cb, rb, phi, 2, 2)
    C0 = oracle.init_mat(m, n, m, 1.0, 2, 3)
    want = C0.copy()
    h = oracle.Handle(1 << 30, sumMode=0)
    h.dsgemm(ta, tb, m, n, k, 1.25, A, ra, B, rb, -0.5, want, m)
    H = oz.Handle(1 << 28, sumMode=0)
    dC = dev(C0)
    H.dsgemm(ta, tb, m, n, k, 1.25, dev(A), ra, dev(B), rb, -0.5, dC, m)
    assert bits_equal(host(dC), want), first_diff(host(dC), want)
    assert (H.nSplitA, H.nSplitB) == (h.nSplitA, h.nSplitB)


@pytest.mark.parametrize("flags", [dict(reproMode=0), dict(nSplitMax=3), dict(splitEpsMode=9),
                                   dict(fastMode=1), dict(nSplitMax=2, reproMode=0)],
                         ids=lambda f: ",".join(f"{k}={v}" for k, v in f.items()))
def test_dsgemm_options_bitwise(oz, oracle, flags):
    m, n, k = 96, 80, 257
    say(f"dsgemm options {flags}")
    A = oracle.init_mat(m, k, m, 2.0, 2, 4)
    B = oracle.init_mat(k, n, k, 2.0, 2, 5)
    want = np.zeros(m * n)
    h = oracle.Handle(1 << 30, sumMode=0, **flags)
    h.dsgemm("n", "n", m, n, k, 1.0, A, m, B, k, 0.0, want, m)
    H = oz.Handle(1 << 28, sumMode=0, **flags)
    dC = dev(np.zeros(m * n))
    H.dsgemm("n", "n", m, n, k, 1.0, dev(A), m, dev(B), k, 0.0, dC, m)
    assert bits_equal(host(dC), want), first_diff(host(dC), want)
    assert (H.nSplitA, H.nSplitB, H.nSplitC) == (h.nSplitA, h.nSplitB, h.nSplitC)


def test_dsgemm_blocked_workspace_bitwise(oz, oracle):
    m, n, k = 300, 260, 200
    A = oracle.init_mat(m, k, m, 1.0, 2, 7)
    B = oracle.init_mat(k, n, k, 1.0, 2, 8)
    want = np.zeros(m * n)
    oracle.Handle(1 << 30, sumMode=0).dsgemm("n", "n", m, n, k, 1.0, A, m, B, k, 0.0, want, m)
    H = oz.Handle(6 << 20, sumMode=0)
    dC = dev(np.zeros(m * n))
    say("dsgemm blocked")
    H.dsgemm("n", "n", m, n, k, 1.0, dev(A), m, dev(B), k, 0.0, dC, m)
    assert (H.mbk, H.nbk) != (m, n)
    assert bits_equal(host(dC), want), first_diff(host(dC), want)


@pytest.mark.parametrize("trans", ["n", "t"])
def test_dsgemv_bitwise(oz, oracle, trans):
    m, n = 301, 177
    A = oracle.init_mat(m, n, m, 2.0, 2, 9)
    lx, ly = (n, m) if trans == "n" else (m, n)
    x = oracle.init_mat(lx, 1, lx, 2.0, 2, 10)
    y0 = oracle.init_mat(ly, 1, ly, 1.0, 2, 11)
    want = y0.copy()
    oracle.Handle(1 << 30, sumMode=0).dsgemv(trans, m, n, 0.75, A, m, x, 2.0, want)
    H = oz.Handle(1 << 28, sumMode=0)
    dy = dev(y0)
    say(f"dsgemv {trans}")
    H.dsgemv(trans, m, n, 0.75, dev(A), m, dev(x), 2.0, dy)
    assert bits_equal(host(dy), want), first_diff(host(dy), want)


@pytest.mark.parametrize("n", [1, 7, 1000, 65536, 100003])
@pytest.mark.parametrize("reproMode", [0, 1])
def test_dsdot_dsnrm2_bitwise(oz, oracle, n, reproMode):
    say(f"dsdot n={n} reproMode={reproMode}")
    x = oracle.init_mat(n, 1, n, 2.0, 2, 21)
    y = oracle.init_mat(n, 1, n, 2.0, 2, 22)
    H = oz.Handle(1 << 28, sumMode=0, reproMode=reproMode)
    h = oracle.Handle(1 << 30, sumMode=0, reproMode=reproMode)
    dx = dev(x)
    got = H.dsdot(n, dx, dev(y))
    want = h.dsdot(n, x, y)
    assert np.float64(got).tobytes() == np.float64(want).tobytes(), (got.hex(), want.hex())
    assert (H.nSplitA, H.nSplitB) == (h.nSplitA, h.nSplitB)
    assert H.dsnrm2(n, dx) == oracle.Handle(1 << 30, sumMode=0, reproMode=reproMode).dsnrm2(n, x)


def test_dsdot_split_edge_cases_bitwise(oz, oracle):
    """Wide ranges, exact-early splits, zeros and inputs of the fallback split loop.  Products whose
    value falls below the binary64 normal range are left out: the reference rounds each scaled
    slice-pair product once when it scales it back (scalbn into the subnormals), the FP64 products
    of unscaled slices round each element product there."""
    rng = np.random.default_rng(51)
    vecs = [rng.uniform(-1, 1, 5000), rng.standard_normal(3000) * np.exp(8 * rng.standard_normal(3000)),
            np.ldexp(rng.integers(-2**20, 2**20, 2000).astype(float), -30), np.zeros(300),
            np.array([1.5 * 2.0 ** 1022, -3.0, 1.0, 2.0 ** 1000])]
    for x in vecs:
        n = len(x)
        y = np.random.default_rng(n).uniform(-1, 1, n)
        h = oracle.Handle(1 << 30, sumMode=0)
        want = h.dsdot(n, x, y)
        H = oz.Handle(1 << 28, sumMode=0)
        say(f"dsdot edge n={n}")
        got = H.dsdot(n, dev(x), dev(y))
        same = np.float64(got).tobytes() == np.float64(want).tobytes()
        assert same or (np.isnan(got) and np.isnan(want)), (n, got, want)


@pytest.mark.parametrize("presplit", [False, True])
@pytest.mark.parametrize("kind", ["laplace", "wide-x"])
def test_dscsrmv_bitwise(oz, oracle, presplit, kind):
    rowptr, colind, val = oracle.laplace3d(12, 1e-3)
    n = 12 ** 3
    nnz = len(val)
    rng = np.random.default_rng(3)
    e = 20 if kind == "laplace" else 120
    x = rng.uniform(-1, 1, n) * np.exp2(rng.integers(-e, e, n))
    y0 = rng.standard_normal(n)
    say(f"dscsrmv {kind} presplit={presplit}")
    h = oracle.Handle(1 << 30, sumMode=0)
    H = oz.Handle(1 << 28, sumMode=0)
    dval, dcol, drow, dx = dev(val), dev(colind), dev(rowptr), dev(x)
    yw = y0.copy()
    dy = dev(y0)
    if presplit:
        p = h.dscsrmv_splitA(n, n, nnz, val, rowptr)
        P = H.dscsrmv_splitA(n, n, nnz, dval, drow)
        assert H.nSplitA_ == h.nSplitA_ and H.splitShift == h.splitShift
        h.dscsrmv(n, n, nnz, 1.5, p, colind, rowptr, x, -0.25, yw)
        H.dscsrmv(n, n, nnz, 1.5, P, dcol, drow, dx, -0.25, dy)
    else:
        h.dscsrmv(n, n, nnz, 1.5, val, colind, rowptr, x, -0.25, yw)
        H.dscsrmv(n, n, nnz, 1.5, dval, dcol, drow, dx, -0.25, dy)
    assert bits_equal(host(dy), yw), first_diff(host(dy), yw)
    assert H.nSplitB == h.nSplitB


@pytest.mark.parametrize("path", ["device", "host"])
def test_dscg_history_bitwise(oz, oracle, path):
    g, iters = 10, 40
    rowptr, colind, val = oracle.laplace3d(g, 1e-3)
    n, nnz = g ** 3, len(val)
    oracle.srand48(3)
    b = np.array([2.0 * oracle.drand48() - 1.0 for _ in range(n)])
    h = oracle.Handle(1 << 30, sumMode=0)
    xw = np.zeros(n)
    want = h.dscg(n, nnz, val, colind, rowptr, b, xw, iters, 0.0)
    H = oz.Handle(1 << 28, sumMode=0, cgDeviceSlices=0 if path == "device" else -1)
    dx = dev(np.zeros(n))
    say(f"dscg n={n} iters={iters} path={path}")
    got = H.dscg(n, nnz, dev(val), dev(colind), dev(rowptr), dev(b), dx, iters, 0.0)
    assert H.cgDevicePath in ((1, 2) if path == "device" else (0,))
    assert len(got) == iters + 1
    assert bits_equal(got, want), first_diff(got, want)
    assert bits_equal(host(dx), xw), first_diff(host(dx), xw)


def test_mixed_nearsum_and_dot2_not_supported(oz):
    """The reference's mixed NearSum and its float Dot2 GEMM are not available."""
    x = dev(np.ones(16))
    for flags in (dict(sumMode=1), dict(sumMode=0, splitEpsMode=2)):
        H = oz.Handle(1 << 24, **flags)
        with pytest.raises(oz.OzblasError) as e:
            H.dsdot(16, x, x)
        assert e.value.status == oz.NOT_SUPPORTED
        C = dev(np.zeros(16))
        with pytest.raises(oz.OzblasError) as e:
            H.dsgemm("n", "n", 4, 4, 4, 1.0, x, 4, x, 4, 0.0, C, 4)
        assert e.value.status == oz.NOT_SUPPORTED
