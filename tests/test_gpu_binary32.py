"""GPU parity of the binary32 entry points (ozblasS{gemm,gemv,dot,nrm2,axpy,csrmv,csrmvSplitA,cg})
through the C ABI, bitwise against the oracle's restatement of the reference's <float,float,float>
instantiations (oracle/ozoracle_s.c)."""
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
    # m, n, k, transA, transB
    (64, 48, 100, "n", "n"),
    (257, 130, 300, "t", "n"),
    (130, 257, 129, "n", "t"),
    (33, 65, 1000, "t", "t"),
    (200, 1, 300, "n", "n"),    # the GEMV branch of the reference GEMM
    (1, 77, 64, "t", "n"),
    (1, 1, 500, "t", "n"),      # the DOT branch
]


@pytest.mark.parametrize("case", GEMM_CASES, ids=lambda c: "x".join(map(str, c[:3])) + c[3] + c[4])
@pytest.mark.parametrize("sumMode", [0, 1])
@pytest.mark.parametrize("phi", [0.5, 4.0])
def test_sgemm_bitwise(oz, oracle, case, sumMode, phi):
    m, n, k, ta, tb = case
    say(f"sgemm {m}x{n}x{k} {ta}{tb} sumMode={sumMode} phi={phi}")
    ra, ca = (m, k) if ta == "n" else (k, m)
    rb, cb = (k, n) if tb == "n" else (n, k)
    A = _mat(oracle, ra, ca, phi, 1)
    B = _mat(oracle, rb, cb, phi, 2)
    C0 = _mat(oracle, m, n, 1.0, 3)
    alpha, beta = np.float32(1.25), np.float32(-0.5)
    want = C0.copy()
    oracle.Handle(1 << 30, sumMode=sumMode).sgemm(ta, tb, m, n, k, alpha, A, ra, B, rb, beta, want, m)
    H = oz.Handle(1 << 28, sumMode=sumMode)
    dC = dev(C0)
    H.sgemm(ta, tb, m, n, k, float(alpha), dev(A), ra, dev(B), rb, float(beta), dC, m)
    got = host(dC)
    assert f32_equal(got, want), f32_diff(got, want)


@pytest.mark.parametrize("flags", [dict(reproMode=0), dict(nSplitMax=3), dict(splitEpsMode=9),
                                   dict(fastMode=1), dict(nSplitMax=2, reproMode=0, sumMode=1)],
                         ids=lambda f: ",".join(f"{k}={v}" for k, v in f.items()))
def test_sgemm_options_bitwise(oz, oracle, flags):
    m, n, k = 96, 80, 257
    say(f"sgemm options {flags}")
    A = _mat(oracle, m, k, 2.0, 4)
    B = _mat(oracle, k, n, 2.0, 5)
    C0 = _mat(oracle, m, n, 1.0, 6)
    want = C0.copy()
    oracle.Handle(1 << 30, **flags).sgemm("n", "n", m, n, k, np.float32(1), A, m, B, k,
                                          np.float32(0), want, m)
    H = oz.Handle(1 << 28, **flags)
    dC = dev(C0)
    H.sgemm("n", "n", m, n, k, 1.0, dev(A), m, dev(B), k, 0.0, dC, m)
    assert f32_equal(host(dC), want), f32_diff(host(dC), want)


def test_sgemm_blocked_workspace_bitwise(oz, oracle):
    """A workspace too small for the whole problem: the m/n blocks give the same bits."""
    m, n, k = 300, 260, 200
    A = _mat(oracle, m, k, 1.0, 7)
    B = _mat(oracle, k, n, 1.0, 8)
    want = np.zeros(m * n, dtype=np.float32)
    oracle.Handle(1 << 30, sumMode=1).sgemm("n", "n", m, n, k, np.float32(1), A, m, B, k,
                                            np.float32(0), want, m)
    H = oz.Handle(6 << 20, sumMode=1)
    dC = dev(np.zeros(m * n, dtype=np.float32))
    say("sgemm blocked")
    H.sgemm("n", "n", m, n, k, 1.0, dev(A), m, dev(B), k, 0.0, dC, m)
    assert (H.mbk, H.nbk) != (m, n)
    assert f32_equal(host(dC), want), f32_diff(host(dC), want)


@pytest.mark.parametrize("trans", ["n", "t"])
@pytest.mark.parametrize("sumMode", [0, 1])
def test_sgemv_bitwise(oz, oracle, trans, sumMode):
    m, n = 301, 177
    A = _mat(oracle, m, n, 2.0, 9)
    lx, ly = (n, m) if trans == "n" else (m, n)
    x = _mat(oracle, lx, 1, 2.0, 10)
    y0 = _mat(oracle, ly, 1, 1.0, 11)
    want = y0.copy()
    oracle.Handle(1 << 30, sumMode=sumMode).sgemv(trans, m, n, np.float32(0.75), A, m, x,
                                                  np.float32(2), want)
    H = oz.Handle(1 << 28, sumMode=sumMode)
    dy = dev(y0)
    say(f"sgemv {trans} sumMode={sumMode}")
    H.sgemv(trans, m, n, 0.75, dev(A), m, dev(x), 2.0, dy)
    assert f32_equal(host(dy), want), f32_diff(host(dy), want)


@pytest.mark.parametrize("n", [1, 2, 7, 1000, 65536, 100003])
@pytest.mark.parametrize("sumMode", [0, 1])
@pytest.mark.parametrize("reproMode", [0, 1])
def test_sdot_snrm2_bitwise(oz, oracle, n, sumMode, reproMode):
    say(f"sdot n={n} sumMode={sumMode} reproMode={reproMode}")
    x = _mat(oracle, n, 1, 2.0, 21)
    y = _mat(oracle, n, 1, 2.0, 22)
    flags = dict(sumMode=sumMode, reproMode=reproMode)
    H = oz.Handle(1 << 28, **flags)
    dx, dy = dev(x), dev(y)
    got = np.float32(H.sdot(n, dx, dy))
    want = oracle.Handle(1 << 30, **flags).sdot(n, x, y)
    assert f32_equal(got, want), (float(got).hex(), float(want).hex())
    assert f32_equal(np.float32(H.snrm2(n, dx)), oracle.Handle(1 << 30, **flags).snrm2(n, x))


def _edge_vectors_f():
    rng = np.random.default_rng(41)
    yield "uniform", rng.uniform(-1, 1, 5000).astype(np.float32)
    yield "wide", (rng.standard_normal(3000) * np.exp(6 * rng.standard_normal(3000))).astype(np.float32)
    yield "few-bits", np.ldexp(rng.integers(-2**12, 2**12, 2000).astype(np.float32), -20)
    v = np.ldexp(np.float32(1), rng.integers(-30, 30, 1500)).astype(np.float32)
    v[::3] *= -1
    yield "powers-of-two", v
    w = rng.uniform(-1, 1, 1200).astype(np.float32)
    w[::7] = np.nextafter(np.float32(0.125), np.float32(1))
    w[1::11] = 0
    yield "just-above-2^e", w
    yield "ties", np.ldexp(rng.integers(0, 2**23, 700).astype(np.float32) + np.float32(0.5), -20)
    yield "zeros", np.zeros(300, dtype=np.float32)
    yield "two-ranges", np.concatenate([rng.uniform(-1, 1, 500),
                                        rng.uniform(-1, 1, 500) * 1e-30]).astype(np.float32)
    # outside the residue model: the one-workgroup reference loop splits these
    yield "huge", np.array([3e38, -3.0, 1.0, 2.0 ** 100], dtype=np.float32)
    yield "subnormal", np.array([1e-40, -3e-42, 1e-45, 0.0], dtype=np.float32)


@pytest.mark.parametrize("sumMode", [0, 1])
def test_sdot_split_edge_cases_bitwise(oz, oracle, sumMode):
    for name, x in _edge_vectors_f():
        n = len(x)
        y = np.random.default_rng(n).uniform(-1, 1, n).astype(np.float32)
        for a, b in ((x, x), (x, y)):
            h = oracle.Handle(1 << 30, sumMode=sumMode)
            want = h.sdot(n, a, b)
            H = oz.Handle(1 << 28, sumMode=sumMode)
            say(f"sdot edge {name} n={n} sumMode={sumMode}")
            got = np.float32(H.sdot(n, dev(a), dev(b)))
            same = f32_equal(got, want)
            assert same or (np.isnan(got) and np.isnan(want)), (name, got, want)
            assert (H.nSplitA, H.nSplitB) == (h.nSplitA, h.nSplitB), name


def test_sdot_is_correctly_rounded_under_nearsum(oz, oracle):
    """With sumMode = 1 the result is the binary32 rounding of the exact dot product."""
    from fractions import Fraction
    rng = np.random.default_rng(9)
    for n in (3, 999, 20000):
        x = (rng.standard_normal(n) * np.exp2(rng.integers(-20, 20, n))).astype(np.float32)
        y = rng.standard_normal(n).astype(np.float32)
        exact = sum(Fraction(float(a)) * Fraction(float(b)) for a, b in zip(x, y))
        H = oz.Handle(1 << 28, sumMode=1)
        got = np.float32(H.sdot(n, dev(x), dev(y)))
        want = np.float32(float(exact))  # double first: exact ties at binary32 cannot arise here
        assert abs(Fraction(float(got)) - exact) <= abs(Fraction(float(want)) - exact)


@pytest.mark.parametrize("reproMode", [0, 1])
def test_saxpy_bitwise(oz, oracle, reproMode):
    n = 10007
    x = _mat(oracle, n, 1, 3.0, 12)
    y0 = _mat(oracle, n, 1, 3.0, 13)
    want = y0.copy()
    oracle.Handle(1 << 24, reproMode=reproMode).saxpy(n, np.float32(-1.75), x, want)
    H = oz.Handle(1 << 24, reproMode=reproMode)
    dy = dev(y0)
    H.saxpy(n, -1.75, dev(x), dy)
    assert f32_equal(host(dy), want), f32_diff(host(dy), want)


def _laplace_f(oracle, g):
    rowptr, colind, val = oracle.laplace3d(g, 1e-3)
    return rowptr, colind, val.astype(np.float32), g ** 3, len(val)


def _random_csr(rng, m, n, maxrow):
    lens = rng.integers(0, maxrow + 1, m)
    lens[::17] = 0  # empty rows
    rowptr = np.zeros(m + 1, dtype=np.int32)
    rowptr[1:] = np.cumsum(lens)
    colind = np.concatenate([np.sort(rng.choice(n, int(L), replace=False)) for L in lens]).astype(np.int32)
    val = (rng.standard_normal(int(rowptr[-1])) * np.exp2(rng.integers(-8, 8, int(rowptr[-1])))).astype(np.float32)
    return rowptr, colind, val


@pytest.mark.parametrize("sumMode", [0, 1])
@pytest.mark.parametrize("presplit", [False, True])
@pytest.mark.parametrize("kind", ["laplace", "random", "wide-x"])
def test_scsrmv_bitwise(oz, oracle, sumMode, presplit, kind):
    rng = np.random.default_rng(2)
    if kind == "laplace":
        rowptr, colind, val, n, nnz = _laplace_f(oracle, 12)
        m = n
    else:
        m, n = 900, 700
        rowptr, colind, val = _random_csr(rng, m, n, 40)
        nnz = len(val)
    x = (rng.uniform(-1, 1, n) * np.exp2(rng.integers(-20, 20, n))).astype(np.float32)
    if kind == "wide-x":  # many vector slices: the general (stored products) kernel
        x = (rng.uniform(-1, 1, n) * np.exp2(rng.integers(-60, 60, n))).astype(np.float32)
    y0 = rng.standard_normal(m).astype(np.float32)
    say(f"scsrmv {kind} sumMode={sumMode} presplit={presplit}")
    h = oracle.Handle(1 << 30, sumMode=sumMode)
    H = oz.Handle(1 << 28, sumMode=sumMode)
    dval, dcol, drow, dx = dev(val), dev(colind), dev(rowptr), dev(x)
    yw = y0.copy()
    dy = dev(y0)
    if presplit:
        p = h.scsrmv_splitA(m, n, nnz, val, rowptr)
        P = H.scsrmv_splitA(m, n, nnz, dval, drow)
        assert H.nSplitA_ == h.nSplitA_ and H.splitShift == h.splitShift
        h.scsrmv(m, n, nnz, np.float32(1.5), p, colind, rowptr, x, np.float32(-0.25), yw)
        H.scsrmv(m, n, nnz, 1.5, P, dcol, drow, dx, -0.25, dy)
    else:
        h.scsrmv(m, n, nnz, np.float32(1.5), val, colind, rowptr, x, np.float32(-0.25), yw)
        H.scsrmv(m, n, nnz, 1.5, dval, dcol, drow, dx, -0.25, dy)
    assert f32_equal(host(dy), yw), f32_diff(host(dy), yw)
    assert H.nSplitB == h.nSplitB


@pytest.mark.parametrize("sumMode", [0, 1])
@pytest.mark.parametrize("reproMode", [0, 1])
def test_scg_history_bitwise(oz, oracle, sumMode, reproMode):
    g, iters = 10, 40
    rowptr, colind, val, n, nnz = _laplace_f(oracle, g)
    oracle.srand48(3)
    b = np.array([2.0 * oracle.drand48() - 1.0 for _ in range(n)], dtype=np.float32)
    say(f"scg n={n} iters={iters} sumMode={sumMode} reproMode={reproMode}")
    flags = dict(sumMode=sumMode, reproMode=reproMode)
    h = oracle.Handle(1 << 30, **flags)
    xw = np.zeros(n, dtype=np.float32)
    want = h.scg(n, nnz, val, colind, rowptr, b, xw, iters, np.float32(0))
    H = oz.Handle(1 << 28, **flags)
    dx = dev(np.zeros(n, dtype=np.float32))
    got = H.scg(n, nnz, dev(val), dev(colind), dev(rowptr), dev(b), dx, iters, 0.0)
    assert len(got) == iters + 1
    assert f32_equal(got.astype(np.float32), want), f32_diff(got, want)
    assert f32_equal(host(dx), xw), f32_diff(host(dx), xw)


def test_scg_tolerance_stop_bitwise(oz, oracle):
    g, iters = 10, 80
    rowptr, colind, val, n, nnz = _laplace_f(oracle, g)
    oracle.srand48(7)
    b = np.array([2.0 * oracle.drand48() - 1.0 for _ in range(n)], dtype=np.float32)
    full = oracle.Handle(1 << 30, sumMode=1).scg(n, nnz, val, colind, rowptr, b,
                                                 np.zeros(n, dtype=np.float32), iters, np.float32(0))
    tol = np.float32(np.nextafter(full[20], np.float32(1)))
    h = oracle.Handle(1 << 30, sumMode=1)
    xw = np.zeros(n, dtype=np.float32)
    want = h.scg(n, nnz, val, colind, rowptr, b, xw, iters, tol)
    assert 1 < len(want) - 1 < iters
    H = oz.Handle(1 << 28, sumMode=1)
    dx = dev(np.zeros(n, dtype=np.float32))
    say(f"scg tol stop n={n}")
    got = H.scg(n, nnz, dev(val), dev(colind), dev(rowptr), dev(b), dx, iters, float(tol))
    assert f32_equal(got.astype(np.float32), want), f32_diff(got, want)
    assert f32_equal(host(dx), xw)


@pytest.mark.parametrize("sumMode", [0, 1])
@pytest.mark.parametrize("path", ["device", "host", "overflow"])
def test_scg_paths_bitwise(oz, oracle, sumMode, path):
    """The binary32 CG iterated device-resident, synchronously, and falling back from the device
    path (a slice bound of 1: the first CSR product gives up) gives the oracle's bits."""
    g, iters = 12, 60
    rowptr, colind, val, n, nnz = _laplace_f(oracle, g)
    oracle.srand48(5)
    b = np.array([2.0 * oracle.drand48() - 1.0 for _ in range(n)], dtype=np.float32)
    say(f"scg paths n={n} iters={iters} sumMode={sumMode} path={path}")
    h = oracle.Handle(1 << 30, sumMode=sumMode)
    xw = np.zeros(n, dtype=np.float32)
    want = h.scg(n, nnz, val, colind, rowptr, b, xw, iters, np.float32(0))
    bound = {"device": 0, "host": -1, "overflow": 1}[path]
    H = oz.Handle(1 << 28, sumMode=sumMode, cgDeviceSlices=bound)
    dx = dev(np.zeros(n, dtype=np.float32))
    got = H.scg(n, nnz, dev(val), dev(colind), dev(rowptr), dev(b), dx, iters, 0.0)
    assert H.cgDevicePath == {"device": 1, "host": 0, "overflow": 2}[path]
    assert len(got) == iters + 1
    assert f32_equal(got.astype(np.float32), want), f32_diff(got, want)
    assert f32_equal(host(dx), xw), f32_diff(host(dx), xw)


@pytest.mark.parametrize("path", ["device", "host"])
def test_scg_tolerance_stop_paths_bitwise(oz, oracle, path):
    g, iters = 10, 80
    rowptr, colind, val, n, nnz = _laplace_f(oracle, g)
    oracle.srand48(7)
    b = np.array([2.0 * oracle.drand48() - 1.0 for _ in range(n)], dtype=np.float32)
    full = oracle.Handle(1 << 30, sumMode=1).scg(n, nnz, val, colind, rowptr, b,
                                                 np.zeros(n, dtype=np.float32), iters, np.float32(0))
    tol = np.float32(np.nextafter(full[20], np.float32(1)))
    h = oracle.Handle(1 << 30, sumMode=1)
    xw = np.zeros(n, dtype=np.float32)
    want = h.scg(n, nnz, val, colind, rowptr, b, xw, iters, tol)
    assert 1 < len(want) - 1 < iters
    H = oz.Handle(1 << 28, sumMode=1, cgDeviceSlices=0 if path == "device" else -1)
    dx = dev(np.zeros(n, dtype=np.float32))
    say(f"scg tol stop path={path} n={n}")
    got = H.scg(n, nnz, dev(val), dev(colind), dev(rowptr), dev(b), dx, iters, float(tol))
    assert H.cgDevicePath == (1 if path == "device" else 0)
    assert f32_equal(got.astype(np.float32), want), f32_diff(got, want)
    assert f32_equal(host(dx), xw)


def test_binary32_dot2_not_supported(oz):
    """The reference's binary32 Dot2 GEMM is a stub; the entry points report NOT_SUPPORTED."""
    H = oz.Handle(1 << 24, splitEpsMode=2)
    x = dev(np.ones(16, dtype=np.float32))
    with pytest.raises(oz.OzblasError) as e:
        H.sdot(16, x, x)
    assert e.value.status == oz.NOT_SUPPORTED
    C = dev(np.zeros(16, dtype=np.float32))
    with pytest.raises(oz.OzblasError) as e:
        H.sgemm("n", "n", 4, 4, 4, 1.0, x, 4, x, 4, 0.0, C, 4)
    assert e.value.status == oz.NOT_SUPPORTED
