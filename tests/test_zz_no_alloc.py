"""ozblasDdotNoAlloc (the DOT on the library's static scratch array) through the C ABI: bitwise
equal to ozblasDdot on two calls in a row, the second with a smaller input, and an input beyond the
64 MiB array rejected with OZBLAS_STATUS_WORKSPACE."""
import ctypes as C

import numpy as np
import pytest

from gpu_util import dev, say, torch_cuda

pytestmark = pytest.mark.gpu

MAX_N = 130432  # two distinct vectors, nSplitMax = 0, reproMode = 1


@pytest.fixture(scope="module")
def oz():
    torch_cuda()
    import ozblas_amd
    f = ozblas_amd._lib.ozblasDdotNoAlloc
    f.restype = C.c_int32
    f.argtypes = [C.POINTER(ozblas_amd.OzblasHandleT), C.c_int32, C.c_void_p, C.c_int32,
                  C.c_void_p, C.c_int32, C.POINTER(C.c_double)]
    return ozblas_amd, f


def _vec(oracle, n, phi, seed):
    return oracle.init_mat(n, 1, n, phi, 2, seed)


def test_no_alloc_dot_matches_ddot(oz, oracle):
    ozb, noalloc = oz
    H = ozb.Handle(1 << 28, sumMode=1)
    cases = [(MAX_N, 4.0, (1, 2)), (1000, 1.0, (3, 4)), (77, 8.0, (5, 6))]
    for n, phi, seeds in cases:
        x, y = _vec(oracle, n, phi, seeds[0]), _vec(oracle, n, phi, seeds[1])
        dx, dy = dev(x), dev(y)
        want = H.ddot(n, dx, dy)
        r = C.c_double(0.0)
        say(f"no-alloc ddot n={n} phi={phi}")
        st = noalloc(H.ref, n, dx.data_ptr(), 1, dy.data_ptr(), 1, C.byref(r))
        assert st == ozb.SUCCESS
        assert np.float64(r.value).view(np.uint64) == np.float64(want).view(np.uint64), (n, r.value, want)
        assert want == oracle.Handle(1 << 30, sumMode=1).dot(n, x, y)
    # the same vector twice (one split) and the Fsum mode
    H.set(sumMode=0)
    x = _vec(oracle, 5000, 2.0, 9)
    dx = dev(x)
    r = C.c_double(0.0)
    assert noalloc(H.ref, 5000, dx.data_ptr(), 1, dx.data_ptr(), 1, C.byref(r)) == ozb.SUCCESS
    assert np.float64(r.value).view(np.uint64) == np.float64(H.ddot(5000, dx, dx)).view(np.uint64)
    H.close()


def test_no_alloc_dot_rejects_beyond_scratch(oz, oracle):
    ozb, noalloc = oz
    H = ozb.Handle(1 << 28, sumMode=1)
    n = MAX_N + 1
    dx, dy = dev(np.ones(n)), dev(np.ones(n))
    r = C.c_double(-1.0)
    assert noalloc(H.ref, n, dx.data_ptr(), 1, dy.data_ptr(), 1, C.byref(r)) == ozb.WORKSPACE
    assert r.value == -1.0  # nothing ran
    # the handle still works afterwards, on its own workspace and on the scratch array
    assert H.ddot(n, dx, dy) == float(n)
    assert noalloc(H.ref, 100, dx.data_ptr(), 1, dy.data_ptr(), 1, C.byref(r)) == ozb.SUCCESS
    assert r.value == 100.0
    H.close()
