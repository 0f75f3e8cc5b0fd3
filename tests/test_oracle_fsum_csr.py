"""Pins of the oracle's CSR split, CSRMV and CG under Fsum (sumMode = 0), whose outputs the reference
ships no log for (its CSRMV/CG logs use matrices that are not in the checkout):

  - CSRMV reduces to GEMV: a CSR matrix whose rows are all stored in full (row nnz = n, explicit
    zeros included) splits row by row with the rho and tau of op(A)'s columns in the GEMV path
    (src/cpu/ozblas_split.cpp:806-809 against :156-159), its slice products are exact in either
    path, and both sum with sumOrder 3 (src/cpu/ozblas_csrmv.cpp:159-166, ozblas_gemm.cpp:290).
    So the oracle's Fsum CSRMV must equal its Fsum GEMV('N') bit for bit -- and that GEMV is pinned
    by the reference's _dgemv_n_sum0 check logs (test_oracle_kat.py).
  - The shifted split of ozblasRcsrmvSplitA (sigma = 0.75 2^(rho + tau) / splitShift,
    src/cpu/ozblas_split.cpp:809, and the doubling search of splitShift, ozblas_csrmv.cpp:238-252)
    against a closed form in exact rationals: slice s of row i is u round-half-even(r / u) with
    u = 2^(rho_i + tau_is - 53) / splitShift (the unit of sigma), tau from the C library's log2,
    and the count and the chosen shift follow from those slices.  A shift applied the wrong way
    (sigma * splitShift) gives units 2^(2 log2 shift) too coarse and fails.
  - CG under Fsum composed from those pinned primitives in the order of src/cpu/ozblas_cg.cpp:24-128
    (SplitA, NRM2, CSRMV, DOT, AXPY = fma, SCAL = IEEE multiply) equals the oracle's CG.
"""
import math
from fractions import Fraction

import numpy as np
import pytest


def dense_rows_csr(A, m, n):
    """CSR of the column-major m x n matrix A with every entry stored (explicit zeros too)."""
    Am = A.reshape(n, m).T
    rowptr = (np.arange(m + 1) * n).astype(np.int32)
    colind = np.tile(np.arange(n, dtype=np.int32), m)
    val = np.ascontiguousarray(Am).ravel()
    return rowptr, colind, val


@pytest.mark.parametrize("m,n,phi", [(40, 64, 1.0), (33, 300, 4.0), (7, 1000, 0.5), (120, 5, 2.0),
                                     (1, 513, 8.0)])
@pytest.mark.parametrize("flags", [dict(), dict(nSplitMax=2), dict(nSplitMax=3), dict(fastMode=1)])
def test_dense_row_csrmv_equals_gemv_fsum(oracle, m, n, phi, flags):
    A = oracle.init_mat(m, n, m, phi, 2, 11)
    A[::5] = 0.0  # explicit zeros stay stored entries of the CSR rows
    x = oracle.init_mat(n, 1, n, phi, 2, 12)
    rowptr, colind, val = dense_rows_csr(A, m, n)
    rng = np.random.default_rng(m * n)
    y0 = rng.standard_normal(m)
    for alpha, beta in ((1.0, 0.0), (-0.75, 1.5)):
        h1 = oracle.Handle(1 << 28, sumMode=0, reproMode=1, **flags)
        y1 = y0.copy()
        h1.csrmv(m, n, len(val), alpha, val, colind, rowptr, x, beta, y1)
        h2 = oracle.Handle(1 << 28, sumMode=0, reproMode=1, **flags)
        y2 = y0.copy()
        h2.gemv("N", m, n, alpha, A, m, x, beta, y2)
        assert (h1.nSplitA, h1.nSplitB) == (h2.nSplitA, h2.nSplitB)
        assert y1.view(np.uint64).tolist() == y2.view(np.uint64).tolist()


def rho_of(d):
    """getRho<double,double>, splitEpsMode 0 (src/cpu/ozblas_split.cpp:20-34), with libm's log2."""
    return int(math.ceil(53 - (53 - math.log2(float(d))) / 2))


def rhe(q):
    """round-half-even of a Fraction to an integer"""
    f = math.floor(q)
    d = q - f
    if d > Fraction(1, 2) or (d == Fraction(1, 2) and f % 2):
        return f + 1
    return f


def closed_form_split(vals, rowptr, shift, cap):
    """Slices (lists of Fractions per slot) and the count of the row split with splitShift = shift,
    from units u = 2^(rho + tau - 53) / shift, for reproMode = 1 and the reference's loop: split 0,
    then splits 1.. until every remainder is zero (the count is then index + 1), at most cap."""
    m = len(rowptr) - 1
    rem = [Fraction(float(v)) for v in vals]
    slices = []
    for s in range(cap):
        out = [Fraction(0)] * len(vals)
        total = Fraction(0)
        for i in range(m):
            a, b = int(rowptr[i]), int(rowptr[i + 1])
            mx = max((abs(rem[q]) for q in range(a, b)), default=Fraction(0))
            if mx != 0:
                tau = int(math.ceil(math.log2(float(mx))))
                u = Fraction(2) ** (rho_of(b - a) + tau - 53) / shift
                for q in range(a, b):
                    out[q] = u * rhe(rem[q] / u)
            nmx = Fraction(0)
            for q in range(a, b):
                rem[q] -= out[q]
                nmx = max(nmx, abs(rem[q]))
            total += nmx
        slices.append(out)
        if s >= 1 and total == 0:
            return slices, s + 1
    return slices, cap


def tuned_shift(vals, rowptr, cap):
    """The splitShift search of ozblasRcsrmvSplitA (nSplitMax = 0): double the shift while the count
    stays the same, up to 512; keep the first shift that changes it, else 1."""
    _, c = closed_form_split(vals, rowptr, 1, cap)
    shift = 1
    while True:
        old = c
        shift *= 2
        _, c = closed_form_split(vals, rowptr, shift, cap)
        if not (old == c and shift < 512):
            break
    if old == c:
        shift = 1
    return shift


def csr_matrices(oracle):
    out = []
    rowptr, colind, val = oracle.laplace3d(3, 1e-3)  # the CG matrix on a small grid
    out.append(("laplace", rowptr, colind, val))
    rng = np.random.default_rng(17)
    m = 12
    nnz_rows = rng.integers(1, 9, m)
    rowptr = np.concatenate([[0], np.cumsum(nnz_rows)]).astype(np.int32)
    val = rng.standard_normal(int(rowptr[-1])) * np.exp2(rng.integers(-20, 20, int(rowptr[-1])).astype(float))
    colind = np.concatenate([np.sort(rng.choice(m, k, replace=False)) for k in nnz_rows]).astype(np.int32)
    out.append(("random rows", rowptr, colind, val))
    vals2 = np.round(val * 2 ** 10) / 2 ** 10  # fewer bits: a split that the shift shortens
    out.append(("short mantissas", rowptr, colind, vals2))
    return out


def test_shifted_csr_split_closed_form(oracle):
    cap = 30  # nSplitMax = 0, reproMode = 1: NumSplitDefaultMax
    shifts_seen = set()
    for name, rowptr, colind, val in csr_matrices(oracle):
        m = len(rowptr) - 1
        nnz = len(val)
        h = oracle.Handle(1 << 28, sumMode=0, reproMode=1)
        ptr = h.csrmv_splitA(m, m, nnz, val, rowptr)
        shift = tuned_shift(val, rowptr, cap)
        slices, count = closed_form_split(val, rowptr, shift, cap)
        assert (int(h.splitShift), int(h.nSplitA_)) == (shift, count), name
        shifts_seen.add(shift)
        lds = oracle.pitch(nnz)
        got = np.ctypeslib.as_array((np.ctypeslib.ctypes.c_double * (lds * count)).from_address(ptr))
        for s in range(count):
            want = np.array([float(v) for v in slices[s]])
            have = got[s * lds: s * lds + nnz]
            assert have.view(np.uint64).tolist() == want.view(np.uint64).tolist(), (name, s)
    assert len(shifts_seen - {1}) > 0  # at least one matrix takes a shift > 1


def test_shift_direction_is_pinned(oracle):
    """The closed form discriminates sigma / shift from sigma * shift: with the units of the other
    reading, slice 0 of a row differs."""
    rowptr = np.array([0, 3], dtype=np.int32)
    val = np.array([1.0 / 3.0, -2.0 / 7.0, 0.1])
    a, _ = closed_form_split(val, rowptr, 4, 30)
    b, _ = closed_form_split(val, rowptr, Fraction(1, 4), 30)
    assert a[0] != b[0]


def _cg_fsum_composed(oracle, val, colind, rowptr, b, iters):
    n = len(rowptr) - 1
    nnz = len(val)
    h = oracle.Handle(1 << 28, sumMode=0, reproMode=1)
    ptr = h.csrmv_splitA(n, n, nnz, val, rowptr)
    nrmb = h.nrm2(n, b)
    x = np.zeros(n)
    r = b.copy()
    h.csrmv(n, n, nnz, -1.0, ptr, colind, rowptr, x, 1.0, r)
    dold = h.dot(n, r, r)
    hist = [math.sqrt(dold) / nrmb]
    p = r.copy()
    for _ in range(iters):
        q = np.zeros(n)
        h.csrmv(n, n, nnz, 1.0, ptr, colind, rowptr, p, 0.0, q)
        alpha = dold / h.dot(n, p, q)
        h.axpy(n, alpha, p, x)
        h.axpy(n, -alpha, q, r)
        dnew = h.dot(n, r, r)
        beta = dnew / dold
        dold = dnew
        p = p * beta
        h.axpy(n, 1.0, r, p)
        hist.append(math.sqrt(dnew) / nrmb)
    return hist, x


@pytest.mark.parametrize("g,iters", [(4, 15), (6, 8)])
def test_cg_fsum_composed_from_pinned_primitives(oracle, g, iters):
    rowptr, colind, val = oracle.laplace3d(g, 1e-3)
    n = g ** 3
    oracle.srand48(3)
    b = np.array([2.0 * oracle.drand48() - 1.0 for _ in range(n)])
    h = oracle.Handle(1 << 28, sumMode=0, reproMode=1)
    x = np.zeros(n)
    hist = h.cg(n, len(val), val, colind, rowptr, b, x, iters, 0.0)
    want, xw = _cg_fsum_composed(oracle, val, colind, rowptr, b, iters)
    assert [float(v).hex() for v in hist] == [float(v).hex() for v in want]
    assert x.view(np.uint64).tolist() == xw.view(np.uint64).tolist()


def test_axpy_is_fma_and_nrm2_is_sqrt_of_dot(oracle):
    """The two primitives of the composition that no log covers: AXPY with reproMode = 1 is
    fma(alpha, x, y) per element (src/cpu/ozblas_axpy.cpp:4-27) and NRM2 is sqrt of the DOT
    (src/cpu/ozblas_nrm2.cpp:27)."""
    rng = np.random.default_rng(2)
    n = 1000
    xv = rng.standard_normal(n) * np.exp2(rng.integers(-30, 30, n).astype(float))
    yv = rng.standard_normal(n)
    h = oracle.Handle(1 << 28, sumMode=0, reproMode=1)
    y = yv.copy()
    h.axpy(n, 0.3, xv, y)
    want = [float(Fraction(0.3) * Fraction(float(a)) + Fraction(float(c))) for a, c in zip(xv, yv)]
    assert y.tolist() == want
    assert h.nrm2(n, xv) == math.sqrt(h.dot(n, xv, xv))
