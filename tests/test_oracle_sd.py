"""Pins the mixed binary32/binary64 oracle restatement (oracle/ozoracle_sd.c, the reference's
<float,double,double> / <float,double> instantiations).  The reference ships no SD logs, so:
  - rho = getRho<float,double> is the least integer with 32 * 4^rho >= d (mode 0) and
    256 * 16^rho >= d (mode 9): checked on integers, with no logarithm;
  - the SD split of a column is the binary32 split of the <float,float> instantiation at the same
    rho (the S split in splitEpsMode 2 has getRho<float,float> = ceil(24 - (53 - log2 d) / 2),
    the SD mode-0 value): SD slice * 2^tau equals the S slice, tau is the tau of the column's
    current maximum, and the slice counts agree under reproMode 1 (no slot rule on either side);
  - the slices sum back to the input exactly;
  - DOT / GEMV / GEMM / CSRMV of integer data are exact (every float partial sum is an integer
    below 2^24), and DOT of general data stays within the bound of the float Fsum of its terms;
  - the mixed NearSum and Dot2 are refused (-3), as the plain bypass.
"""
from fractions import Fraction as F
import math

import numpy as np
import pytest


def _split(oracle, x, cols, h):
    L = oracle.lib()
    n = x.size // cols
    cap = 32
    out = np.zeros(x.size, dtype=np.float32)
    split = np.zeros(n * cols * cap, dtype=np.float64)
    exps = np.zeros(cols * cap, dtype=np.int16)
    mx = np.zeros(max(n, cols), dtype=np.float32)
    cnt = L.ozo_sd_split(h._h, n, cols, x, n, out, n, split, n, exps, cols, mx)
    return cnt, split.reshape(cap, cols, n), exps.reshape(cap, cols)


def _s_split(oracle, x, cols, h):
    L = oracle.lib()
    n = x.size // cols
    cap = 32
    out = np.zeros(x.size, dtype=np.float32)
    split = np.zeros(n * cols * cap, dtype=np.float32)
    mx = np.zeros(max(n, cols), dtype=np.float32)
    cnt = L.ozo_s_split(h._h, n, cols, x, n, out, n, split, n, mx)
    return cnt, split.reshape(cap, cols, n)


def test_sd_rho_is_least_exact_rho(oracle):
    for d in list(range(1, 5000)) + [2 ** 20 - 1, 2 ** 20, 2 ** 20 + 1, 10 ** 6, 2 ** 31 - 1]:
        r = oracle.sd_rho(d, 0)
        assert F(32) * F(4) ** r >= d and F(32) * F(4) ** (r - 1) < d, (d, r)
        r9 = oracle.sd_rho(d, 9)
        assert F(256) * F(16) ** r9 >= d and F(256) * F(16) ** (r9 - 1) < d, (d, r9)
        assert r == oracle.s_rho(d, 2), d  # getRho<float,float> in Dot2 mode: the same formula


@pytest.mark.parametrize("cols", [1, 5])
@pytest.mark.parametrize("n", [7, 200, 3000])
def test_sd_split_equals_binary32_split_scaled(oracle, cols, n):
    rng = np.random.default_rng(n + cols)
    x = (rng.standard_normal(n * cols) * np.exp2(rng.integers(-30, 30, n * cols))).astype(np.float32)
    x[::11] = 0.0
    hs = oracle.Handle(1 << 26, reproMode=1, splitEpsMode=2)
    hsd = oracle.Handle(1 << 26, reproMode=1)
    cs, ss = _s_split(oracle, x.copy(), cols, hs)
    cd, sd, ed = _split(oracle, x.copy(), cols, hsd)
    assert cs == cd
    rem = x.reshape(cols, n).astype(np.float64)
    for s in range(cd):
        for j in range(cols):
            # tau of the column's current maximum (ozo_s_tau is pinned against libm elsewhere)
            mx = np.float32(np.max(np.abs(rem[j]))) if n else np.float32(0)
            assert ed[s, j] == oracle.s_tau(mx), (s, j)
            got = np.ldexp(sd[s, j], int(ed[s, j]))
            assert np.array_equal(got, ss[s, j].astype(np.float64)), (s, j)
            rem[j] = rem[j] - got
    assert not np.any(rem), "slices do not sum back to the input"


def test_sd_split_of_tiny_values_is_exact(oracle):
    """Scaled slices are multiples of 2^(rho-24) of magnitude at most 1, so scalbn(split, -tau)
    stays normal and exact in binary32 even for data near the bottom of the float range."""
    rng = np.random.default_rng(4)
    n = 500
    x = (rng.standard_normal(n) * 2.0 ** -120).astype(np.float32)
    c, sd, ed = _split(oracle, x.copy(), 1, oracle.Handle(1 << 26))
    tot = np.zeros(n)
    for s in range(c):
        tot += np.ldexp(sd[s, 0], int(ed[s, 0]))
    assert np.array_equal(tot, x.astype(np.float64))


def test_sd_integer_data_is_exact(oracle):
    rng = np.random.default_rng(5)
    h = oracle.Handle(1 << 28, sumMode=0)
    for n in (129, 1000, 4096):
        x = rng.integers(-40, 41, n).astype(np.float32)
        y = rng.integers(-40, 41, n).astype(np.float32)
        assert h.sddot(n, x, y) == np.float32(int(np.dot(x.astype(np.int64), y.astype(np.int64))))
        assert h.sdnrm2(n, x) == np.float32(math.sqrt(int(np.dot(x.astype(np.int64), x.astype(np.int64)))))
    m, n, k = 37, 23, 300
    A = rng.integers(-20, 21, m * k).astype(np.float32)  # column-major m x k
    B = rng.integers(-20, 21, k * n).astype(np.float32)
    C = np.zeros(m * n, dtype=np.float32)
    h.sdgemm("n", "n", m, n, k, 1.0, A, m, B, k, 0.0, C, m)
    want = A.reshape(k, m).T.astype(np.int64) @ B.reshape(n, k).T.astype(np.int64)
    assert np.array_equal(C.reshape(n, m).T, want.astype(np.float32))
    y = np.zeros(m, dtype=np.float32)
    h.sdgemv("n", m, k, 1.0, A, m, B[:k].copy(), 0.0, y)
    assert np.array_equal(y, (A.reshape(k, m).T.astype(np.int64) @ B[:k].astype(np.int64)).astype(np.float32))


def test_sd_csrmv_integer_data_is_exact(oracle):
    rng = np.random.default_rng(6)
    m = 400
    rowptr = np.zeros(m + 1, dtype=np.int32)
    cols, vals = [], []
    for r in range(m):
        c = np.sort(rng.choice(m, size=int(rng.integers(0, 12)), replace=False))
        cols.extend(c)
        vals.extend(rng.integers(-30, 31, c.size))
        rowptr[r + 1] = len(cols)
    colind = np.array(cols, dtype=np.int32)
    val = np.array(vals, dtype=np.float32)
    x = rng.integers(-50, 51, m).astype(np.float32)
    y0 = rng.integers(-50, 51, m).astype(np.float32)
    y = y0.copy()
    oracle.Handle(1 << 28, sumMode=0).sdcsrmv(m, m, colind.size, 2.0, val, colind, rowptr, x, -1.0, y)
    want = np.array([2 * sum(int(val[i]) * int(x[colind[i]]) for i in range(rowptr[r], rowptr[r + 1]))
                     - int(y0[r]) for r in range(m)], dtype=np.float32)
    assert np.array_equal(y, want)


def test_sd_dot_within_float_fsum_bound(oracle):
    """|SD dot - exact| <= (terms + 1) * u * sum|x_i y_i| with u = 2^-24: every term is a float
    rounding of an exact slice product and the Fsum adds them in float."""
    rng = np.random.default_rng(7)
    h = oracle.Handle(1 << 28, sumMode=0)
    for t in range(20):
        n = int(rng.integers(129, 5000))
        x = (rng.standard_normal(n) * np.exp2(rng.integers(-8, 8, n))).astype(np.float32)
        y = rng.standard_normal(n).astype(np.float32)
        got = F(float(h.sddot(n, x, y)))
        exact = sum(F(float(a)) * F(float(b)) for a, b in zip(x, y))
        absum = sum(abs(F(float(a)) * F(float(b))) for a, b in zip(x, y))
        terms = int(h.nSplitA * h.nSplitB)
        assert abs(got - exact) <= F(terms + 1, 2 ** 24) * absum, t


def test_sd_refuses_nearsum_and_dot2(oracle):
    L = oracle.lib()
    x = np.ones(200, dtype=np.float32)
    C = np.zeros(1, dtype=np.float32)
    for flags in (dict(sumMode=1), dict(sumMode=0, splitEpsMode=2), dict(reproMode=0, nSplitMax=1)):
        h = oracle.Handle(1 << 26, **flags)
        assert L.ozo_sd_gemm(h._h, b"t", b"n", 1, 1, 200, 1.0, x, 200, x, 200, 0.0, C, 1) == -3, flags


def test_sd_cg_solves_laplacian(oracle):
    n1 = 6
    nnz = oracle.lib().ozo_laplace3d_nnz(n1)
    n = n1 ** 3
    rowptr = np.zeros(n + 1, dtype=np.int32)
    colind = np.zeros(nnz, dtype=np.int32)
    valD = np.zeros(nnz, dtype=np.float64)
    oracle.lib().ozo_laplace3d(n1, 1.0, rowptr, colind, valD)
    val = valD.astype(np.float32)
    b = np.ones(n, dtype=np.float32)
    x = np.zeros(n, dtype=np.float32)
    hist = oracle.Handle(1 << 28, sumMode=0).sdcg(n, nnz, val, colind, rowptr, b, x, 200, 1e-5)
    assert hist[-1] < 1e-5 and len(hist) < 200
    r = b.astype(np.float64).copy()
    for i in range(n):
        for j in range(rowptr[i], rowptr[i + 1]):
            r[i] -= float(val[j]) * float(x[colind[j]])
    assert np.linalg.norm(r) / np.linalg.norm(b) < 1e-4
