"""Replays the reference's check-mode sweeps with a given BLAS implementation.

It restates the harness of the reference (testing/testing_level3.cpp:16-297 for GEMM,
testing/testing_level1b.cpp for DOT, testing/testing_level2.cpp for GEMV, options and defaults at
testing/testing_common.cpp:233-458, inputs at :718-763, the check at :839-930) so that the
reference's logs in tests/golden/a100_kat.json can be compared with an implementation's output:
every row holds the max relative error (as the harness computes it in binary64) and the number of
entries that differ from RN(exact).

``impl`` is any object with ``gemm``, ``dot`` and ``gemv`` methods taking the oracle's argument
convention (column-major flat float64 buffers).  The same handle/implementation object is used for
the whole sweep, as in the harness.
"""
from __future__ import annotations

import numpy as np

from oracle import oracle as O


def harness_opts(opts):
    """Harness defaults (testing/testing_common.cpp:242-262) overridden by the command line."""
    th = dict(phi=1, reproMode=0, nSplitMax=0, sumMode=0, fastMode=0, useBatchedGemmFlag=0,
              splitEpsMode=0, tranA="N", tranB="N", m_const=0, n_const=0, k_const=0,
              start=16, stop=128, step=16, step_mode=0)
    if "summode" in opts:
        th["sumMode"] = int(opts["summode"])
    if "fastmode" in opts:
        th["fastMode"] = int(opts["fastmode"])
    if "usebatchedgemm" in opts:
        th["useBatchedGemmFlag"] = int(opts["usebatchedgemm"])
    if "repromode" in opts:
        th["reproMode"] = int(opts["repromode"])
    if "degree" in opts:
        th["nSplitMax"] = int(opts["degree"])
    if "phi" in opts:
        th["phi"] = int(opts["phi"])
    if "transa" in opts:
        th["tranA"] = opts["transa"][0]
    if "transb" in opts:
        th["tranB"] = opts["transb"][0]
    for key in ("m", "n", "k"):
        if key in opts:
            th[key + "_const"] = int(opts[key])
    if "range" in opts:
        a, b, c = opts["range"].split(":")
        th["start"], th["stop"] = int(a), int(b)
        if c == "pow":
            th["step_mode"] = 2
        else:
            th["step"] = int(c)
    return th


def handle_flags(th):
    return dict(nSplitMax=th["nSplitMax"], fastMode=th["fastMode"], reproMode=th["reproMode"],
                sumMode=th["sumMode"], useBatchedGemmFlag=th["useBatchedGemmFlag"],
                splitEpsMode=th["splitEpsMode"])


def check(target, ref, prec="d"):
    """mublasCheckMatrix comparison on FP_TYPE (MPFR_CHECK undefined): ref = ToFptype(mpref), which
    for binary32 is the binary64 RN of the exact value rounded again to float
    (testing/testing_common.cpp:621-640); differences and relative errors in FP_TYPE."""
    if prec == "s":
        target = np.asarray(target, dtype=np.float32).ravel()
        ref = np.asarray(ref, dtype=np.float64).astype(np.float32).ravel()
        with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
            dif = target - ref
            rerr = np.where(ref == 0, np.float32(0), np.abs(dif / np.where(ref == 0, np.float32(1), ref)))
        finite = np.isfinite(target)
        nicnt = int((~finite).sum())
        errcnt = int(((dif != 0) & finite).sum())
        rmax = np.float32(0)
        for r in rerr:
            if rmax < r:
                rmax = r
        return float(rmax), errcnt, nicnt
    target = np.asarray(target, dtype=np.float64).ravel()
    ref = np.asarray(ref, dtype=np.float64).ravel()
    with np.errstate(divide="ignore", invalid="ignore"):
        dif = target - ref
        rerr = np.where(ref == 0.0, 0.0, np.abs(dif / np.where(ref == 0.0, 1.0, ref)))
    finite = np.isfinite(target)
    nicnt = int((~finite).sum())
    errcnt = int(((dif != 0) & finite).sum())
    rmax = 0.0
    for r in rerr:  # "if (rerrmax < rerr) rerrmax = rerr" in order (NaN never replaces)
        if rmax < r:
            rmax = float(r)
    return rmax, errcnt, nicnt


def _sizes(th, dims):
    cur = {d: (th[d + "_const"] or th["start"]) for d in dims}
    while True:
        if any(th[d + "_const"] == 0 and cur[d] > th["stop"] for d in dims):
            return
        if all(cur[d] != 0 for d in dims):
            yield dict(cur)
        for d in dims:
            if th[d + "_const"] == 0:
                cur[d] = cur[d] * 2 if th["step_mode"] == 2 else cur[d] + th["step"]
        if all(th[d + "_const"] for d in dims):
            return


def _gen(prec):
    return O.s_init_mat if prec == "s" else O.init_mat


def run_gemm_sweep(opts, impl, max_size=None, prec="d"):
    th = harness_opts(opts)
    stop = th["stop"]
    if th["tranA"] in "Nn":
        rda, cda = (th["m_const"] or stop), (th["k_const"] or stop)
    else:
        rda, cda = (th["k_const"] or stop), (th["m_const"] or stop)
    if th["tranB"] in "Nn":
        rdb, cdb = (th["k_const"] or stop), (th["n_const"] or stop)
    else:
        rdb, cdb = (th["n_const"] or stop), (th["k_const"] or stop)
    rdc, cdc = (th["m_const"] or stop), (th["n_const"] or stop)
    lda, ldb, ldc = rda, rdb, rdc
    O.srand(1)  # rand() is never seeded by the harness
    A = _gen(prec)(rda, cda, lda, th["phi"], 3, 1)
    B = _gen(prec)(rdb, cdb, ldb, th["phi"], 3, 2)
    Cm = np.zeros(ldc * cdc, dtype=A.dtype)
    out = []
    for d in _sizes(th, ("m", "n", "k")):
        m, n, k = d["m"], d["n"], d["k"]
        if max_size is not None and max(m, n, k) > max_size:
            break
        impl.gemm(th["tranA"], th["tranB"], m, n, k, 1.0, A, lda, B, ldb, 0.0, Cm, ldc)
        got = Cm.reshape(cdc, ldc).T[:m, :n]
        Am = A.reshape(cda, lda).T
        Bm = B.reshape(cdb, ldb).T
        ref = O.exact_gemm_rn(th["tranA"], th["tranB"], m, n, k, A.astype(np.float64), lda,
                              B.astype(np.float64), ldb)
        out.append((m, n, k) + check(got, ref, prec))
        del Am, Bm
    return out


def run_dot_sweep(opts, impl, max_size=None, prec="d"):
    th = harness_opts(opts)
    nh = th["n_const"] or th["stop"]
    O.srand(1)
    X = _gen(prec)(nh, 1, 0, th["phi"], 3, 1)
    Y = _gen(prec)(nh, 1, 0, th["phi"], 3, 2)
    out = []
    for d in _sizes(th, ("n",)):
        n = d["n"]
        if max_size is not None and n > max_size:
            break
        r = impl.dot(n, X[:n].copy(), Y[:n].copy())
        ref = O.exact_dot_rn(X[:n].astype(np.float64), Y[:n].astype(np.float64))
        out.append((None, n, None) + check([r], [ref], prec))
    return out


def run_gemv_sweep(opts, impl, max_size=None, prec="d"):
    th = harness_opts(opts)
    stop = th["stop"]
    rda, cda, vlx, vly = stop, stop, stop, stop
    if th["tranA"] in "Nn":
        vlx = th["n_const"] or vlx
        vly = th["m_const"] or vly
    else:
        vlx = th["m_const"] or vlx
        vly = th["n_const"] or vly
    rda = th["m_const"] or rda
    cda = th["n_const"] or cda
    lda = rda
    O.srand(1)
    A = _gen(prec)(rda, cda, lda, th["phi"], 3, 1)
    X = _gen(prec)(vlx, 1, 0, th["phi"], 3, 2)
    Y = np.zeros(vly, dtype=A.dtype)
    out = []
    for d in _sizes(th, ("m", "n")):
        m, n = d["m"], d["n"]
        if max_size is not None and max(m, n) > max_size:
            break
        lx, ly = (n, m) if th["tranA"] in "Nn" else (m, n)
        impl.gemv(th["tranA"], m, n, 1.0, A, lda, X, 0.0, Y)
        A64, X64 = A.astype(np.float64), X.astype(np.float64)
        if th["tranA"] in "Nn":
            ref = O.exact_gemm_rn("N", "N", m, 1, n, A64, lda, X64, max(n, 1))
        else:
            ref = O.exact_gemm_rn("T", "N", n, 1, m, A64, lda, X64, max(m, 1))
        out.append((m, n, None) + check(Y[:ly], ref, prec))
    return out


RUNNERS = {"gemm": run_gemm_sweep, "dot": run_dot_sweep, "gemv": run_gemv_sweep}


def compare(case, rows, limit=None):
    """List of mismatches between a sweep result and the reference log rows."""
    bad = []
    exp_rows = case["rows"][: len(rows)] if limit is None else case["rows"][:limit]
    for got, exp in zip(rows, exp_rows):
        rmax, nwrong, nicnt = got[3], got[4], got[5]
        want = float.fromhex(exp["relerr_hex"])
        if nicnt or rmax != want or nwrong != exp["nwrong"]:
            bad.append({"got": got, "want": exp})
    return bad


class OracleImpl:
    """Adapter: the oracle restatement driven with one persistent handle (as the harness does)."""

    def __init__(self, th_flags, work_bytes=1 << 32):
        self.h = O.Handle(work_bytes, **th_flags)
        # the harness workspace (WORK_MEM_SIZE, testing/testing_setting.h:8) decides blocking
        self.h.set(workSizeBytes=int(16e9) if work_bytes >= (1 << 32) else work_bytes)

    def gemm(self, *a):
        return self.h.gemm(*a)

    def dot(self, n, x, y):
        return self.h.dot(n, x, y)

    def gemv(self, *a):
        return self.h.gemv(*a)


class OracleImplS(OracleImpl):
    """The binary32 restatement with the same harness conventions."""

    def gemm(self, *a):
        return self.h.sgemm(*a)

    def dot(self, n, x, y):
        return self.h.sdot(n, x, y)

    def gemv(self, *a):
        return self.h.sgemv(*a)


class OracleImplDS(OracleImpl):
    """The mixed <double,float,float> restatement with the same harness conventions."""

    def gemm(self, *a):
        return self.h.dsgemm(*a)

    def dot(self, n, x, y):
        return self.h.dsdot(n, x, y)

    def gemv(self, *a):
        return self.h.dsgemv(*a)
