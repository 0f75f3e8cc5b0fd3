"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import this
module.  It loads ``oracle/build/libozoracle.so`` (built by ``oracle/Makefile``), a plain-C
restatement of the reference OzBLAS CPU path plus independent exact and double-double checkers.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "libozoracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class OzoHandle(C.Structure):
    _fields_ = [
        ("nSplitMax", C.c_int32), ("splitMode", C.c_int32), ("fastMode", C.c_int32),
        ("reproMode", C.c_int32), ("sumMode", C.c_int32), ("useBatchedGemmFlag", C.c_int32),
        ("splitEpsMode", C.c_int32), ("precxFlag", C.c_int32),
        ("workSizeBytes", C.c_uint64), ("memAddr", C.c_uint64), ("nSplitA_", C.c_int32),
        ("splitShift", C.c_int32), ("memMaskSplitA", C.c_uint64),
        ("work", C.c_void_p), ("workCap", C.c_uint64),
        ("nSplitA", C.c_float), ("nSplitB", C.c_float), ("nSplitC", C.c_float),
        ("mbk", C.c_int32), ("nbk", C.c_int32),
        ("trueresFlag", C.c_int32), ("verbose", C.c_int32), ("cg_numiter", C.c_int32),
        ("cg_verbose1", C.c_void_p), ("cg_verbose2", C.c_void_p),
        ("fastModeRule", C.c_int32), ("dot2Chunks", C.c_int32),
    ]


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        H = C.POINTER(OzoHandle)
        L.ozo_handle_new.restype = H
        L.ozo_handle_new.argtypes = [C.c_uint64]
        L.ozo_set_cblas_dgemm.restype = None
        L.ozo_set_cblas_dgemm.argtypes = [C.c_void_p]
        L.ozo_handle_free.argtypes = [H]
        dp = C.POINTER(C.c_double)
        L.ozo_dd_dot2i.argtypes = [C.c_double, C.c_double, dp, dp]
        L.ozo_dd_quad_add.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, dp, dp]
        L.ozo_dd_quad_mul.argtypes = [C.c_double, C.c_double, C.c_double, dp, dp]
        L.ozo_dot2_chunked.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_int32, dp, dp]
        L.ozo_rho.restype = C.c_int32
        L.ozo_rho.argtypes = [C.c_int32, C.c_int32]
        L.ozo_tau.restype = C.c_int32
        L.ozo_tau.argtypes = [C.c_double]
        L.ozo_tau_batch.argtypes = [C.c_int64, _dp, np.ctypeslib.ndpointer(dtype=np.int32,
                                                                         flags="C_CONTIGUOUS")]
        L.ozo_pitch.restype = C.c_int32
        L.ozo_pitch.argtypes = [C.c_int32]
        L.ozo_nearsum.restype = C.c_double
        L.ozo_nearsum.argtypes = [C.c_int32, _dp, C.c_int32]
        L.ozo_split.restype = C.c_int32
        L.ozo_split.argtypes = [H, C.c_int32, C.c_int32, _dp, C.c_int32, _dp, C.c_int32, _dp,
                                C.c_int32, _dp]
        L.ozo_gemm.restype = C.c_int32
        L.ozo_gemm.argtypes = [H, C.c_char, C.c_char, C.c_int32, C.c_int32, C.c_int32, C.c_double,
                               _dp, C.c_int32, _dp, C.c_int32, C.c_double, _dp, C.c_int32]
        L.ozo_dot.restype = C.c_double
        L.ozo_dot.argtypes = [H, C.c_int32, _dp, _dp]
        L.ozo_nrm2.restype = C.c_double
        L.ozo_nrm2.argtypes = [H, C.c_int32, _dp]
        L.ozo_gemv.restype = C.c_int32
        L.ozo_gemv.argtypes = [H, C.c_char, C.c_int32, C.c_int32, C.c_double, _dp, C.c_int32, _dp,
                               C.c_double, _dp]
        L.ozo_axpy.restype = C.c_int32
        L.ozo_axpy.argtypes = [H, C.c_int32, C.c_double, _dp, _dp]
        L.ozo_csrmv.restype = C.c_int32
        L.ozo_csrmv.argtypes = [H, C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_void_p, _ip,
                                _ip, _dp, C.c_double, _dp]
        L.ozo_csrmv_splitA.restype = C.c_void_p
        L.ozo_csrmv_splitA.argtypes = [H, C.c_int32, C.c_int32, C.c_int32, _dp, _ip]
        L.ozo_cg.restype = C.c_int32
        L.ozo_cg.argtypes = [H, C.c_int32, C.c_int32, _dp, _ip, _ip, _dp, _dp, C.c_int32,
                             C.c_double]
        L.ozo_exact_dot_rn.restype = C.c_double
        L.ozo_exact_dot_rn.argtypes = [C.c_int64, _dp, C.c_int64, _dp, C.c_int64]
        L.ozo_exact_sum_rn.restype = C.c_double
        L.ozo_exact_sum_rn.argtypes = [C.c_int64, _dp]
        L.ozo_exact_gemm_rn.argtypes = [C.c_char, C.c_char, C.c_int32, C.c_int32, C.c_int32, _dp,
                                        C.c_int32, _dp, C.c_int32, _dp, C.c_int32, C.c_int32]
        L.ozo_exact_gemm_rn_entries.argtypes = [C.c_char, C.c_char, C.c_int32, _dp, C.c_int32, _dp,
                                                C.c_int32, C.c_int64, _ip, _ip, _dp]
        L.ozo_dd_dot.argtypes = [C.c_int64, _dp, C.c_int64, _dp, C.c_int64, C.POINTER(C.c_double),
                                 C.POINTER(C.c_double)]
        L.ozo_srand48.argtypes = [C.c_int64]
        L.ozo_drand48.restype = C.c_double
        L.ozo_srand.argtypes = [C.c_uint32]
        L.ozo_rand.restype = C.c_int32
        L.ozo_init_mat.argtypes = [C.c_int32, C.c_int32, C.c_int32, _dp, C.c_double, C.c_int32,
                                   C.c_int32]
        _fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
        L.ozo_s_rho.restype = C.c_int32
        L.ozo_s_rho.argtypes = [C.c_int32, C.c_int32]
        L.ozo_s_tau.restype = C.c_int32
        L.ozo_s_tau.argtypes = [C.c_float]
        L.ozo_s_nearsum.restype = C.c_float
        L.ozo_s_nearsum.argtypes = [C.c_int32, _fp, C.c_int32]
        L.ozo_s_split.restype = C.c_int32
        L.ozo_s_split.argtypes = [H, C.c_int32, C.c_int32, _fp, C.c_int32, _fp, C.c_int32, _fp,
                                  C.c_int32, _fp]
        L.ozo_s_gemm.restype = C.c_int32
        L.ozo_s_gemm.argtypes = [H, C.c_char, C.c_char, C.c_int32, C.c_int32, C.c_int32, C.c_float,
                                 _fp, C.c_int32, _fp, C.c_int32, C.c_float, _fp, C.c_int32]
        L.ozo_s_dot.restype = C.c_float
        L.ozo_s_dot.argtypes = [H, C.c_int32, _fp, _fp]
        L.ozo_s_nrm2.restype = C.c_float
        L.ozo_s_nrm2.argtypes = [H, C.c_int32, _fp]
        L.ozo_s_gemv.restype = C.c_int32
        L.ozo_s_gemv.argtypes = [H, C.c_char, C.c_int32, C.c_int32, C.c_float, _fp, C.c_int32, _fp,
                                 C.c_float, _fp]
        L.ozo_s_axpy.restype = C.c_int32
        L.ozo_s_axpy.argtypes = [H, C.c_int32, C.c_float, _fp, _fp]
        L.ozo_s_csrmv.restype = C.c_int32
        L.ozo_s_csrmv.argtypes = [H, C.c_int32, C.c_int32, C.c_int32, C.c_float, C.c_void_p, _ip,
                                  _ip, _fp, C.c_float, _fp]
        L.ozo_s_csrmv_splitA.restype = C.c_void_p
        L.ozo_s_csrmv_splitA.argtypes = [H, C.c_int32, C.c_int32, C.c_int32, _fp, _ip]
        L.ozo_s_cg.restype = C.c_int32
        L.ozo_s_cg.argtypes = [H, C.c_int32, C.c_int32, _fp, _ip, _ip, _fp, _fp, C.c_int32,
                               C.c_float]
        L.ozo_s_init_mat.argtypes = [C.c_int32, C.c_int32, C.c_int32, _fp, C.c_double, C.c_int32,
                                     C.c_int32]
        _sp = np.ctypeslib.ndpointer(dtype=np.int16, flags="C_CONTIGUOUS")
        L.ozo_ds_rho.restype = C.c_int32
        L.ozo_ds_rho.argtypes = [C.c_int32, C.c_int32]
        L.ozo_ds_split.restype = C.c_int32
        L.ozo_ds_split.argtypes = [H, C.c_int32, C.c_int32, _dp, C.c_int32, _dp, C.c_int32, _fp,
                                   C.c_int32, _sp, C.c_int32, _dp]
        L.ozo_ds_gemm.restype = C.c_int32
        L.ozo_ds_gemm.argtypes = [H, C.c_char, C.c_char, C.c_int32, C.c_int32, C.c_int32, C.c_double,
                                  _dp, C.c_int32, _dp, C.c_int32, C.c_double, _dp, C.c_int32]
        L.ozo_ds_dot.restype = C.c_double
        L.ozo_ds_dot.argtypes = [H, C.c_int32, _dp, _dp]
        L.ozo_ds_nrm2.restype = C.c_double
        L.ozo_ds_nrm2.argtypes = [H, C.c_int32, _dp]
        L.ozo_ds_gemv.restype = C.c_int32
        L.ozo_ds_gemv.argtypes = [H, C.c_char, C.c_int32, C.c_int32, C.c_double, _dp, C.c_int32, _dp,
                                  C.c_double, _dp]
        L.ozo_ds_csrmv.restype = C.c_int32
        L.ozo_ds_csrmv.argtypes = [H, C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_void_p, _ip,
                                   _ip, _dp, C.c_double, _dp]
        L.ozo_ds_csrmv_splitA.restype = C.c_void_p
        L.ozo_ds_csrmv_splitA.argtypes = [H, C.c_int32, C.c_int32, C.c_int32, _dp, _ip]
        L.ozo_ds_cg.restype = C.c_int32
        L.ozo_ds_cg.argtypes = [H, C.c_int32, C.c_int32, _dp, _ip, _ip, _dp, _dp, C.c_int32,
                                C.c_double]
        L.ozo_sd_rho.restype = C.c_int32
        L.ozo_sd_rho.argtypes = [C.c_int32, C.c_int32]
        L.ozo_sd_split.restype = C.c_int32
        L.ozo_sd_split.argtypes = [H, C.c_int32, C.c_int32, _fp, C.c_int32, _fp, C.c_int32, _dp,
                                   C.c_int32, _sp, C.c_int32, _fp]
        L.ozo_sd_gemm.restype = C.c_int32
        L.ozo_sd_gemm.argtypes = [H, C.c_char, C.c_char, C.c_int32, C.c_int32, C.c_int32, C.c_float,
                                  _fp, C.c_int32, _fp, C.c_int32, C.c_float, _fp, C.c_int32]
        L.ozo_sd_dot.restype = C.c_float
        L.ozo_sd_dot.argtypes = [H, C.c_int32, _fp, _fp]
        L.ozo_sd_nrm2.restype = C.c_float
        L.ozo_sd_nrm2.argtypes = [H, C.c_int32, _fp]
        L.ozo_sd_gemv.restype = C.c_int32
        L.ozo_sd_gemv.argtypes = [H, C.c_char, C.c_int32, C.c_int32, C.c_float, _fp, C.c_int32, _fp,
                                  C.c_float, _fp]
        L.ozo_sd_csrmv.restype = C.c_int32
        L.ozo_sd_csrmv.argtypes = [H, C.c_int32, C.c_int32, C.c_int32, C.c_float, C.c_void_p, _ip,
                                   _ip, _fp, C.c_float, _fp]
        L.ozo_sd_csrmv_splitA.restype = C.c_void_p
        L.ozo_sd_csrmv_splitA.argtypes = [H, C.c_int32, C.c_int32, C.c_int32, _fp, _ip]
        L.ozo_sd_cg.restype = C.c_int32
        L.ozo_sd_cg.argtypes = [H, C.c_int32, C.c_int32, _fp, _ip, _ip, _fp, _fp, C.c_int32,
                                C.c_float]
        L.ozo_laplace3d_nnz.restype = C.c_int64
        L.ozo_laplace3d_nnz.argtypes = [C.c_int32]
        L.ozo_laplace3d.argtypes = [C.c_int32, C.c_double, _ip, _ip, _dp]
        _lib = L
    return _lib


def use_openblas(enable=True):
    """Slice products of the binary64 oracle through OpenBLAS's cblas_dgemm (the LP64 build that
    scipy bundles, entry scipy_cblas_dgemm), as the reference CPU OzBLAS does them; False restores
    the built-in loop.  Returns the library path, or None when it is not found (the loop stays)."""
    import glob
    if not enable:
        lib().ozo_set_cblas_dgemm(None)
        return None
    try:
        import scipy
    except ImportError:
        return None
    libs = glob.glob(os.path.join(os.path.dirname(os.path.dirname(scipy.__file__)), "scipy.libs",
                                  "libscipy_openblas-*.so"))
    if not libs:
        return None
    global _openblas
    _openblas = C.CDLL(libs[0])
    fn = _openblas.scipy_cblas_dgemm
    lib().ozo_set_cblas_dgemm(C.cast(fn, C.c_void_p))
    return libs[0]


_openblas = None


class Handle:
    """Oracle handle with the reference's defaults (src/cpu/ozblas.cpp:22-61)."""

    def __init__(self, work_bytes: int = 1 << 31, **flags):
        self._h = lib().ozo_handle_new(int(work_bytes))
        for k, v in flags.items():
            setattr(self._h.contents, k, v)

    def __getattr__(self, name):
        return getattr(self._h.contents, name)

    def set(self, **flags):
        for k, v in flags.items():
            setattr(self._h.contents, k, v)
        return self

    def close(self):
        if self._h is not None:
            lib().ozo_handle_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # routines -----------------------------------------------------------------------------
    def gemm(self, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc):
        rc = lib().ozo_gemm(self._h, transA.encode(), transB.encode(), m, n, k, alpha, A, lda, B,
                            ldb, beta, C, ldc)
        if rc != 0:
            raise RuntimeError(f"ozo_gemm failed ({rc})")
        return C

    def dot(self, n, x, y):
        return lib().ozo_dot(self._h, n, x, y)

    def nrm2(self, n, x):
        return lib().ozo_nrm2(self._h, n, x)

    def gemv(self, trans, m, n, alpha, A, lda, x, beta, y):
        rc = lib().ozo_gemv(self._h, trans.encode(), m, n, alpha, A, lda, x, beta, y)
        if rc != 0:
            raise RuntimeError(f"ozo_gemv failed ({rc})")
        return y

    def axpy(self, n, alpha, x, y):
        lib().ozo_axpy(self._h, n, alpha, x, y)
        return y

    def csrmv(self, m, n, nnz, alpha, val, colind, rowptr, x, beta, y):
        ptr = val if isinstance(val, int) else val.ctypes.data
        rc = lib().ozo_csrmv(self._h, m, n, nnz, alpha, ptr, colind, rowptr, x, beta, y)
        if rc != 0:
            raise RuntimeError(f"ozo_csrmv failed ({rc})")
        return y

    def csrmv_splitA(self, m, n, nnz, val, rowptr):
        return lib().ozo_csrmv_splitA(self._h, m, n, nnz, val, rowptr)

    def cg(self, n, nnz, val, colind, rowptr, b, x, maxiter, tol, history=True):
        hist = np.zeros(maxiter + 2, dtype=np.float64)
        if history:
            self._h.contents.verbose = 1
            self._h.contents.cg_verbose1 = hist.ctypes.data
        rc = lib().ozo_cg(self._h, n, nnz, val, colind, rowptr, b, x, maxiter, tol)
        if rc != 0:
            raise RuntimeError(f"ozo_cg failed ({rc})")
        self._h.contents.cg_verbose1 = None
        return hist[: self._h.contents.cg_numiter + 1]


    # binary32 routines (the <float,float,float> instantiations, ozoracle_s.c) ---------------
    def sgemm(self, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc):
        rc = lib().ozo_s_gemm(self._h, transA.encode(), transB.encode(), m, n, k, alpha, A, lda, B,
                              ldb, beta, C, ldc)
        if rc != 0:
            raise RuntimeError(f"ozo_s_gemm failed ({rc})")
        return C

    def sdot(self, n, x, y):
        return np.float32(lib().ozo_s_dot(self._h, n, x, y))

    def snrm2(self, n, x):
        return np.float32(lib().ozo_s_nrm2(self._h, n, x))

    def sgemv(self, trans, m, n, alpha, A, lda, x, beta, y):
        rc = lib().ozo_s_gemv(self._h, trans.encode(), m, n, alpha, A, lda, x, beta, y)
        if rc != 0:
            raise RuntimeError(f"ozo_s_gemv failed ({rc})")
        return y

    def saxpy(self, n, alpha, x, y):
        lib().ozo_s_axpy(self._h, n, alpha, x, y)
        return y

    def scsrmv(self, m, n, nnz, alpha, val, colind, rowptr, x, beta, y):
        ptr = val if isinstance(val, int) else val.ctypes.data
        rc = lib().ozo_s_csrmv(self._h, m, n, nnz, alpha, ptr, colind, rowptr, x, beta, y)
        if rc != 0:
            raise RuntimeError(f"ozo_s_csrmv failed ({rc})")
        return y

    def scsrmv_splitA(self, m, n, nnz, val, rowptr):
        return lib().ozo_s_csrmv_splitA(self._h, m, n, nnz, val, rowptr)

    def scg(self, n, nnz, val, colind, rowptr, b, x, maxiter, tol, history=True):
        hist = np.zeros(maxiter + 2, dtype=np.float64)
        if history:
            self._h.contents.verbose = 1
            self._h.contents.cg_verbose1 = hist.ctypes.data
        rc = lib().ozo_s_cg(self._h, n, nnz, val, colind, rowptr, b, x, maxiter, tol)
        if rc != 0:
            raise RuntimeError(f"ozo_s_cg failed ({rc})")
        self._h.contents.cg_verbose1 = None
        return hist[: self._h.contents.cg_numiter + 1].astype(np.float32)


    # mixed routines (the <double,float,float> instantiations, ozoracle_ds.c) -----------------
    def dsgemm(self, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc):
        rc = lib().ozo_ds_gemm(self._h, transA.encode(), transB.encode(), m, n, k, alpha, A, lda, B,
                               ldb, beta, C, ldc)
        if rc != 0:
            raise RuntimeError(f"ozo_ds_gemm failed ({rc})")
        return C

    def dsdot(self, n, x, y):
        return lib().ozo_ds_dot(self._h, n, x, y)

    def dsnrm2(self, n, x):
        return lib().ozo_ds_nrm2(self._h, n, x)

    def dsgemv(self, trans, m, n, alpha, A, lda, x, beta, y):
        rc = lib().ozo_ds_gemv(self._h, trans.encode(), m, n, alpha, A, lda, x, beta, y)
        if rc != 0:
            raise RuntimeError(f"ozo_ds_gemv failed ({rc})")
        return y

    def dscsrmv(self, m, n, nnz, alpha, val, colind, rowptr, x, beta, y):
        ptr = val if isinstance(val, int) else val.ctypes.data
        rc = lib().ozo_ds_csrmv(self._h, m, n, nnz, alpha, ptr, colind, rowptr, x, beta, y)
        if rc != 0:
            raise RuntimeError(f"ozo_ds_csrmv failed ({rc})")
        return y

    def dscsrmv_splitA(self, m, n, nnz, val, rowptr):
        return lib().ozo_ds_csrmv_splitA(self._h, m, n, nnz, val, rowptr)

    def dscg(self, n, nnz, val, colind, rowptr, b, x, maxiter, tol, history=True):
        hist = np.zeros(maxiter + 2, dtype=np.float64)
        if history:
            self._h.contents.verbose = 1
            self._h.contents.cg_verbose1 = hist.ctypes.data
        rc = lib().ozo_ds_cg(self._h, n, nnz, val, colind, rowptr, b, x, maxiter, tol)
        if rc != 0:
            raise RuntimeError(f"ozo_ds_cg failed ({rc})")
        self._h.contents.cg_verbose1 = None
        return hist[: self._h.contents.cg_numiter + 1]


    # mixed routines (the <float,double,double> instantiations, ozoracle_sd.c) -----------------
    def sdgemm(self, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc):
        rc = lib().ozo_sd_gemm(self._h, transA.encode(), transB.encode(), m, n, k, alpha, A, lda, B,
                               ldb, beta, C, ldc)
        if rc != 0:
            raise RuntimeError(f"ozo_sd_gemm failed ({rc})")
        return C

    def sddot(self, n, x, y):
        return np.float32(lib().ozo_sd_dot(self._h, n, x, y))

    def sdnrm2(self, n, x):
        return np.float32(lib().ozo_sd_nrm2(self._h, n, x))

    def sdgemv(self, trans, m, n, alpha, A, lda, x, beta, y):
        rc = lib().ozo_sd_gemv(self._h, trans.encode(), m, n, alpha, A, lda, x, beta, y)
        if rc != 0:
            raise RuntimeError(f"ozo_sd_gemv failed ({rc})")
        return y

    def sdcsrmv(self, m, n, nnz, alpha, val, colind, rowptr, x, beta, y):
        ptr = val if isinstance(val, int) else val.ctypes.data
        rc = lib().ozo_sd_csrmv(self._h, m, n, nnz, alpha, ptr, colind, rowptr, x, beta, y)
        if rc != 0:
            raise RuntimeError(f"ozo_sd_csrmv failed ({rc})")
        return y

    def sdcsrmv_splitA(self, m, n, nnz, val, rowptr):
        return lib().ozo_sd_csrmv_splitA(self._h, m, n, nnz, val, rowptr)

    def sdcg(self, n, nnz, val, colind, rowptr, b, x, maxiter, tol, history=True):
        hist = np.zeros(maxiter + 2, dtype=np.float64)
        if history:
            self._h.contents.verbose = 1
            self._h.contents.cg_verbose1 = hist.ctypes.data
        rc = lib().ozo_sd_cg(self._h, n, nnz, val, colind, rowptr, b, x, maxiter, tol)
        if rc != 0:
            raise RuntimeError(f"ozo_sd_cg failed ({rc})")
        self._h.contents.cg_verbose1 = None
        return hist[: self._h.contents.cg_numiter + 1].astype(np.float32)


# building blocks -------------------------------------------------------------------------------
def ds_rho(dim, mode=0):
    return lib().ozo_ds_rho(dim, mode)


def sd_rho(dim, mode=0):
    return lib().ozo_sd_rho(dim, mode)


def s_rho(dim, mode=0):
    return lib().ozo_s_rho(dim, mode)


def s_tau(x):
    return lib().ozo_s_tau(float(np.float32(x)))


def s_nearsum(v):
    v = np.ascontiguousarray(v, dtype=np.float32).copy()
    return np.float32(lib().ozo_s_nearsum(len(v), v, 1))


def s_init_mat(m, n, ld, phi, mode, seed):
    """mublasInitMat of the binary32 harness: column-major buffer of ld*n floats."""
    ld = max(ld, m)
    buf = np.zeros(ld * n, dtype=np.float32)
    lib().ozo_s_init_mat(m, n, ld, buf, float(phi), mode, seed)
    return buf


def rho(dim, mode=0):
    return lib().ozo_rho(dim, mode)


def tau(x):
    return lib().ozo_tau(float(x))


def tau_batch(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(len(x), dtype=np.int32)
    lib().ozo_tau_batch(len(x), x, out)
    return out


def pitch(n):
    return lib().ozo_pitch(n)


def nearsum(v):
    v = np.ascontiguousarray(v, dtype=np.float64).copy()
    return lib().ozo_nearsum(len(v), v, 1)


def exact_sum_rn(v):
    v = np.ascontiguousarray(v, dtype=np.float64)
    return lib().ozo_exact_sum_rn(len(v), v)


def exact_dot_rn(x, y):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    return lib().ozo_exact_dot_rn(len(x), x, 1, y, 1)


def dd_dot(x, y):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    hi, lo = C.c_double(), C.c_double()
    lib().ozo_dd_dot(len(x), x, 1, y, 1, C.byref(hi), C.byref(lo))
    return hi.value, lo.value


def exact_gemm_rn(transA, transB, m, n, k, A, lda, B, ldb):
    """RN(exact op(A) op(B)) as an m x n column-major (Fortran-order) array."""
    Cm = np.zeros(max(m, 1) * n, dtype=np.float64)
    lib().ozo_exact_gemm_rn(transA.encode(), transB.encode(), m, n, k, _flat(A), lda, _flat(B),
                            ldb, Cm, max(m, 1), os.cpu_count() or 1)
    return Cm.reshape(n, max(m, 1)).T


def exact_gemm_rn_entries(transA, transB, k, A, lda, B, ldb, rows, cols):
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    cols = np.ascontiguousarray(cols, dtype=np.int32)
    out = np.zeros(len(rows), dtype=np.float64)
    lib().ozo_exact_gemm_rn_entries(transA.encode(), transB.encode(), k, _flat(A), lda, _flat(B),
                                    ldb, len(rows), rows, cols, out)
    return out


def _flat(M):
    """Column-major flat view of a 2-D array (Fortran order) or a 1-D buffer."""
    if M.ndim == 1:
        return np.ascontiguousarray(M, dtype=np.float64)
    return np.ascontiguousarray(M.T, dtype=np.float64).reshape(-1)


# generators ------------------------------------------------------------------------------------
def srand48(seed):
    lib().ozo_srand48(seed)


def drand48():
    return lib().ozo_drand48()


def srand(seed):
    lib().ozo_srand(seed)


def rand():
    return lib().ozo_rand()


def init_mat(m, n, ld, phi, mode, seed):
    """mublasInitMat (testing/testing_common.cpp:664-806): column-major buffer of ld*n doubles."""
    ld = max(ld, m)
    buf = np.zeros(ld * n, dtype=np.float64)
    lib().ozo_init_mat(m, n, ld, buf, float(phi), mode, seed)
    return buf


def laplace3d(g, shift):
    n = g ** 3
    nnz = lib().ozo_laplace3d_nnz(g)
    rowptr = np.zeros(n + 1, dtype=np.int32)
    colind = np.zeros(nnz, dtype=np.int32)
    val = np.zeros(nnz, dtype=np.float64)
    lib().ozo_laplace3d(g, float(shift), rowptr, colind, val)
    return rowptr, colind, val


# Dot2 building blocks (splitEpsMode = 2) ------------------------------------------------------
def dd_dot2i(a, b, c=(0.0, 0.0)):
    x, y = C.c_double(c[0]), C.c_double(c[1])
    lib().ozo_dd_dot2i(float(a), float(b), C.byref(x), C.byref(y))
    return x.value, y.value


def dd_quad_add(a, b):
    x, y = C.c_double(), C.c_double()
    lib().ozo_dd_quad_add(a[0], a[1], b[0], b[1], C.byref(x), C.byref(y))
    return x.value, y.value


def dd_quad_mul(a, b):
    x, y = C.c_double(), C.c_double()
    lib().ozo_dd_quad_mul(float(a), b[0], b[1], C.byref(x), C.byref(y))
    return x.value, y.value


def dot2_chunked(x, y, T):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    hi, lo = C.c_double(), C.c_double()
    lib().ozo_dot2_chunked(len(x), x.ctypes.data, y.ctypes.data, int(T), C.byref(hi), C.byref(lo))
    return hi.value, lo.value
