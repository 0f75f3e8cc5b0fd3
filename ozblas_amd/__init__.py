"""ozblas_amd -- MI355X (gfx950) Ozaki-scheme BLAS behind the OzBLAS handle and C API.

The compute path is ``libozblas_mi355x.so`` (hand-written HIP kernels, C ABI declared in
``include/ozblas_mi355x.h``).  This module is a thin ctypes front end that mirrors the reference
operator interface (``ozblasRgemm`` / ``cuozblasRgemm`` ... in ``include/cuozblas.h``): a handle
with the reference's option fields, and one call per routine taking device buffers (torch tensors
on ``cuda:0`` or raw device addresses).  There is no fallback: if the library is missing the
import fails.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import sys
import weakref

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libozblas_mi355x.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"ozblas_amd: {LIB_PATH} is not built (run __graft_entry__.build())")

_lib = C.CDLL(LIB_PATH)

SUCCESS, NOT_INITIALIZED, ALLOC_FAILED, INVALID_VALUE, NOT_SUPPORTED, WORKSPACE, EXEC_FAILED = range(7)
STATUS_NAMES = {0: "SUCCESS", 1: "NOT_INITIALIZED", 2: "ALLOC_FAILED", 3: "INVALID_VALUE",
                4: "NOT_SUPPORTED", 5: "WORKSPACE", 6: "EXECUTION_FAILED"}
SLICE_FP64, SLICE_INT8, SLICE_FP16, SLICE_BF16 = 0, 1, 2, 3
SLICE_NAMES = {"fp64": SLICE_FP64, "int8": SLICE_INT8, "fp16": SLICE_FP16, "bf16": SLICE_BF16}
SCHEME_AUTO, SCHEME_DIGITS, SCHEME_RESIDUES = 0, 1, 2
SCHEME_NAMES = {"auto": SCHEME_AUTO, "digits": SCHEME_DIGITS, "residues": SCHEME_RESIDUES}


class OzblasHandleT(C.Structure):
    """ozblasHandle_t (include/ozblas_mi355x.h), field for field."""
    _fields_ = [
        ("stream", C.c_void_p), ("backend", C.c_void_p),
        ("devWork", C.c_void_p), ("devWorkCommon", C.c_void_p), ("workSizeBytes", C.c_uint64),
        ("memAddr", C.c_uint64), ("nSplitMax", C.c_int32), ("hstBatchAddr", C.c_void_p),
        ("devBatchAddr", C.c_void_p), ("initialized", C.c_int32),
        ("splitMode", C.c_int32), ("fastMode", C.c_int32), ("reproMode", C.c_int32),
        ("sumMode", C.c_int32), ("useBatchedGemmFlag", C.c_int32), ("splitEpsMode", C.c_int32),
        ("precxFlag", C.c_int32),
        ("nSplitA", C.c_float), ("nSplitB", C.c_float), ("nSplitC", C.c_float),
        ("t_SplitA", C.c_float), ("t_SplitB", C.c_float), ("t_comp", C.c_float),
        ("t_sum", C.c_float), ("t_total", C.c_float), ("n_comp", C.c_float),
        ("mbk", C.c_int32), ("nbk", C.c_int32),
        ("nSplitA_", C.c_int32), ("splitShift", C.c_int32), ("memMaskSplitA", C.c_uint64),
        ("trueresFlag", C.c_int32), ("verbose", C.c_int32), ("cg_numiter", C.c_int32),
        ("cg_verbose1", C.c_void_p), ("cg_verbose2", C.c_void_p),
        ("t_SplitMat_total", C.c_double), ("t_SplitVec_total", C.c_double),
        ("t_Sum_total", C.c_double), ("t_AXPY_SCAL_total", C.c_double),
        ("t_DOT_NRM2_total", C.c_double), ("t_SpMV_SpMM_total", C.c_double),
        ("sliceType", C.c_int32), ("lastStatus", C.c_int32), ("nLevelBins", C.c_int32),
        ("batchTile", C.c_int32), ("batchReplay", C.c_int32), ("batchGrouped", C.c_int32),
        ("digitLevels", C.c_int32), ("fallbackTiles", C.c_int32),
        ("cgDeviceSlices", C.c_int32), ("cgDevicePath", C.c_int32),
        ("digitScheme", C.c_int32), ("schemeUsed", C.c_int32),
    ]


_H = C.POINTER(OzblasHandleT)
_p, _i, _d, _c = C.c_void_p, C.c_int32, C.c_double, C.c_char


def _sig(name, res, *args):
    f = getattr(_lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


_create = _sig("ozblasCreate", _i, _H, C.c_uint64)
_destroy = _sig("ozblasDestroy", _i, _H)
_setstream = _sig("ozblasSetStream", _i, _H, _p)
_dgemm = _sig("ozblasDgemm", _i, _H, _c, _c, _i, _i, _i, _d, _p, _i, _p, _i, _d, _p, _i)
_dgemv = _sig("ozblasDgemv", _i, _H, _c, _i, _i, _d, _p, _i, _p, _i, _d, _p, _i)
_ddot = _sig("ozblasDdot", _i, _H, _i, _p, _i, _p, _i, C.POINTER(C.c_double))
_dnrm2 = _sig("ozblasDnrm2", _i, _H, _i, _p, _i, C.POINTER(C.c_double))
_daxpy = _sig("ozblasDaxpy", _i, _H, _i, _d, _p, _i, _p, _i)
_dcsrmv = _sig("ozblasDcsrmv", _i, _H, _c, _i, _i, _i, _d, _p, _p, _p, _p, _p, _d, _p)
_dcsrmv_split = _sig("ozblasDcsrmvSplitA", _p, _H, _c, _i, _i, _i, _p, _p, _p)
_dcg = _sig("ozblasDcg", _i, _H, _c, _i, _i, _p, _p, _p, _p, _p, _p, _i, _d)
_dgemm_batched = _sig("ozblasDgemmBatched", _i, _H, _i, C.c_char_p, C.c_char_p,
                      C.POINTER(_i), C.POINTER(_i), C.POINTER(_i), C.POINTER(_d),
                      C.POINTER(_p), C.POINTER(_i), C.POINTER(_p), C.POINTER(_i),
                      C.POINTER(_d), C.POINTER(_p), C.POINTER(_i))
_f = C.c_float
_sgemm = _sig("ozblasSgemm", _i, _H, _c, _c, _i, _i, _i, _f, _p, _i, _p, _i, _f, _p, _i)
_sgemv = _sig("ozblasSgemv", _i, _H, _c, _i, _i, _f, _p, _i, _p, _i, _f, _p, _i)
_sdot = _sig("ozblasSdot", _i, _H, _i, _p, _i, _p, _i, C.POINTER(C.c_float))
_snrm2 = _sig("ozblasSnrm2", _i, _H, _i, _p, _i, C.POINTER(C.c_float))
_saxpy = _sig("ozblasSaxpy", _i, _H, _i, _f, _p, _i, _p, _i)
_scsrmv = _sig("ozblasScsrmv", _i, _H, _c, _i, _i, _i, _f, _p, _p, _p, _p, _p, _f, _p)
_scsrmv_split = _sig("ozblasScsrmvSplitA", _p, _H, _c, _i, _i, _i, _p, _p, _p)
_scg = _sig("ozblasScg", _i, _H, _c, _i, _i, _p, _p, _p, _p, _p, _p, _i, _f)
_dsgemm = _sig("ozblasDSgemm", _i, _H, _c, _c, _i, _i, _i, _d, _p, _i, _p, _i, _d, _p, _i)
_dsgemv = _sig("ozblasDSgemv", _i, _H, _c, _i, _i, _d, _p, _i, _p, _i, _d, _p, _i)
_dsdot = _sig("ozblasDSdot", _i, _H, _i, _p, _i, _p, _i, C.POINTER(C.c_double))
_dsnrm2 = _sig("ozblasDSnrm2", _i, _H, _i, _p, _i, C.POINTER(C.c_double))
_dscsrmv = _sig("ozblasDScsrmv", _i, _H, _c, _i, _i, _i, _d, _p, _p, _p, _p, _p, _d, _p)
_dscsrmv_split = _sig("ozblasDScsrmvSplitA", _p, _H, _c, _i, _i, _i, _p, _p, _p)
_dscg = _sig("ozblasDScg", _i, _H, _c, _i, _i, _p, _p, _p, _p, _p, _p, _i, _d)
_sdgemm = _sig("ozblasSDgemm", _i, _H, _c, _c, _i, _i, _i, _f, _p, _i, _p, _i, _f, _p, _i)
_sdgemv = _sig("ozblasSDgemv", _i, _H, _c, _i, _i, _f, _p, _i, _p, _i, _f, _p, _i)
_sddot = _sig("ozblasSDdot", _i, _H, _i, _p, _i, _p, _i, C.POINTER(C.c_float))
_sdnrm2 = _sig("ozblasSDnrm2", _i, _H, _i, _p, _i, C.POINTER(C.c_float))
_sdcsrmv = _sig("ozblasSDcsrmv", _i, _H, _c, _i, _i, _i, _f, _p, _p, _p, _p, _p, _f, _p)
_sdcsrmv_split = _sig("ozblasSDcsrmvSplitA", _p, _H, _c, _i, _i, _i, _p, _p, _p)
_sdcg = _sig("ozblasSDcg", _i, _H, _c, _i, _i, _p, _p, _p, _p, _p, _p, _i, _f)
_version = _sig("ozblasGetVersionString", C.c_char_p)
_gen_phi = _sig("ozblasGenPhi", None, _i, _i, _i, _p, _d, C.c_int64)
internal_tau = _sig("ozblas_internal_tau", _i, _d)
internal_rho = _sig("ozblas_internal_rho", _i, _i, _i)
internal_rho_row = _sig("ozblas_internal_rho_row", _i, _i)
internal_rho_row_mode = _sig("ozblas_internal_rho_row_mode", _i, _i, _i)
internal_rho_ds = _sig("ozblas_internal_rho_ds", _i, _i, _i)
internal_rho_row_ds = _sig("ozblas_internal_rho_row_ds", _i, _i, _i)
internal_rho_sd = _sig("ozblas_internal_rho_sd", _i, _i, _i)
internal_rho_row_sd = _sig("ozblas_internal_rho_row_sd", _i, _i, _i)
internal_f32_term = _sig("ozblas_internal_f32_term", C.c_float, _d)
internal_tau_f = _sig("ozblas_internal_tau_f", _i, C.c_float)
internal_nearsum_exp_f = _sig("ozblas_internal_nearsum_exp_f", _i, C.c_float)
internal_rho_f = _sig("ozblas_internal_rho_f", _i, _i, _i)
internal_rho_row_f = _sig("ozblas_internal_rho_row_f", _i, _i, _i)


class OzblasError(RuntimeError):
    def __init__(self, where, status):
        super().__init__(f"{where} failed: {STATUS_NAMES.get(status, status)}")
        self.status = status


def version() -> str:
    return _version().decode()


def _ptr(x):
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        return x.data_ptr()
    if hasattr(x, "ctypes"):
        return x.ctypes.data
    raise TypeError(f"cannot take the address of {type(x)}")


def gen_phi(m, n, ld, phi, seed):
    """The reference harness's phi-mode generator (testing/testing_common.cpp:703-717) into a new
    column-major float64 numpy buffer of ld*n elements."""
    import numpy as np
    buf = np.zeros(max(ld, m) * n, dtype=np.float64)
    _gen_phi(m, n, max(ld, m), buf.ctypes.data, float(phi), int(seed))
    return buf


# Library teardown is explicit: at interpreter exit (before the modules are cleared, and before
# torch's own exit handlers when this package is imported after torch) the registered callbacks
# run, then every live handle is destroyed.  A handle's destructor makes no HIP call once the
# interpreter is finalising: by then the streams a handle was given (torch's) may be gone.
_LIVE = weakref.WeakSet()
_TEARDOWN = []


def register_teardown(fn):
    """Runs fn() at interpreter exit, before the live handles are destroyed."""
    _TEARDOWN.append(fn)


@atexit.register
def _teardown():
    while _TEARDOWN:
        _TEARDOWN.pop()()
    for h in list(_LIVE):
        h.close()


class Handle:
    """ozblasHandle_t wrapper: ``Handle(work_bytes, sumMode=1, sliceType='int8', ...)``."""

    def __init__(self, work_bytes: int = 1 << 30, **flags):
        self._h = OzblasHandleT()
        st = _create(C.byref(self._h), int(work_bytes))
        if st != SUCCESS:
            raise OzblasError("ozblasCreate", st)
        _LIVE.add(self)
        self.set(**flags)

    def set(self, **flags):
        for k, v in flags.items():
            if k == "sliceType" and isinstance(v, str):
                v = SLICE_NAMES[v]
            if k == "digitScheme" and isinstance(v, str):
                v = SCHEME_NAMES[v]
            if not hasattr(self._h, k):
                raise AttributeError(k)
            setattr(self._h, k, v)
        return self

    def __getattr__(self, name):
        if name.startswith("_"):
            raise AttributeError(name)
        return getattr(self._h, name)

    @property
    def ref(self):
        return C.byref(self._h)

    def set_stream(self, stream_ptr):
        st = _setstream(self.ref, stream_ptr)
        if st != SUCCESS:
            raise OzblasError("ozblasSetStream", st)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.initialized:
            _destroy(self.ref)

    def __del__(self, _finalizing=sys.is_finalizing):
        # bound at definition: the module's globals (sys among them) may already be None when the
        # interpreter tears the module down
        if _finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _check(where, st):
        if st != SUCCESS:
            raise OzblasError(where, st)

    # ---- routines (device buffers; column-major; host scalars) -------------------------------
    def dgemm(self, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, Cm, ldc):
        self._check("ozblasDgemm", _dgemm(self.ref, transA.encode(), transB.encode(), m, n, k,
                                          alpha, _ptr(A), lda, _ptr(B), ldb, beta, _ptr(Cm), ldc))

    def dgemv(self, trans, m, n, alpha, A, lda, x, beta, y):
        self._check("ozblasDgemv", _dgemv(self.ref, trans.encode(), m, n, alpha, _ptr(A), lda,
                                          _ptr(x), 1, beta, _ptr(y), 1))

    def ddot(self, n, x, y) -> float:
        r = C.c_double(0.0)
        self._check("ozblasDdot", _ddot(self.ref, n, _ptr(x), 1, _ptr(y), 1, C.byref(r)))
        return r.value

    def dnrm2(self, n, x) -> float:
        r = C.c_double(0.0)
        self._check("ozblasDnrm2", _dnrm2(self.ref, n, _ptr(x), 1, C.byref(r)))
        return r.value

    def daxpy(self, n, alpha, x, y):
        self._check("ozblasDaxpy", _daxpy(self.ref, n, alpha, _ptr(x), 1, _ptr(y), 1))

    def dcsrmv(self, m, n, nnz, alpha, val, colind, rowptr, x, beta, y):
        self._check("ozblasDcsrmv", _dcsrmv(self.ref, b"N", m, n, nnz, alpha, None, _ptr(val),
                                            _ptr(colind), _ptr(rowptr), _ptr(x), beta, _ptr(y)))

    def dcsrmv_splitA(self, m, n, nnz, val, rowptr):
        p = _dcsrmv_split(self.ref, b"N", m, n, nnz, None, _ptr(val), _ptr(rowptr))
        if not p:
            raise OzblasError("ozblasDcsrmvSplitA", self._h.lastStatus)
        return p

    def _cg(self, fn, name, dtype, n, nnz, val, colind, rowptr, b, x, maxiter, tol, history,
            second):
        import numpy as np
        hist = np.zeros(maxiter + 2, dtype=dtype)
        v2 = None
        if history:
            self._h.verbose = 1
            self._h.cg_verbose1 = hist.ctypes.data
            if second is not None:  # "trueres": |b - A x| / |b| (dtype); "times": seconds (float64)
                v2 = np.zeros(maxiter + 2, dtype=dtype if second == "trueres" else np.float64)
                self._h.trueresFlag = 1 if second == "trueres" else 0
                self._h.cg_verbose2 = v2.ctypes.data
        st = fn(self.ref, b"N", n, nnz, None, _ptr(val), _ptr(colind), _ptr(rowptr), _ptr(b),
                _ptr(x), maxiter, tol)
        self._h.cg_verbose1 = None
        self._h.cg_verbose2 = None
        self._h.trueresFlag = 0
        self._check(name, st)
        k = self._h.cg_numiter + 1
        return hist[:k] if v2 is None else (hist[:k], v2[:k])

    def dcg(self, n, nnz, val, colind, rowptr, b, x, maxiter, tol, history=True, second=None):
        """Residual history resi/nrmb (cg_verbose1); with second="trueres" or "times" also the
        second history (cg_verbose2) as a tuple."""
        return self._cg(_dcg, "ozblasDcg", "float64", n, nnz, val, colind, rowptr, b, x, maxiter,
                        tol, history, second)

    # ---- binary32 (the reference's <float,float,float> instantiations) -----------------------
    def sgemm(self, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, Cm, ldc):
        self._check("ozblasSgemm", _sgemm(self.ref, transA.encode(), transB.encode(), m, n, k,
                                          alpha, _ptr(A), lda, _ptr(B), ldb, beta, _ptr(Cm), ldc))

    def sgemv(self, trans, m, n, alpha, A, lda, x, beta, y):
        self._check("ozblasSgemv", _sgemv(self.ref, trans.encode(), m, n, alpha, _ptr(A), lda,
                                          _ptr(x), 1, beta, _ptr(y), 1))

    def sdot(self, n, x, y) -> float:
        r = C.c_float(0.0)
        self._check("ozblasSdot", _sdot(self.ref, n, _ptr(x), 1, _ptr(y), 1, C.byref(r)))
        return r.value

    def snrm2(self, n, x) -> float:
        r = C.c_float(0.0)
        self._check("ozblasSnrm2", _snrm2(self.ref, n, _ptr(x), 1, C.byref(r)))
        return r.value

    def saxpy(self, n, alpha, x, y):
        self._check("ozblasSaxpy", _saxpy(self.ref, n, alpha, _ptr(x), 1, _ptr(y), 1))

    def scsrmv(self, m, n, nnz, alpha, val, colind, rowptr, x, beta, y):
        self._check("ozblasScsrmv", _scsrmv(self.ref, b"N", m, n, nnz, alpha, None, _ptr(val),
                                            _ptr(colind), _ptr(rowptr), _ptr(x), beta, _ptr(y)))

    def scsrmv_splitA(self, m, n, nnz, val, rowptr):
        p = _scsrmv_split(self.ref, b"N", m, n, nnz, None, _ptr(val), _ptr(rowptr))
        if not p:
            raise OzblasError("ozblasScsrmvSplitA", self._h.lastStatus)
        return p

    def scg(self, n, nnz, val, colind, rowptr, b, x, maxiter, tol, history=True, second=None):
        """As dcg; the histories of the binary32 solver are float arrays, as the reference's."""
        return self._cg(_scg, "ozblasScg", "float32", n, nnz, val, colind, rowptr, b, x, maxiter,
                        tol, history, second)

    # ---- mixed: binary64 data, binary32 slices (<double,float,float>) -----------------------
    def dsgemm(self, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, Cm, ldc):
        self._check("ozblasDSgemm", _dsgemm(self.ref, transA.encode(), transB.encode(), m, n, k,
                                            alpha, _ptr(A), lda, _ptr(B), ldb, beta, _ptr(Cm), ldc))

    def dsgemv(self, trans, m, n, alpha, A, lda, x, beta, y):
        self._check("ozblasDSgemv", _dsgemv(self.ref, trans.encode(), m, n, alpha, _ptr(A), lda,
                                            _ptr(x), 1, beta, _ptr(y), 1))

    def dsdot(self, n, x, y) -> float:
        r = C.c_double(0.0)
        self._check("ozblasDSdot", _dsdot(self.ref, n, _ptr(x), 1, _ptr(y), 1, C.byref(r)))
        return r.value

    def dsnrm2(self, n, x) -> float:
        r = C.c_double(0.0)
        self._check("ozblasDSnrm2", _dsnrm2(self.ref, n, _ptr(x), 1, C.byref(r)))
        return r.value

    def dscsrmv(self, m, n, nnz, alpha, val, colind, rowptr, x, beta, y):
        self._check("ozblasDScsrmv", _dscsrmv(self.ref, b"N", m, n, nnz, alpha, None, _ptr(val),
                                              _ptr(colind), _ptr(rowptr), _ptr(x), beta, _ptr(y)))

    def dscsrmv_splitA(self, m, n, nnz, val, rowptr):
        p = _dscsrmv_split(self.ref, b"N", m, n, nnz, None, _ptr(val), _ptr(rowptr))
        if not p:
            raise OzblasError("ozblasDScsrmvSplitA", self._h.lastStatus)
        return p

    def dscg(self, n, nnz, val, colind, rowptr, b, x, maxiter, tol, history=True, second=None):
        return self._cg(_dscg, "ozblasDScg", "float64", n, nnz, val, colind, rowptr, b, x, maxiter,
                        tol, history, second)

    # ---- mixed: binary32 data, binary64-held slices (<float,double,double>) ------------------
    def sdgemm(self, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, Cm, ldc):
        self._check("ozblasSDgemm", _sdgemm(self.ref, transA.encode(), transB.encode(), m, n, k,
                                            alpha, _ptr(A), lda, _ptr(B), ldb, beta, _ptr(Cm), ldc))

    def sdgemv(self, trans, m, n, alpha, A, lda, x, beta, y):
        self._check("ozblasSDgemv", _sdgemv(self.ref, trans.encode(), m, n, alpha, _ptr(A), lda,
                                            _ptr(x), 1, beta, _ptr(y), 1))

    def sddot(self, n, x, y) -> float:
        r = C.c_float(0.0)
        self._check("ozblasSDdot", _sddot(self.ref, n, _ptr(x), 1, _ptr(y), 1, C.byref(r)))
        return r.value

    def sdnrm2(self, n, x) -> float:
        r = C.c_float(0.0)
        self._check("ozblasSDnrm2", _sdnrm2(self.ref, n, _ptr(x), 1, C.byref(r)))
        return r.value

    def sdcsrmv(self, m, n, nnz, alpha, val, colind, rowptr, x, beta, y):
        self._check("ozblasSDcsrmv", _sdcsrmv(self.ref, b"N", m, n, nnz, alpha, None, _ptr(val),
                                              _ptr(colind), _ptr(rowptr), _ptr(x), beta, _ptr(y)))

    def sdcsrmv_splitA(self, m, n, nnz, val, rowptr):
        p = _sdcsrmv_split(self.ref, b"N", m, n, nnz, None, _ptr(val), _ptr(rowptr))
        if not p:
            raise OzblasError("ozblasSDcsrmvSplitA", self._h.lastStatus)
        return p

    def sdcg(self, n, nnz, val, colind, rowptr, b, x, maxiter, tol, history=True, second=None):
        return self._cg(_sdcg, "ozblasSDcg", "float32", n, nnz, val, colind, rowptr, b, x, maxiter,
                        tol, history, second)

    def dgemm_batched(self, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, Cm, ldc):
        bsz = len(m)
        arr = lambda T, v: (T * bsz)(*v)
        st = _dgemm_batched(self.ref, bsz, "".join(transA).encode(), "".join(transB).encode(),
                            arr(_i, m), arr(_i, n), arr(_i, k), arr(_d, alpha),
                            arr(_p, [_ptr(a) for a in A]), arr(_i, lda),
                            arr(_p, [_ptr(b) for b in B]), arr(_i, ldb), arr(_d, beta),
                            arr(_p, [_ptr(c) for c in Cm]), arr(_i, ldc))
        self._check("ozblasDgemmBatched", st)
