"""Differentiable Ozaki GEMM: ``torch.autograd.Function`` over the C-ABI library.

forward:  C  = A @ B            (Ozaki DGEMM, slice form ``fwd_form``)
backward: dA = dC @ B^T, dB = A^T @ dC   (Ozaki DGEMM, slice form ``bwd_form``, FP16 by default)

Every product is an Ozaki DGEMM with NearSum and automatic degree, i.e. the correctly rounded
(RN-even) exact product, so forward and backward results are bit-reproducible run to run and
independent of the slice form.  Tensors are float64 on the GPU; torch's row-major buffers are
passed to the column-major C ABI as transposed views (C^T = B^T A^T), so no copies are made.
"""
from __future__ import annotations

import torch

from . import Handle, register_teardown

_HANDLES = {}


def _close_handles():
    """Destroys the cached handles while HIP and torch are still up (run at interpreter exit)."""
    while _HANDLES:
        _, h = _HANDLES.popitem()
        h.close()


register_teardown(_close_handles)


def _handle(form: str, work_bytes: int) -> Handle:
    """The cached handle of (form, device), its work issued on torch's current stream: the stream
    is set on every call, so the work is ordered after whatever produced the inputs there."""
    key = (form, torch.cuda.current_device())
    h = _HANDLES.get(key)
    if h is None or h.workSizeBytes < work_bytes:
        if h is not None:
            h.close()
        h = Handle(work_bytes, sumMode=1, reproMode=1, nSplitMax=0, sliceType=form)
        _HANDLES[key] = h
    h.set_stream(torch.cuda.current_stream().cuda_stream)
    return h


def _work_bytes(m, n, k):
    return int(max(1 << 28, 64 * (m * k + k * n + m * n) * 8))


def ozaki_mm(A: torch.Tensor, B: torch.Tensor, form: str = "int8", transA=False,
             transB=False) -> torch.Tensor:
    """C = op(A) @ op(B) for row-major float64 CUDA tensors with an Ozaki DGEMM."""
    if A.dtype != torch.float64 or B.dtype != torch.float64 or not A.is_cuda or not B.is_cuda:
        raise TypeError("ozaki_mm needs float64 CUDA tensors")
    A = A.contiguous()
    B = B.contiguous()
    m, k = (A.shape[1], A.shape[0]) if transA else (A.shape[0], A.shape[1])
    k2, n = (B.shape[1], B.shape[0]) if transB else (B.shape[0], B.shape[1])
    if k != k2:
        raise ValueError(f"inner dimensions differ: {k} vs {k2}")
    C = torch.empty((m, n), dtype=torch.float64, device=A.device)
    h = _handle(form, _work_bytes(m, n, k))
    # column-major view: C^T (n x m) = op(B)^T (n x k) * op(A)^T (k x m)
    # a row-major R x S buffer is the column-major S x R matrix with ld = S
    ldb = B.shape[1]
    lda = A.shape[1]
    tb = "T" if transB else "N"
    ta = "T" if transA else "N"
    h.dgemm(tb, ta, n, m, k, 1.0, B, ldb, A, lda, 0.0, C, n)
    return C


class OzakiMatmul(torch.autograd.Function):
    @staticmethod
    def forward(ctx, A, B, fwd_form="int8", bwd_form="fp16"):
        ctx.save_for_backward(A, B)
        ctx.bwd_form = bwd_form
        return ozaki_mm(A, B, fwd_form)

    @staticmethod
    def backward(ctx, dC):
        A, B = ctx.saved_tensors
        dC = dC.contiguous()
        dA = dB = None
        if ctx.needs_input_grad[0]:
            dA = ozaki_mm(dC, B, ctx.bwd_form, transB=True)   # dC @ B^T
        if ctx.needs_input_grad[1]:
            dB = ozaki_mm(A, dC, ctx.bwd_form, transA=True)   # A^T @ dC
        return dA, dB, None, None


def ozaki_matmul(A, B, fwd_form="int8", bwd_form="fp16"):
    return OzakiMatmul.apply(A, B, fwd_form, bwd_form)
