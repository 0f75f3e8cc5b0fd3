"""Differentiable Ozaki GEMM (config 5): torch.autograd.Function whose forward and backward are
Ozaki DGEMMs (backward in the FP16-slice form).

Checks: gradcheck against FP64 finite differences; gradients bit-reproducible run to run and equal
across slice forms; sampled gradient entries equal RN(exact) from the oracle, and their error
against a double-double reference gradient is within half an ulp."""
import numpy as np
import pytest

from gpu_util import say, torch_cuda

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mods():
    torch = torch_cuda()
    from ozblas_amd import autograd as AG
    return torch, AG


def test_gradcheck_small(mods):
    torch, AG = mods
    g = torch.Generator(device="cpu").manual_seed(0)
    A = torch.randn(7, 5, dtype=torch.float64, generator=g).cuda().requires_grad_()
    B = torch.randn(5, 6, dtype=torch.float64, generator=g).cuda().requires_grad_()
    assert torch.autograd.gradcheck(lambda a, b: AG.ozaki_matmul(a, b, "int8", "fp16"), (A, B))


def test_forward_matches_exact(mods, oracle):
    torch, AG = mods
    A = oracle.init_mat(33, 21, 33, 1.0, 2, 5).reshape(21, 33).T.copy()   # 33 x 21
    B = oracle.init_mat(21, 17, 21, 1.0, 2, 6).reshape(17, 21).T.copy()   # 21 x 17
    C = AG.ozaki_mm(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), "int8").cpu().numpy()
    Af = np.asfortranarray(A).T.ravel()
    Bf = np.asfortranarray(B).T.ravel()
    ex = oracle.exact_gemm_rn("N", "N", 33, 17, 21, Af.copy(), 33, Bf.copy(), 21)
    assert C.tobytes() == np.ascontiguousarray(ex).tobytes()


def test_config5_gradients(mods, oracle):
    """A 2048x1024, B 1024x2048 (phi-mode, seeds 5/6), loss = sum(C * W), W ~ N(0,1)."""
    torch, AG = mods
    m, k, n = 2048, 1024, 2048
    A0 = oracle.init_mat(m, k, m, 1.0, 2, 5).reshape(k, m).T.copy()
    B0 = oracle.init_mat(k, n, k, 1.0, 2, 6).reshape(n, k).T.copy()
    W0 = np.random.default_rng(7).standard_normal((m, n))
    grads = []
    for rep, form in enumerate(("fp16", "fp16", "int8")):
        say(f"config5 backward form={form} rep={rep}")
        A = torch.from_numpy(A0).cuda().requires_grad_()
        B = torch.from_numpy(B0).cuda().requires_grad_()
        W = torch.from_numpy(W0).cuda()
        C = AG.ozaki_matmul(A, B, "int8", form)
        (C * W).sum().backward()
        grads.append((A.grad.cpu().numpy(), B.grad.cpu().numpy()))
    # bit-reproducible run to run and across slice forms
    for g in grads[1:]:
        assert g[0].tobytes() == grads[0][0].tobytes()
        assert g[1].tobytes() == grads[0][1].tobytes()
    dA, dB = grads[0]
    # sampled entries against RN(exact) and a double-double reference gradient
    rng = np.random.default_rng(1)
    rows = rng.integers(0, m, 64)
    cols = rng.integers(0, k, 64)
    Wf = np.asfortranarray(W0).T.ravel().copy()     # m x n column-major
    Bf = np.asfortranarray(B0).T.ravel().copy()     # k x n column-major
    exact_dA = oracle.exact_gemm_rn_entries("N", "T", n, Wf, m, Bf, k, rows, cols)
    assert dA[rows, cols].tobytes() == exact_dA.tobytes()
    worst = 0.0
    for r, c, e in zip(rows, cols, exact_dA):
        hi, lo = oracle.dd_dot(W0[r, :], B0[c, :])
        worst = max(worst, abs((e - hi) - lo) / abs(hi))
    assert worst <= 2.0 ** -53
    rows2 = rng.integers(0, k, 64)
    cols2 = rng.integers(0, n, 64)
    Af = np.asfortranarray(A0).T.ravel().copy()     # m x k column-major
    exact_dB = oracle.exact_gemm_rn_entries("T", "N", m, Af, m, Wf, m, rows2, cols2)
    assert dB[rows2, cols2].tobytes() == exact_dB.tobytes()
