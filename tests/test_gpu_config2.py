"""Ozaki DGEMM at m = n = k = 4096, phi 0.5 and 4: every fast form against the FP64-slice form.

Inputs are the reference harness's phi-mode generator (testing/testing_common.cpp:664-717), seeds
1 and 2, degree auto, NearSum.  For each phi:
  - the INT8, FP16 and BF16 forms, each with residue slices and with digit planes, are bitwise
    equal to the FP64-slice plain form over the whole of C (the plain form being the restatement of the reference CPU OzBLAS, pinned bitwise
    against the oracle at smaller sizes in test_gpu_gemm.py);
  - 512 sampled entries (plus the largest and the smallest in magnitude) equal RN of the exact dot
    product from the oracle's fixed-point accumulator;
  - the relative error of those entries against a double-double dot product (the oracle's
    ozo_dd_dot: TwoProd / TwoSum per term) is printed and bounded by 2^-53 (plus the double-double
    reference's own error), with the error of a native binary64 GEMM (torch.matmul, for scale)
    printed beside it.
"""
import numpy as np
import pytest

from gpu_util import bits_equal, dev, first_diff, host, say, torch_cuda

pytestmark = pytest.mark.gpu

N = 4096
NSAMPLE = 512


@pytest.fixture(scope="module")
def oz():
    torch_cuda()
    import ozblas_amd
    return ozblas_amd


@pytest.fixture(scope="module", params=[0.5, 4.0], ids=["phi0.5", "phi4"])
def problem(request, oz, oracle):
    phi = request.param
    A = oracle.init_mat(N, N, N, phi, 2, 1)
    B = oracle.init_mat(N, N, N, phi, 2, 2)
    dA, dB = dev(A), dev(B)
    H = oz.Handle(24 << 30, sumMode=1, reproMode=1, sliceType="fp64")
    dC = dev(np.zeros(N * N))
    say(f"config 2: plain FP64-slice DGEMM {N}^3 phi={phi}")
    H.dgemm("N", "N", N, N, N, 1.0, dA, N, dB, N, 0.0, dC, N)
    plain = host(dC)
    say(f"  slices {H.nSplitA:.0f} x {H.nSplitB:.0f}, {H.nSplitC:.0f} products")
    yield dict(phi=phi, A=A, B=B, dA=dA, dB=dB, plain=plain)
    H.close()


@pytest.mark.parametrize("scheme", ["residues", "digits"])
@pytest.mark.parametrize("slice_type", ["int8", "fp16", "bf16"])
def test_fast_form_bitwise_equal_to_plain(oz, problem, slice_type, scheme):
    phi = problem["phi"]
    H = oz.Handle(24 << 30, sumMode=1, reproMode=1, sliceType=slice_type, digitScheme=scheme)
    dC = dev(np.zeros(N * N))
    say(f"config 2: {slice_type} {scheme} DGEMM {N}^3 phi={phi}")
    H.dgemm("N", "N", N, N, N, 1.0, problem["dA"], N, problem["dB"], N, 0.0, dC, N)
    got = host(dC)
    if scheme == "digits":
        say(f"  digit planes {H.nSplitA:.0f} x {H.nSplitB:.0f}, first-pass pairs {H.nSplitC:.0f}, "
            f"second-pass tiles {H.fallbackTiles}")
    else:
        say(f"  moduli {H.nSplitC:.0f}")
    assert H.schemeUsed == (1 if scheme == "digits" else 2)
    H.close()
    assert bits_equal(got, problem["plain"]), first_diff(got, problem["plain"])


def _samples(C):
    rng = np.random.default_rng(2024)
    rows = rng.integers(0, N, NSAMPLE)
    cols = rng.integers(0, N, NSAMPLE)
    Cm = C.reshape(N, N)  # Cm[j, i] = C(i, j) (column-major)
    mag = np.abs(Cm)
    jmax, imax = np.unravel_index(np.argmax(mag), mag.shape)
    jmin, imin = np.unravel_index(np.argmin(np.where(mag > 0, mag, np.inf)), mag.shape)
    rows = np.concatenate([rows, [imax, imin]]).astype(np.int32)
    cols = np.concatenate([cols, [jmax, jmin]]).astype(np.int32)
    return rows, cols


def test_sampled_entries_exact_and_error_vs_double_double(oz, oracle, problem):
    import torch
    phi, A, B, C = problem["phi"], problem["A"], problem["B"], problem["plain"]
    rows, cols = _samples(C)
    got = C[cols.astype(np.int64) * N + rows]
    ex = oracle.exact_gemm_rn_entries("N", "N", N, A, N, B, N, rows, cols)
    assert bits_equal(got, ex), first_diff(got, ex)
    Am = A.reshape(N, N)  # Am[kk, i] = A(i, kk)
    Bm = B.reshape(N, N)  # Bm[j, kk] = B(kk, j)
    # native binary64 GEMM on the same GPU, for scale only (column-major buffers viewed transposed)
    dAm = problem["dA"].reshape(N, N).T.contiguous()   # (i, kk)
    dBm = problem["dB"].reshape(N, N).T.contiguous()   # (kk, j)
    native = torch.matmul(dAm, dBm).cpu().numpy()      # native[i, j]
    err_oz, err_nat, nz = 0.0, 0.0, 0
    for q in range(len(rows)):
        i, j = int(rows[q]), int(cols[q])
        x = np.ascontiguousarray(Am[:, i])
        y = np.ascontiguousarray(Bm[j, :])
        hi, lo = oracle.dd_dot(x, y)
        if hi == 0.0:
            continue
        nz += 1
        # relative error of a binary64 value v against the double-double hi + lo, evaluated as
        # ((v - hi) - lo) / hi: v - hi is exact when v and hi are within a factor 2 (Sterbenz)
        err_oz = max(err_oz, abs((got[q] - hi) - lo) / abs(hi))
        err_nat = max(err_nat, abs((native[i, j] - hi) - lo) / abs(hi))
    say(f"config 2 phi={phi}: max relative error vs double-double over {nz} entries: "
        f"Ozaki (every form, same bits) {err_oz:.3e}, native binary64 GEMM {err_nat:.3e}")
    assert nz >= NSAMPLE // 2
    # RN(exact) is within half an ulp, 2^-53 relative; the double-double reference adds its own
    # error, far below that for these inputs
    assert err_oz <= 2.0 ** -53 * (1 + 2.0 ** -20)
