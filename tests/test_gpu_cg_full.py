"""Reproducible CG at full size on the GPU: 3-D 7-point Laplacian on a 100^3 grid (n = 10^6,
nnz = 6,940,000, diagonal shift 1e-3), b_i = 2 drand48() - 1 (srand48(3)), x0 = 0, 500 iterations,
tol 0.  The residual history (resi / nrmb per iteration) and the final x must be bitwise identical
to the CPU oracle's (tests/golden/cg_laplace100.json, written by tests/golden/make_cg_fixture.py),
for NearSum and for Fsum."""
import hashlib
import json
import os
import time

import numpy as np
import pytest

from gpu_util import dev, host, say, torch_cuda

pytestmark = pytest.mark.gpu

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cg_laplace100.json")


@pytest.mark.parametrize("sumMode", [1, 0])
def test_cg_laplace100_history_bitwise(oracle, sumMode):
    torch_cuda()
    import ozblas_amd as oz
    fx = json.load(open(FIX))
    case = fx["cases"][str(sumMode)]
    g, iters = fx["grid"], fx["iters"]
    rowptr, colind, val = oracle.laplace3d(g, fx["shift"])
    n, nnz = g ** 3, len(val)
    assert (n, nnz) == (fx["n"], fx["nnz"])
    oracle.srand48(fx["seed_b"])
    b = np.array([2.0 * oracle.drand48() - 1.0 for _ in range(n)])
    H = oz.Handle(4 << 30, sumMode=sumMode)
    dx = dev(np.zeros(n))
    say(f"cg n={n} iters={iters} sumMode={sumMode}")
    t0 = time.time()
    hist = H.dcg(n, nnz, dev(val), dev(colind), dev(rowptr), dev(b), dx, iters, 0.0)
    say(f"cg {iters} iterations: {time.time() - t0:.2f} s")
    want = [float.fromhex(v) for v in case["history_hex"]]
    got = [float(v) for v in hist]
    assert len(got) == len(want)
    bad = [i for i, (u, v) in enumerate(zip(got, want)) if u.hex() != v.hex()]
    assert not bad, (bad[0], got[bad[0]].hex(), want[bad[0]].hex(), len(bad))
    assert (H.splitShift, H.nSplitA_) == (case["splitShift"], case["nSplitA_"])
    assert hashlib.sha256(np.ascontiguousarray(host(dx)).tobytes()).hexdigest() == case["x_sha256"]
