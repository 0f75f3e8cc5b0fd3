"""The committed CG fixture (tests/golden/cg_laplace100.json) against the oracle itself: its first
iterations are recomputed here (the full 500 iterations take minutes on the CPU), and the history
is checked for the properties any CG residual history of this system has."""
import json
import os

import numpy as np

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cg_laplace100.json")


def test_cg_fixture_head_matches_oracle(oracle):
    fx = json.load(open(FIX))
    g = fx["grid"]
    rowptr, colind, val = oracle.laplace3d(g, fx["shift"])
    n, nnz = g ** 3, len(val)
    assert (n, nnz) == (fx["n"], fx["nnz"]) == (10 ** 6, 7 * 10 ** 6 - 6 * 100 ** 2)
    oracle.srand48(fx["seed_b"])
    b = np.array([2.0 * oracle.drand48() - 1.0 for _ in range(n)])
    for sumMode in (1, 0):
        h = oracle.Handle(32 << 30, sumMode=sumMode)
        head = h.cg(n, nnz, val, colind, rowptr, b, np.zeros(n), 3, 0.0)
        want = [float.fromhex(v) for v in fx["cases"][str(sumMode)]["history_hex"][:4]]
        assert [v.hex() for v in head] == [v.hex() for v in want]


def test_cg_fixture_history_shape():
    fx = json.load(open(FIX))
    for sumMode in ("1", "0"):
        h = np.array([float.fromhex(v) for v in fx["cases"][sumMode]["history_hex"]])
        assert len(h) == fx["iters"] + 1 and h[0] == 1.0  # x0 = 0: r0 = b
        assert np.all(np.isfinite(h)) and h[-1] < 1e-10  # converged well within 500 iterations
    # NearSum and Fsum agree to rounding level along the whole history
    a = np.array([float.fromhex(v) for v in fx["cases"]["1"]["history_hex"]])
    b = np.array([float.fromhex(v) for v in fx["cases"]["0"]["history_hex"]])
    assert np.max(np.abs(a - b) / a) < 1e-9
