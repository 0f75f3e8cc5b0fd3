"""Golden residual history of the reproducible CG on the 3-D 7-point Laplacian (100^3 grid,
n = 10^6, diagonal shift 1e-3), b_i = 2 drand48() - 1 with srand48(3), x0 = 0, 500 iterations,
tol 0, nSplitMax 0, reproMode 1; sumMode 1 (NearSum) and 0 (Fsum).

Computed with the CPU oracle (oracle/ozoracle.c, the restatement of src/cpu/ozblas_cg.cpp and the
Ozaki DOT/NRM2/CSRMV it calls).  Stores the residual history (resi / nrmb per iteration, 501
doubles, as hex bit patterns) and a SHA-256 of the final x for each sumMode.

Usage: python tests/golden/make_cg_fixture.py [grid] [iters]   (defaults 100 500)
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))

from oracle import oracle  # noqa: E402


def main():
    g = int(sys.argv[1]) if len(sys.argv) > 1 else 100
    iters = int(sys.argv[2]) if len(sys.argv) > 2 else 500
    rowptr, colind, val = oracle.laplace3d(g, 1e-3)
    n, nnz = g ** 3, len(val)
    oracle.srand48(3)
    b = np.array([2.0 * oracle.drand48() - 1.0 for _ in range(n)])
    out = {"grid": g, "n": n, "nnz": nnz, "iters": iters, "shift": 1e-3, "seed_b": 3, "cases": {}}
    for sumMode in (1, 0):
        h = oracle.Handle(32 << 30, sumMode=sumMode)  # reserved, not touched: sized for L^2 product slots
        x = np.zeros(n)
        t0 = time.time()
        hist = h.cg(n, nnz, val, colind, rowptr, b, x, iters, 0.0)
        dt = time.time() - t0
        out["cases"][str(sumMode)] = {
            "history_hex": [float(v).hex() for v in hist],
            "x_sha256": hashlib.sha256(np.ascontiguousarray(x).tobytes()).hexdigest(),
            "splitShift": int(h.splitShift), "nSplitA_": int(h.nSplitA_),
            "cpu_seconds": round(dt, 2),
        }
        print(f"sumMode={sumMode}: {iters} iterations in {dt:.1f} s, last residual {hist[-1]!r}", flush=True)
    path = os.path.join(HERE, f"cg_laplace{g}.json")
    with open(path, "w") as f:
        json.dump(out, f)
    print("wrote", path)


if __name__ == "__main__":
    main()
