"""Extract the reference's own accuracy logs into a JSON fixture (run in the build container only).

The reference ships check-mode logs of its CPU OzBLAS build under script/a100/*.dat together with
the sweep scripts that produced them (script/a100/_a100_{gemm,dot,gemv}.sh).  Each log row is
``m n k OK|NG max_rel_err (hex) n_wrong`` where n_wrong counts entries that differ from
RN(exact) (testing/testing_common.cpp:839-930, MPFR 2048-bit reference).  This script pairs every
``testing_ozblas_{d,s,ds}{gemm,dot,gemv}`` command line (binary64, binary32 and mixed builds) with its log and
stores the parsed rows -- data only, no reference source -- in tests/golden/a100_kat.json.

Usage: python tests/golden/make_kat_fixtures.py [/root/reference]
"""
import json
import os
import re
import shlex
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
SCRIPTS = ["_a100_gemm.sh", "_a100_dot.sh", "_a100_gemv.sh"]
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "a100_kat.json")


def parse_cmd(line):
    line = line.split("#")[0].strip()
    if ">" not in line:
        return None
    cmd, out = line.split(">", 1)
    toks = shlex.split(cmd)
    mt = re.search(r"testing_ozblas_(ds|d|s)(gemm|dot|gemv)$", toks[0]) if toks else None
    if mt is None:
        return None
    prec, routine = mt.group(1), mt.group(2)
    opts = {}
    for t in toks[1:]:
        if t.startswith("--") and "=" in t:
            k, v = t[2:].split("=", 1)
            opts[k] = v
    return prec, routine, opts, out.strip().split()[0]


def parse_dat(path):
    rows = []
    for ln in open(path):
        f = ln.split()
        if len(f) < 6:
            continue
        m, n, k = (None if x == "--" else int(x) for x in f[:3])
        hexv = re.search(r"\((0x[0-9a-fp+\-.]+)\)", ln).group(1)
        rows.append({"m": m, "n": n, "k": k, "status": f[3],
                     "relerr_hex": hexv, "nwrong": int(f[-1])})
    return rows


def main():
    base = os.path.join(REF, "script", "a100")
    cases = []
    seen = set()
    for sh in SCRIPTS:
        for line in open(os.path.join(base, sh)):
            p = parse_cmd(line)
            if p is None:
                continue
            prec, routine, opts, dat = p
            if dat in seen:
                continue
            path = os.path.join(base, dat)
            if not os.path.exists(path):
                continue
            seen.add(dat)
            cases.append({"prec": prec, "routine": routine, "opts": opts, "dat": dat,
                          "rows": parse_dat(path)})
    json.dump({"source": "script/a100 (OzBLAS 1.6a reference logs)", "cases": cases},
              open(OUT, "w"), indent=1)
    print(f"wrote {len(cases)} cases to {OUT}")


if __name__ == "__main__":
    main()
