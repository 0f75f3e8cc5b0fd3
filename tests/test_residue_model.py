"""CPU checks of the residue form's arithmetic (ozblas_amd/csrc/oz_residue.hip) against exact
integers:
  - the host constants the kernels read (moduli, (M/p)^-1 mod p, the limbs of M and M/p, 256^a
    mod p), taken from the library itself, against Python integers;
  - the element residues (res_group: (+-f mod P) (2^sh mod P) mod P in binary64 for the product P of
    three moduli; res_member: the balanced residue of that value modulo each member in binary32),
    restated in IEEE arithmetic (numpy), against (+-f 2^sh) mod p;
  - the reconstruction (res_elem_value: c_t = r_t u_t mod p_t, K = rint(sum c_t / p_t) in
    binary64, X = sum c_t M/p_t - K M), restated the same way, against X itself for X at 0, +-1,
    +-M/4 and random values in between: the claim that |X| <= M/4 makes K exact.
A plausible mistake (a wrong inverse, a dropped limb, K rounded from the wrong side, a residue
taken modulo the wrong power of two) fails one of them.  No GPU call is made."""
import ctypes
import math
import os
import random

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "ozblas_amd", "libozblas_mi355x.so")
MAXMOD, MAXLIMB, MAXSH = 49, 18, 512
NGROUP = (MAXMOD + 2) // 3


def _dtype():
    return np.dtype([("nmod", "<i4"), ("nl", "<i4"), ("p", "<i4", MAXMOD), ("invpf", "<f4", MAXMOD),
                     ("invpd", "<f8", MAXMOD), ("u", "<i4", MAXMOD), ("Mp", "<u4", (MAXMOD, MAXLIMB)),
                     ("M", "<u4", MAXLIMB), ("pow256", "u1", (MAXMOD, MAXSH // 8)),
                     ("Pg", "<f8", NGROUP), ("invPg", "<f8", NGROUP), ("pow2g", "<u4", (NGROUP, MAXSH))],
                    align=True)


@pytest.fixture(scope="module")
def rc():
    if not os.path.exists(LIB):
        pytest.fail("libozblas_mi355x.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(LIB)
    f = lib.ozblas_internal_res_const
    f.restype = ctypes.c_int32
    f.argtypes = [ctypes.c_int32, ctypes.c_void_p, ctypes.c_uint64]
    lg = lib.ozblas_internal_res_log2
    lg.restype = ctypes.c_double
    lg.argtypes = [ctypes.c_int32]
    dt = _dtype()
    assert f(1, None, 0) == dt.itemsize
    out = {}
    for nmod in range(1, MAXMOD + 1):
        buf = np.zeros(1, dtype=dt)
        f(nmod, buf.ctypes.data, dt.itemsize)
        out[nmod] = (buf[0], lg(nmod))
    return out


def limbs(v, n):
    return [(v >> (32 * q)) & 0xFFFFFFFF for q in range(n)]


def test_moduli_and_constants(rc):
    c, _ = rc[MAXMOD]
    ps = [int(p) for p in c["p"][:MAXMOD]]
    assert ps[0] == 256 and all(2 < p <= 256 for p in ps)
    for a in range(len(ps)):
        for b in range(a):
            assert math.gcd(ps[a], ps[b]) == 1
    for nmod in (1, 2, 7, 20, 33, MAXMOD):
        c, lg = rc[nmod]
        assert int(c["nmod"]) == nmod
        M = math.prod(ps[:nmod])
        assert abs(lg - math.log2(M)) < 1e-9
        # X and sum_t c_t M/p_t < nmod M fit the carried limbs with a sign bit
        assert 32 * int(c["nl"]) >= (nmod * M).bit_length() + 1
        assert [int(w) for w in c["M"]] == limbs(M, MAXLIMB)
        for t in range(nmod):
            p = ps[t]
            assert int(c["p"][t]) == p
            assert [int(w) for w in c["Mp"][t]] == limbs(M // p, MAXLIMB)
            assert (int(c["u"][t]) * (M // p)) % p == 1
            assert c["invpd"][t] == np.float64(1.0) / np.float64(p)
            assert c["invpf"][t] == np.float32(1.0) / np.float32(p)
            assert [int(v) for v in c["pow256"][t]] == [pow(256, a, p) for a in range(MAXSH // 8)]


def res_group(fd, sh, neg, P, invP, pw2):
    """res_group restated in binary64 (numpy rounds each operation like the device; the fma is exact)."""
    if fd == 0:
        return 0
    fd = np.float64(fd)
    q = np.floor(fd * invP)
    r = np.float64(fd - q * P)  # fma(-q, P, f): exact, q P < 2^53
    r = r + P if r < 0 else r
    r = r - P if r >= P else r
    x = np.float64(r * np.float64(pw2[sh]))
    assert x < 2.0 ** 53
    q = np.floor(x * invP)
    r = np.float64(float(x) - float(q) * float(P))  # exact: |x - q P| < 2 P
    r = r + P if r < 0 else r
    r = r - P if r >= P else r
    if neg and r != 0:
        r = P - r
    return int(r)


def res_member(rx, p, invpf):
    """res_member restated in binary32 / int32 (rint: nearest, ties to even, as the device)."""
    q = int(np.rint(np.float32(np.float32(rx) * invpf)))
    r = rx - q * p
    hi = p >> 1 if p & 1 else (p >> 1) - 1
    lo = -(p >> 1)
    r = r - p if r > hi else r
    r = r + p if r < lo else r
    return r


def balanced(v, p):
    r = v % p
    thr = p >> 1 if p & 1 else (p >> 1) - 1
    return r - p if r > thr else r


def test_group_tables(rc):
    c, _ = rc[MAXMOD]
    ps = [int(p) for p in c["p"][:MAXMOD]]
    for g in range(NGROUP):
        P = math.prod(ps[3 * g: 3 * g + 3])
        assert P < 2 ** 24 and c["Pg"][g] == P and c["invPg"][g] == np.float64(1.0) / np.float64(P)
        assert [int(v) for v in c["pow2g"][g]] == [pow(2, sh, P) for sh in range(MAXSH)]


def test_element_residues_against_integers(rc):
    c, _ = rc[MAXMOD]
    rng = random.Random(7)
    fs = [1, 3, (1 << 53) - 1, (1 << 52) + 1, 255, 257, 65535, 1234567890123457]
    fs += [rng.randrange(1, 1 << 53) | 1 for _ in range(200)]
    shs = [0, 1, 7, 8, 9, 63, 64, 255, 256, 511] + [rng.randrange(0, MAXSH) for _ in range(30)]
    for g in range(NGROUP):
        P = float(c["Pg"][g])
        cases = [(f, sh) for f in (fs if g < 2 else fs[:30]) for sh in (shs if g < 2 else shs[:12])]
        for f, sh in cases:
            for neg in (0, 1):
                v = (-1 if neg else 1) * f * pow(2, sh, int(P))
                rx = res_group(f, sh, neg, P, c["invPg"][g], c["pow2g"][g])
                assert rx == v % int(P), (P, f, sh, neg)
                for t in range(3 * g, min(3 * g + 3, MAXMOD)):
                    p = int(c["p"][t])
                    got = res_member(rx, p, c["invpf"][t])
                    want = balanced((-1 if neg else 1) * f * 2 ** sh, p)
                    assert got == want, (p, f, sh, neg)
                    assert -128 <= got <= 127


def reconstruct(c, nmod, X):
    """res_elem_value restated: residues of X -> X, with K from the binary64 sum."""
    ps = [int(v) for v in c["p"][:nmod]]
    M = math.prod(ps)
    S = np.float64(0.0)
    acc = 0
    for t in range(nmod):
        r = X % ps[t]  # the GEMM stores X mod p in [0, p)
        ct = (r * int(c["u"][t])) % ps[t]
        S = np.float64(S + np.float64(ct) * c["invpd"][t])
        acc += ct * (M // ps[t])
    K = int(np.rint(S))
    return acc - K * M


@pytest.mark.parametrize("nmod", [2, 5, 20, 21, 34, MAXMOD])
def test_reconstruction_is_exact_within_quarter(rc, nmod):
    c, _ = rc[nmod]
    M = math.prod(int(v) for v in c["p"][:nmod])
    rng = random.Random(nmod)
    q = M // 4
    xs = [0, 1, -1, 2, -2, q, -q, q - 1, -(q - 1), M // 5, -(M // 5)]
    xs += [rng.randrange(-q, q + 1) for _ in range(400)]
    xs += [rng.randrange(-1 << 60, 1 << 60) for _ in range(100)] if M > 1 << 64 else []
    for X in xs:
        assert reconstruct(c, nmod, X) == X, (nmod, X)
    # beyond the quarter the reconstruction is not promised: the count rule keeps |X| < M / 4
    # (M >= 4 k 2^(spanA + spanB) > 4 |X|), checked here on the rule the host applies
    for need in (5.0, 60.5, 155.3, 300.0):
        n = 1
        while n <= MAXMOD and rc[n][1] < need + 1e-6:
            n += 1
        if n <= MAXMOD:
            assert rc[n][1] >= need and (n == 1 or rc[n - 1][1] < need + 1e-6)


def mod_i32(v, p, invpf):
    """mod_i32 of the MFMA epilogues restated in binary32 (numpy) and int32 arithmetic."""
    q1 = int(np.rint(np.float32(np.float32(v) * invpf)))
    r = v - q1 * p
    assert -(2 ** 31) <= q1 * p < 2 ** 31
    q2 = int(np.floor(np.float32(np.float32(r) * invpf)))
    r -= q2 * p
    r = r + p if r < 0 else r
    return r - p if r >= p else r


def test_epilogue_mod_against_integers(rc):
    """acc mod p in the MFMA epilogues for |acc| <= 2^30 (65,536 k x 2^14 per chunk), every modulus
    (256 included), at the ends of the range and near multiples of p."""
    c, _ = rc[MAXMOD]
    rng = random.Random(11)
    vs = [0, 1, -1, 2 ** 30, -(2 ** 30), 2 ** 30 - 1, -(2 ** 30) + 1, 2 ** 24, 2 ** 24 + 1, -(2 ** 24) - 3]
    vs += [rng.randrange(-(2 ** 30), 2 ** 30 + 1) for _ in range(300)]
    for t in range(MAXMOD):
        p = int(c["p"][t])
        near = [q * p + d for q in (rng.randrange(-(2 ** 30) // p, 2 ** 30 // p) for _ in range(20))
                for d in (-1, 0, 1)]
        for v in vs + [x for x in near if abs(x) <= 2 ** 30]:
            assert mod_i32(v, p, c["invpf"][t]) == v % p, (p, v)


def mod_i32_short(v, p, invpf):
    """mod_i32_short restated: one binary32 quotient floor(RN(RN(v) * invpf)), then two corrections."""
    q = int(np.floor(np.float32(np.float32(v) * invpf)))
    r = v - q * p
    assert -p <= r < 2 * p, (v, p, r)  # the bound the two corrections rely on
    r = r + p if r < 0 else r
    return r - p if r >= p else r


def test_short_chunk_epilogue_mod_against_integers(rc):
    """acc mod p with the one-step quotient for |acc| <= 2^27 (chunks of at most 8192 k): every
    modulus (29 the smallest), at the ends of the range, near multiples of p and where the binary32
    rounding of acc is largest."""
    c, _ = rc[MAXMOD]
    rng = random.Random(13)
    top = 2 ** 27
    vs = [0, 1, -1, top, -top, top - 1, -top + 1, 2 ** 24 + 1, -(2 ** 24) - 1, 2 ** 26 + 3, -(2 ** 26) - 5]
    vs += [rng.randrange(-top, top + 1) for _ in range(400)]
    for t in range(MAXMOD):
        p = int(c["p"][t])
        near = [q * p + d for q in (rng.randrange(-top // p, top // p) for _ in range(40)) for d in (-1, 0, 1)]
        for v in vs + [x for x in near if abs(x) <= top]:
            assert mod_i32_short(v, p, c["invpf"][t]) == v % p, (p, v)
