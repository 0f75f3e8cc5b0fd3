"""CPU checks of the residue form's arithmetic (ozblas_amd/csrc/oz_residue.hip) against exact
integers:
  - the host constants the kernels read (moduli, (M/p)^-1 mod p, the limbs of M and M/p, 256^a
    mod p), taken from the library itself, against Python integers;
  - the element residues (res_group: (+-f mod P) (2^sh mod P) mod P in binary64 for the product P of
    three moduli, left in [-P, 2P); res_member: the balanced residue of that value modulo each member
    from one exact binary64 quotient),
    restated in IEEE arithmetic (numpy), against (+-f 2^sh) mod p;
  - the element decomposition (res_elem: x 2^-lo = sf 2^sh) against exact rationals;
  - the op(A) group tables with the factors u_t = (M/p_t)^-1 mod p_t folded in (pow2gU);
  - the reconstruction (res_group_values: c_t = X u_t mod p_t, K = rint(sum c_t / p_t) in
    binary64, X = sum c_t M/p_t - K M in binary64 chunks of 39 bits, each sum checked exact),
    restated the same way, against X itself for X at 0, +-1, +-M/4 and random values in between:
    the claim that |X| <= M/4 makes K exact.
  - the MFMA epilogues' acc mod p (an exact binary64 quotient, no correction) for |acc| <= 2^30.
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
CHUNK = 39
MAXCHUNK = (64 * 9 + CHUNK - 1) // CHUNK


def _dtype():
    return np.dtype([("nmod", "<i4"), ("nl", "<i4"), ("p", "<i4", MAXMOD), ("invpf", "<f4", MAXMOD),
                     ("invpd", "<f8", MAXMOD), ("u", "<i4", MAXMOD), ("Mp", "<u4", (MAXMOD, MAXLIMB)),
                     ("M", "<u4", MAXLIMB), ("pow256", "u1", (MAXMOD, MAXSH // 8)),
                     ("Pg", "<f8", NGROUP), ("invPg", "<f8", NGROUP), ("pow2g", "<u4", (NGROUP, MAXSH)),
                     ("pow2gU", "<u4", (NGROUP, MAXSH)),
                     ("Mpd", "<f8", (MAXMOD, MAXCHUNK)), ("Md", "<f8", MAXCHUNK)],
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
        assert [int(v) for v in c["Md"]] == [(M >> (CHUNK * q)) & ((1 << CHUNK) - 1) for q in range(MAXCHUNK)]
        for t in range(nmod):
            p = ps[t]
            assert int(c["p"][t]) == p
            assert [int(w) for w in c["Mp"][t]] == limbs(M // p, MAXLIMB)
            assert [int(v) for v in c["Mpd"][t]] == [((M // p) >> (CHUNK * q)) & ((1 << CHUNK) - 1)
                                                     for q in range(MAXCHUNK)]
            assert (int(c["u"][t]) * (M // p)) % p == 1
            assert c["invpd"][t] == np.float64(1.0) / np.float64(p)
            assert c["invpf"][t] == np.float32(1.0) / np.float32(p)
            assert [int(v) for v in c["pow256"][t]] == [pow(256, a, p) for a in range(MAXSH // 8)]


def fma_d(a, b, c):
    from fractions import Fraction as Fr
    return float(Fr(float(a)) * Fr(float(b)) + Fr(float(c)))


def res_elem(x, lo):
    """res_elem restated: x 2^-lo = sf 2^sh, sh the exponent of x's last place relative to lo."""
    eb = (np.float64(x).view(np.uint64) >> np.uint64(52)) & np.uint64(0x7FF)
    sh = max(max(int(eb), 1) - 1075 - lo, 0)
    return float(np.ldexp(np.float64(x), -(lo + sh))), sh


def res_group(sf, sh, P, invP, pw2):
    """res_group restated in binary64 (numpy rounds each product like the device; the fmas are exact):
    sf mod P left in [-P, 2P), times 2^sh mod P, reduced again to [-P, 2P)."""
    sf = np.float64(sf)
    q = np.floor(sf * invP)
    r0 = fma_d(-q, P, sf)  # one rounding of an integer below 2^25: exact
    assert -P <= r0 < 2 * P and r0 == int(sf) - int(q) * int(P)
    x = np.float64(r0 * float(pw2[sh]))
    assert abs(float(x)) < 2.0 ** 49 and float(x) == r0 * float(pw2[sh])
    q = np.floor(x * invP)
    r = fma_d(-q, P, x)
    assert -P <= r < 2 * P and r == int(x) - int(q) * int(P)
    return int(r)


def res_member(rx, p, invpd):
    """res_member restated: q = floor(RN(rx invp + 1/2)) through exact rationals, r = rx - q p."""
    assert abs(rx) < 2 ** 25
    q = math.floor(fma_d(rx, invpd, 0.5))
    return rx - q * p


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
    # the op(A) tables: 2^sh u_t modulo each member p_t < nmod, 2^sh modulo the members beyond
    for nmod in (1, 2, 3, 7, 20, 23, 34, MAXMOD):
        c, _ = rc[nmod]
        for g in range(NGROUP):
            P = math.prod(ps[3 * g: 3 * g + 3])
            for sh in (0, 1, 7, 100, MAXSH - 1):
                v = int(c["pow2gU"][g][sh])
                assert v < P
                for t in range(3 * g, min(3 * g + 3, MAXMOD)):
                    ut = int(c["u"][t]) if t < nmod else 1
                    assert v % ps[t] == (pow(2, sh, ps[t]) * ut) % ps[t], (nmod, g, t, sh)


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
                rx = res_group((-1 if neg else 1) * f, sh, P, c["invPg"][g], c["pow2g"][g])
                assert rx % int(P) == v % int(P), (P, f, sh, neg)
                for t in range(3 * g, min(3 * g + 3, MAXMOD)):
                    p = int(c["p"][t])
                    got = res_member(rx, p, c["invpd"][t])
                    want = balanced((-1 if neg else 1) * f * 2 ** sh, p)
                    assert got == want, (p, f, sh, neg)
                    assert -128 <= got <= 127


def test_member_residue_at_the_half_points(rc):
    """res_member at the representatives next to the balanced range's ends (rx = k p +- (p-1)/2,
    +-(p+1)/2, and the ties k 256 + 128 of the even modulus), over |rx| < 2^25."""
    c, _ = rc[MAXMOD]
    rng = random.Random(23)
    for t in range(MAXMOD):
        p = int(c["p"][t])
        ks = [0, 1, -1, (2 ** 25 - p) // p, -((2 ** 25 - p) // p)] + [rng.randrange(-(2 ** 25) // p + 1, (2 ** 25) // p - 1) for _ in range(40)]
        for k in ks:
            for d in (-(p + 1) // 2, -(p // 2), -(p - 1) // 2, 0, (p - 1) // 2, p // 2, (p + 1) // 2):
                rx = k * p + d
                if abs(rx) >= 2 ** 25:
                    continue
                assert res_member(rx, p, c["invpd"][t]) == balanced(rx, p), (p, rx)


def test_element_decomposition_against_integers():
    """res_elem: x 2^-lo == sf 2^sh exactly, sf an integer below 2^53, 0 <= sh <= the shift of x's
    lowest set bit -- for normal, subnormal and zero x and every lo up to x's lowest set bit."""
    rng = np.random.default_rng(17)
    xs = [0.0, -0.0, 1.0, -3.0, 2.0 ** -1074, -(2.0 ** -1074) * 12345, 2.0 ** -1022, 1.5 * 2.0 ** 1000]
    xs += list(rng.standard_normal(200) * np.exp2(rng.integers(-1060, 1000, 200)))
    for x in xs:
        if x == 0.0:
            lows = [0, -1074, 5]
        else:
            m, e = math.frexp(abs(x))  # x = m 2^e
            mi = int(m * 2 ** 53)
            low = e - 53 + ((mi & -mi).bit_length() - 1)  # exponent of the lowest set bit
            lows = [low, low - 1, low - 60, max(low - 500, -1074), -1074]
        for lo in lows:
            sf, sh = res_elem(x, lo)
            assert sf == int(sf) and abs(sf) < 2.0 ** 53 and sh >= 0
            from fractions import Fraction as Fr
            assert Fr(x) * Fr(2) ** (-lo) == Fr(sf) * Fr(2) ** sh, (x, lo)
            if x != 0.0:
                assert sh <= low - lo


def reconstruct(c, nmod, X):
    """res_group_values restated: c_t = X u_t mod p_t, the binary64 chunk sums of c_t M/p_t (each
    checked exact), K = rint of the fma sum of c_t / p_t, the chunks minus K M's, carried in binary64
    (each carry checked exact) so that all chunks but the top one lie in [0, 2^39)."""
    ps = [int(v) for v in c["p"][:nmod]]
    S = 0.0
    A = [0.0] * MAXCHUNK
    for t in range(nmod):
        ct = (X * int(c["u"][t])) % ps[t]  # the GEMM of op(A)'s u-scaled residues stores X u mod p
        assert ct < 256
        S = fma_d(ct, c["invpd"][t], S)
        for q in range(MAXCHUNK):
            A[q] = fma_d(ct, c["Mpd"][t][q], A[q])
            assert A[q] == int(A[q]) and A[q] < 2.0 ** 53
    K = float(np.rint(S))
    nc = MAXCHUNK
    while nc > 1 and c["Md"][nc - 1] == 0.0 and all(c["Mpd"][t][nc - 1] == 0.0 for t in range(nmod)):
        nc -= 1
    nc += 1  # one chunk above M's: it takes the sign and the last carry
    nc = min(nc, MAXCHUNK)
    d = [fma_d(-K, c["Md"][q], A[q]) for q in range(nc)]
    for q in range(nc - 1):  # the carries in binary64, each step checked exact
        assert d[q] == int(d[q]) and abs(d[q]) < 2.0 ** 53
        hi = math.floor(d[q] * 2.0 ** -CHUNK)
        r = fma_d(-hi, 2.0 ** CHUNK, d[q])
        assert 0 <= r < 2.0 ** CHUNK and r == d[q] - hi * 2.0 ** CHUNK
        nxt = d[q + 1] + hi
        assert nxt == int(d[q + 1]) + hi and abs(nxt) < 2.0 ** 53
        d[q], d[q + 1] = r, nxt
    assert abs(d[nc - 1]) < 2.0 ** 51  # the top chunk's integer conversion
    out = 0
    for q in range(nc):  # disjoint bit ranges: the sum is the OR of the chunks
        out += int(d[q]) << (CHUNK * q)
    return out


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


def mod_f64(v, p, invpd):
    """mod_f64 of the MFMA epilogues restated: q = floor(RN(v invp + 2^-10)) (one rounding, through
    exact rationals), r = v - q p (exact); checked to be the canonical residue without a correction."""
    from fractions import Fraction as Fr
    t = float(Fr(int(v)) * Fr(float(invpd)) + Fr(1, 1024))  # fma: one rounding of the exact value
    q = math.floor(t)
    r = v - q * p
    assert 0 <= r < p, (v, p, r)
    return r


def test_epilogue_mod_against_integers(rc):
    """acc mod p in the MFMA epilogues for |acc| <= 2^30 (65,536 k x 2^14 per chunk), every modulus
    (256 included): the ends of the range, multiples of p and their neighbours (where the quotient's
    rounding matters), and random values."""
    c, _ = rc[MAXMOD]
    rng = random.Random(11)
    top = 2 ** 30
    vs = [0, 1, -1, top, -top, top - 1, -top + 1, 2 ** 24, 2 ** 24 + 1, -(2 ** 24) - 3, 2 ** 27, -(2 ** 27)]
    vs += [rng.randrange(-top, top + 1) for _ in range(300)]
    for t in range(MAXMOD):
        p = int(c["p"][t])
        assert c["invpd"][t] == np.float64(1.0) / np.float64(p)
        near = [q * p + d for q in (rng.randrange(-top // p, top // p) for _ in range(60)) for d in (-1, 0, 1)]
        near += [q * p + d for q in (top // p, -(top // p), 1, -1) for d in (-1, 0, 1)]
        for v in vs + [x for x in near if abs(x) <= top]:
            assert mod_f64(v, p, c["invpd"][t]) == v % p, (p, v)


def mod_f(v, p, invpf):
    """mod_f restated (0 <= v < 2^24): the same quotient, corrected either way."""
    q = int(np.floor(np.float32(np.float32(v) * invpf)))
    r = v - q * p
    assert -p <= r < 2 * p, (v, p, r)
    r = r + p if r < 0 else r
    return r - p if r >= p else r


def test_reconstruction_chunk_sum_mod_against_integers(rc):
    """The reduction of the k-chunk sums (up to 255 x 32768 < 2^24) modulo every p, random and near
    multiples of p."""
    c, _ = rc[MAXMOD]
    rng = random.Random(19)
    for t in range(MAXMOD):
        p, ipf = int(c["p"][t]), c["invpf"][t]
        vs = [rng.randrange(0, 1 << 24) for _ in range(300)] + [(1 << 24) - 1, 255 * 32768]
        vs += [q * p + d for q in (rng.randrange(0, (1 << 24) // p) for _ in range(50)) for d in (-1, 0, 1)]
        for v in vs:
            if 0 <= v < (1 << 24):
                assert mod_f(v, p, ipf) == v % p, (p, v)
