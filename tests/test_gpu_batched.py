"""GPU parity of the grouped batched Ozaki DGEMM (ozblasDgemmBatched, INT8 digit form).

Every C[p] must be bitwise equal to ozblasDgemm on problem p alone in the plain form (FP64 slices,
NearSum), i.e. RN of the exact product -- checked against the oracle's exact accumulator on a few
problems too.  Covers uneven sizes (multiples of no tile), every transposition, alpha/beta, an
empty and a k = 0 problem, a problem with a NaN (computed by the plain form), the replay of the
captured graph, and a new plan when new operands need more digits than the cached plan holds.
"""
import numpy as np
import pytest

from gpu_util import bits_equal, dev, first_diff, host, say, torch_cuda

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def oz():
    torch_cuda()
    import ozblas_amd
    return ozblas_amd


def _shape(trans, rows, cols):
    return (rows, cols) if trans in "Nn" else (cols, rows)


def make_batch(oz, oracle, sizes, trans, phi=2.0, seed0=1000, nan_in=None):
    probs = []
    for p, ((m, n, k), (ta, tb)) in enumerate(zip(sizes, trans)):
        ar, ac = _shape(ta, m, k)
        br, bc = _shape(tb, k, n)
        lda, ldb, ldc = max(1, ar + (p % 3)), max(1, br + (p % 2)), max(1, m + 1)
        A = oz.gen_phi(ar, ac, lda, phi, seed0 + 2 * p) if ar * ac else np.zeros(max(1, lda * ac))
        B = oz.gen_phi(br, bc, ldb, phi, seed0 + 2 * p + 1) if br * bc else np.zeros(max(1, ldb * bc))
        if nan_in == p:
            A[0] = np.nan
        alpha = 1.0 if p % 4 else 0.75
        beta = 0.0 if p % 5 else -0.5
        C0 = np.zeros(ldc * n) if beta == 0.0 else oz.gen_phi(m, n, ldc, 1.0, 7 + p)
        probs.append(dict(m=m, n=n, k=k, ta=ta, tb=tb, A=A, B=B, lda=lda, ldb=ldb, ldc=ldc,
                          alpha=alpha, beta=beta, C0=C0))
    return probs


def solo_plain(oz, probs):
    """each problem alone through ozblasDgemm, plain form (FP64 slices), NearSum"""
    H = oz.Handle(1 << 30, sumMode=1, sliceType="fp64")
    out = []
    for q in probs:
        dC = dev(q["C0"])
        H.dgemm(q["ta"], q["tb"], q["m"], q["n"], q["k"], q["alpha"], dev(q["A"]), q["lda"],
                dev(q["B"]), q["ldb"], q["beta"], dC, q["ldc"])
        out.append(host(dC))
    return out


def run_batched(H, probs, dA=None, dB=None, dC=None):
    dA = dA or [dev(q["A"]) for q in probs]
    dB = dB or [dev(q["B"]) for q in probs]
    dC = dC or [dev(q["C0"]) for q in probs]
    H.dgemm_batched([q["ta"] for q in probs], [q["tb"] for q in probs], [q["m"] for q in probs],
                    [q["n"] for q in probs], [q["k"] for q in probs], [q["alpha"] for q in probs],
                    dA, [q["lda"] for q in probs], dB, [q["ldb"] for q in probs],
                    [q["beta"] for q in probs], dC, [q["ldc"] for q in probs])
    return dA, dB, dC


def check_all(probs, dC, want):
    for p, (q, d, w) in enumerate(zip(probs, dC, want)):
        got = host(d)
        assert bits_equal(got, w), (p, q["m"], q["n"], q["k"], first_diff(got, w))


def sizes_of(seed, count, lo, hi):
    rng = np.random.default_rng(seed)
    return [tuple(int(v) for v in rng.integers(lo, hi + 1, 3)) for _ in range(count)]


def test_batched_bitwise_and_replay(oz, oracle):
    sizes = sizes_of(4, 22, 1, 300) + [(0, 5, 7), (9, 11, 0), (256, 256, 256)]
    trans = [("NN", "NT", "TN", "TT")[p % 4] for p in range(len(sizes))]
    probs = make_batch(oz, oracle, sizes, trans, nan_in=3)
    say(f"batched: {len(probs)} problems, sizes {sizes[:4]}...")
    want = solo_plain(oz, probs)
    H = oz.Handle(1 << 30, sumMode=1, sliceType="int8")
    dA, dB, dC = run_batched(H, probs)
    assert H.batchReplay == 0 and H.batchTile in (128, 256)
    assert H.batchGrouped == len(probs) - 3  # empty, k = 0 and the NaN problem go the plain way
    check_all(probs, dC, want)
    # the same call again replays the captured graph
    for q, d in zip(probs, dC):
        d.copy_(dev(q["C0"]))
    say("batched: replay")
    run_batched(H, probs, dA, dB, dC)
    assert H.batchReplay == 1
    check_all(probs, dC, want)
    # a few problems against the oracle's exact accumulator (RN of the exact product)
    for p in (0, 7, 24):
        q = probs[p]
        if q["beta"] != 0.0 or q["alpha"] != 1.0:
            continue
        ex = oracle.exact_gemm_rn(q["ta"], q["tb"], q["m"], q["n"], q["k"], q["A"].copy(), q["lda"],
                                  q["B"].copy(), q["ldb"])
        got = host(dC[p]).reshape(q["n"], q["ldc"])[:, :q["m"]].ravel()
        assert bits_equal(got, np.ascontiguousarray(np.asarray(ex).T).ravel())


def test_batched_new_operands_replan(oz, oracle):
    sizes = sizes_of(11, 10, 16, 200)
    trans = ["NN"] * len(sizes)
    probs = make_batch(oz, oracle, sizes, trans, phi=1.0)
    H = oz.Handle(1 << 30, sumMode=1, sliceType="int8")
    dA, dB, dC = run_batched(H, probs)
    check_all(probs, dC, solo_plain(oz, probs))
    # same buffers, new values that need more digits (a wider exponent span): the cached plan's
    # check flags it and the call builds a new plan
    for p, q in enumerate(probs):
        q["A"] = q["A"] * np.where(np.arange(len(q["A"])) % 7 == 3, 2.0 ** -60, 1.0)
        dA[p].copy_(dev(q["A"]))
        dC[p].copy_(dev(q["C0"]))
    say("batched: new operands")
    run_batched(H, probs, dA, dB, dC)
    assert H.batchReplay == 0
    check_all(probs, dC, solo_plain(oz, probs))
    # and values that fit the plan again replay it
    for p, q in enumerate(probs):
        dC[p].copy_(dev(q["C0"]))
    run_batched(H, probs, dA, dB, dC)
    assert H.batchReplay == 1
    check_all(probs, dC, solo_plain(oz, probs))


def test_batched_other_forms_loop(oz, oracle):
    """The FP64 slice form (Fsum and NearSum) runs problem by problem, as ozblasDgemm; a form that
    ozblasDgemm refuses (digit slices with Fsum) is refused by the batch too."""
    sizes = sizes_of(5, 5, 8, 120)
    probs = make_batch(oz, oracle, sizes, ["NT"] * 5)
    with pytest.raises(oz.OzblasError):
        run_batched(oz.Handle(1 << 30, sumMode=0, sliceType="int8"), probs)
    for flags in (dict(sumMode=0, sliceType="fp64"), dict(sumMode=1, sliceType="fp64")):
        H1 = oz.Handle(1 << 30, **flags)
        _, _, dC = run_batched(H1, probs)
        H2 = oz.Handle(1 << 30, **flags)
        for q, d in zip(probs, dC):
            dc = dev(q["C0"])
            H2.dgemm(q["ta"], q["tb"], q["m"], q["n"], q["k"], q["alpha"], dev(q["A"]), q["lda"],
                     dev(q["B"]), q["ldb"], q["beta"], dc, q["ldc"])
            assert bits_equal(host(d), host(dc))


def test_batched_config4_full_size(oz, oracle):
    """The full batched workload: 256 problems with m, n, k = 64 + floor(961 drand48()) after
    srand48(4), phi = 2 inputs (seeds 1000 + 2p, 1001 + 2p), alpha = 1, beta = 0, in one grouped
    launch; every C[p] bitwise equal to its solo plain-form (FP64 slices, NearSum) ozblasDgemm, and
    the three smallest problems equal to RN of the exact product."""
    oracle.srand48(4)
    sizes = [tuple(64 + int(961 * oracle.drand48()) for _ in range(3)) for _ in range(256)]
    probs = []
    for p, (m, n, k) in enumerate(sizes):
        probs.append(dict(m=m, n=n, k=k, ta="N", tb="N", A=oz.gen_phi(m, k, m, 2.0, 1000 + 2 * p),
                          B=oz.gen_phi(k, n, k, 2.0, 1001 + 2 * p), lda=m, ldb=k, ldc=m, alpha=1.0,
                          beta=0.0, C0=np.zeros(m * n)))
    say(f"batched config 4: {len(probs)} problems, sum 2mnk = {sum(2.0 * m * n * k for m, n, k in sizes):.3e}")
    want = solo_plain(oz, probs)
    H = oz.Handle(1 << 30, sumMode=1, sliceType="int8")
    dA, dB, dC = run_batched(H, probs)
    assert H.batchGrouped == len(probs)
    check_all(probs, dC, want)
    say("batched config 4: replay")
    for q, d in zip(probs, dC):
        d.zero_()
    run_batched(H, probs, dA, dB, dC)
    assert H.batchReplay == 1
    check_all(probs, dC, want)
    for p in sorted(range(len(sizes)), key=lambda q: sizes[q][0] * sizes[q][1] * sizes[q][2])[:3]:
        q = probs[p]
        ex = oracle.exact_gemm_rn("N", "N", q["m"], q["n"], q["k"], q["A"], q["m"], q["B"], q["k"])
        got = host(dC[p]).reshape(q["n"], q["m"]).T
        assert bits_equal(got, ex), (p, first_diff(got, ex))
