"""Pins the CPU oracle (oracle/episdet_oracle.c) without trusting it.

Checks, each chosen so that a plausible mistake fails one of them:
  * bit order and plane layout: decoding the golden planes gives back the
    genotypes the golden script packed by explicit bit arithmetic;
  * the 2^k AND-count + derived-cell method equals per-sample counting on
    random datasets (any wrong derived cell, wrong marginal or swapped index
    breaks this), and tables obey their invariants;
  * K2 equals ln prod (a+b+1)*C(a+b,a) computed on exact integers, and the
    single-non-empty-cell closed form;
  * MI equals sklearn's mutual information, 0 on independent tables and
    H(Y) on a separating table;
  * the exhaustive search agrees with a brute force over the golden sets
    (including the tie rule) and recovers a planted interaction;
  * n_choose_k equals math.comb and the README's C(4096,3).
"""
import itertools
import math

import numpy as np
import pytest

import episdet
import oracle as O


def rand_planes(rng, n, N, W=None, maf=(0.05, 0.5)):
    W = W or 32 * ((N + 1023) // 1024)
    p = rng.uniform(*maf, n)
    g = np.array([rng.choice(3, N, p=[(1 - q) ** 2, 2 * q * (1 - q), q * q]) for q in p], dtype=np.uint8)
    planes = np.zeros((n, 2, W), dtype=np.uint32)
    s = np.arange(N)
    for v in range(n):
        for pl in range(2):
            bits = (g[v] == pl)
            words = np.zeros(W, dtype=np.uint64)
            np.bitwise_or.at(words, s[bits] // 32, (1 << (31 - s[bits] % 32)).astype(np.uint64))
            planes[v, pl] = words.astype(np.uint32)
    return g, planes


def test_n_choose_k():
    for n in range(0, 200, 7):
        for k in range(0, 4):
            assert O.n_choose_k(n, k) == math.comb(n, k)
    assert O.n_choose_k(4096, 3) == 11_444_858_880          # README.md:54
    assert O.n_choose_k(8192, 3) == 91_592_417_280
    assert O.n_choose_k(512, 2) == 130_816


def test_decode_matches_golden_bit_order(golden, golden_planes):
    cases, controls, Nc, Nt = golden_planes
    assert (O.decode(cases, Nc) == np.array(golden["genotypes_cases"])).all()
    assert (O.decode(controls, Nt) == np.array(golden["genotypes_controls"])).all()
    assert O.check_planes(cases, Nc) == 0 and O.check_planes(controls, Nt) == 0


def test_check_planes_flags_overlap_and_padding(golden_planes):
    cases, _, Nc, _ = golden_planes
    bad = cases.copy()
    bad[0, 0, 0] |= 0x80000000      # sample 0 set in both planes
    bad[0, 1, 0] |= 0x80000000
    assert O.check_planes(bad, Nc) == 1
    pad = cases.copy()
    pad[1, 0, 1] |= 1          # sample 63 >= Nc = 45 lies in the padding
    assert O.check_planes(pad, Nc) == 1


@pytest.mark.parametrize("N,W", [(1, None), (31, None), (32, None), (45, None), (1000, None), (1025, None),
                                 (300, 40), (2048, None)])
def test_bits_method_equals_direct_counting(N, W):
    rng = np.random.default_rng(N * 7 + (W or 0))
    n = 9
    g, planes = rand_planes(rng, n, N, W)
    assert (O.decode(planes, N) == g).all()
    for order in (2, 3):
        for snps in itertools.combinations(range(n), order):
            direct = O.table_direct(g, N, snps)
            bits = O.table_bits(planes, N, snps)
            assert (direct == bits).all(), (snps, direct, bits)
            assert int(bits.sum()) == N


def test_table_marginals():
    rng = np.random.default_rng(5)
    g, planes = rand_planes(rng, 6, 777)
    for x, y, z in itertools.combinations(range(6), 3):
        t = O.table_bits(planes, 777, (x, y, z)).reshape(3, 3, 3).astype(np.int64)
        assert (t.sum(axis=2).ravel() == O.table_bits(planes, 777, (x, y))).all()
        assert (t.sum(axis=1).ravel() == O.table_bits(planes, 777, (x, z))).all()
        assert (t.sum(axis=0).ravel() == O.table_bits(planes, 777, (y, z))).all()


def test_golden_tables(golden, golden_planes):
    cases, controls, Nc, Nt = golden_planes
    for order in (2, 3):
        for s in golden[f"order{order}"]["sets"]:
            assert list(O.table_bits(cases, Nc, s["snps"])) == s["cases"]
            assert list(O.table_bits(controls, Nt, s["snps"])) == s["controls"]


def test_k2_exact_integer_form(golden):
    for order in (2, 3):
        for s in golden[f"order{order}"]["sets"]:
            v = O.score_k2(s["cases"], s["controls"])
            assert v == pytest.approx(s["k2"], rel=1e-12, abs=1e-12)


def test_k2_single_cell_closed_form():
    for Nc, Nt in [(1, 1), (45, 70), (1024, 1024), (131072, 131072)]:
        cas = np.zeros(27, dtype=np.uint32)
        ctl = np.zeros(27, dtype=np.uint32)
        cas[13], ctl[13] = Nc, Nt
        expect = math.lgamma(Nc + Nt + 2) - math.lgamma(Nc + 1) - math.lgamma(Nt + 1)
        assert O.score_k2(cas, ctl) == pytest.approx(expect, rel=1e-13)


def test_k2_rejects_plausible_mistakes():
    # A table where the "+1" in (a+b+1)! matters and cases/controls differ.
    cas = np.array([3, 0, 5, 1, 0, 0, 2, 0, 4], dtype=np.uint32)
    ctl = np.array([1, 2, 0, 0, 6, 1, 0, 3, 2], dtype=np.uint32)
    exact = math.log(math.prod((a + b + 1) * math.comb(a + b, a) for a, b in zip(cas.tolist(), ctl.tolist())))
    assert O.score_k2(cas, ctl) == pytest.approx(exact, rel=1e-13)
    no_plus_one = math.log(math.prod(math.comb(a + b, a) for a, b in zip(cas.tolist(), ctl.tolist())))
    assert abs(O.score_k2(cas, ctl) - no_plus_one) > 1.0


def test_mi_matches_sklearn(golden):
    for order in (2, 3):
        for s in golden[f"order{order}"]["sets"]:
            assert O.score_mi(s["cases"], s["controls"]) == pytest.approx(s["mi"], abs=1e-12)


def test_mi_special_cases():
    # independent: controls proportional to cases -> MI = 0
    cas = np.array([2, 4, 6, 0, 8, 10, 2, 0, 4], dtype=np.uint32)
    assert O.score_mi(cas, 3 * cas) == pytest.approx(0.0, abs=1e-14)
    # perfectly separating: MI = H(Y)
    cas = np.array([5, 0, 0, 0, 0, 0, 0, 0, 0], dtype=np.uint32)
    ctl = np.array([0, 3, 4, 0, 0, 0, 0, 0, 0], dtype=np.uint32)
    q = 5 / 12
    assert O.score_mi(cas, ctl) == pytest.approx(-(q * math.log2(q) + (1 - q) * math.log2(1 - q)), abs=1e-14)


def test_golden_search_and_tie_rule(golden, golden_planes):
    cases, controls, Nc, Nt = golden_planes
    for order in (2, 3):
        sets = golden[f"order{order}"]["sets"]
        for score, key, sign in ((O.K2, "k2", 1), (O.MI, "mi", -1)):
            r = O.search(cases, controls, Nc, Nt, order, score)
            best = min(sign * s[key] for s in sets)
            tied = [s["packed"] for s in sets if abs(sign * s[key] - best) <= 1e-9 * max(1.0, abs(best))]
            assert r["packed"] == min(tied)
            assert sign * r["score"] == pytest.approx(best, rel=1e-12, abs=1e-12)
            assert r["sets"] == len(sets)
        assert golden[f"order{order}"]["best_k2"] == list(episdet.unpack(order, O.search(
            cases, controls, Nc, Nt, order, O.K2)["packed"])[:order])


def test_duplicate_snps_tie_to_lowest_index():
    rng = np.random.default_rng(11)
    gc, pc = rand_planes(rng, 8, 300)
    gt, pt = rand_planes(rng, 8, 280)
    # SNP 6 duplicates SNP 1 in both classes, so every set with SNP 1 has a
    # twin with SNP 6 and an identical table; SNP 4 copies SNP 1 in cases
    # only, which makes (1,4) and its twin (4,6) the strongest pairs.  The
    # lower packed index of a tie must win.
    pc[6], pt[6] = pc[1], pt[1]
    pc[4] = pc[1]
    for order in (2, 3):
        r = O.search(pc, pt, 300, 280, order, O.K2)
        n = pc.shape[0]
        scores = {}
        for s in itertools.combinations(range(n), order):
            cas = O.table_bits(pc, 300, s)
            ctl = O.table_bits(pt, 280, s)
            scores[s] = O.score_k2(cas, ctl)
        b = min(scores.values())
        tied = [s for s, v in scores.items() if v == b]
        assert len(tied) >= 2
        packed = [s[0] | (s[1] << 21) | (s[2] << 42) if order == 3 else s[0] | (s[1] << 32) for s in tied]
        assert r["packed"] == min(packed)


def test_search_slices_cover_full_search():
    rng = np.random.default_rng(3)
    _, pc = rand_planes(rng, 30, 500)
    _, pt = rand_planes(rng, 30, 600)
    for order in (2, 3):
        full = O.search(pc, pt, 500, 600, order, O.K2)
        parts = [O.search(pc, pt, 500, 600, order, O.K2, last=(lo, min(lo + 7, 30))) for lo in range(order - 1, 30, 7)]
        assert sum(p["sets"] for p in parts) == full["sets"] == math.comb(30, order)
        best = min(parts, key=lambda p: (p["score"], p["packed"]))
        assert (best["packed"], best["score"]) == (full["packed"], full["score"])


@pytest.mark.parametrize("n,N,W", [(7, 64, 2), (75, 700, 32), (150, 96, 3)])
def test_search_checksum_equals_per_set_tables(n, N, W):
    """The search's table checksum (X rows blocked and counted two at a time,
    an odd last row paired with itself) equals the per-set bit-plane count of
    every table (orc_all_tables), for both orders; 150 SNPs on 3 words span
    several X blocks of 64."""
    rng = np.random.default_rng(n + N)
    _, pc = rand_planes(rng, n, N, W)
    _, pt = rand_planes(rng, n, N + 5 if N + 5 <= 32 * W else N, W)
    Nt = N + 5 if N + 5 <= 32 * W else N
    for order in (2, 3):
        _, ref = O.all_tables(pc, pt, N, Nt, order, want_tables=False)
        r = O.search(pc, pt, N, Nt, order, O.K2, checksum=True)
        assert r["checksum"] == ref
        parts = [O.search(pc, pt, N, Nt, order, O.MI, checksum=True, last=(lo, min(lo + 11, n)))
                 for lo in range(order - 1, n, 11)]
        assert sum(p["checksum"] for p in parts) % 2**64 == ref


def test_oracle_recovers_planted_triplet():
    syn = episdet.generate(48, 4000, 4000, seed=7, planted_sets=1, planted_order=3)
    r = O.search(syn.cases, syn.controls, 4000, 4000, 3, O.K2)
    assert tuple(r["idx"]) == syn.planted[0]
    r = O.search(syn.cases, syn.controls, 4000, 4000, 3, O.MI)
    assert tuple(r["idx"]) == syn.planted[0]


def test_oracle_recovers_planted_pair():
    syn = episdet.generate(64, 3000, 4500, seed=9, planted_sets=1, planted_order=2)
    for score in (O.K2, O.MI):
        r = O.search(syn.cases, syn.controls, 3000, 4500, 2, score)
        assert tuple(r["idx"]) == syn.planted[0]


@pytest.mark.parametrize("order,score", [(2, 0), (2, 1), (3, 0), (3, 1)])
def test_ranked_agrees_with_search_and_direct_counting(order, score):
    """oracle.ranked (used to check top-k) puts orc_search's winner first,
    is monotone, breaks exact ties by packed index, and scores each set as a
    per-sample genotype count does."""
    rng = np.random.default_rng(40 + order + 2 * score)
    n, Nc, Nt = (14, 300, 260) if order == 3 else (30, 300, 260)
    gc, cases = rand_planes(rng, n, Nc)
    gt, controls = rand_planes(rng, n, Nt)
    cases[5], controls[5] = cases[2], controls[2]          # a duplicated SNP makes exact ties
    gc[5], gt[5] = gc[2], gt[2]
    packed, sc = O.ranked(cases, controls, Nc, Nt, order, score)
    assert len(packed) == math.comb(n, order)
    best = O.search(cases, controls, Nc, Nt, order, score)
    assert int(packed[0]) == best["packed"] and sc[0] == best["score"]
    d = np.diff(sc)
    assert (d >= 0).all() if score == 0 else (d <= 0).all()
    tie = d == 0
    assert tie.any() and (np.diff(packed.astype(np.int64))[tie] > 0).all()
    T = O.lnfact_table(Nc + Nt + 1) if score == 0 else O.log2_table(Nc + Nt)
    for i in list(range(5)) + list(rng.integers(0, len(packed), 20)):
        p = int(packed[i])
        s = [p & 0x1FFFFF, (p >> 21) & 0x1FFFFF, p >> 42] if order == 3 else [p & 0xFFFFFFFF, p >> 32]
        a = O.table_direct(gc, Nc, s)
        b = O.table_direct(gt, Nt, s)
        ref = O.score_k2(a, b, T) if score == 0 else O.score_mi(a, b, T)
        assert ref == sc[i]
