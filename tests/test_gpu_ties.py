"""GPU: exact answers under heavy ties and candidate overflow, through the C ABI.

The fused kernels keep every set within the FP32 error margin of the running
best as a candidate for FP64 re-scoring.  When more sets than the candidate
buffer holds fall within it (exact ties, duplicated SNPs, a small
candidate_cap), the search reruns in windows of work units that fit
(result "fallback" = 1).  The plain path keeps each round's best by (exact
score, lowest packed index) and needs no candidate list.  In every case the
answer is the oracle's: the lowest packed index among the sets with the best
FP64 score (the reference's best-set reduction, helper.cu:552-583 and
reduction.hpp:64-122, with its nondeterministic index race made a total
order).
"""
import sys

import numpy as np
import pytest

import episdet
import oracle as O

pytestmark = pytest.mark.gpu

M, V, P = episdet.PATH_FUSED_MFMA, episdet.PATH_FUSED_VALU, episdet.PATH_PLAIN


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def monomorphic(n, Nc, Nt):
    """Every SNP genotype 0 in every sample: every table has one full cell,
    so all C(n, k) scores tie exactly."""
    cases = np.zeros((n, 2, episdet.padded_words(Nc)), dtype=np.uint32)
    controls = np.zeros((n, 2, episdet.padded_words(Nt)), dtype=np.uint32)
    for arr, N in ((cases, Nc), (controls, Nt)):
        full, rem = divmod(N, 32)
        arr[:, 0, :full] = 0xFFFFFFFF
        if rem:
            arr[:, 0, full] = (0xFFFFFFFF << (32 - rem)) & 0xFFFFFFFF
    return cases, controls


def colex_first(order, k):
    """Packed indices of the k sets with the lowest packed index: packing puts
    the largest index highest, so they are the first k sets in colex order."""
    out, z = [], order - 1
    while len(out) < k:
        if order == 3:
            out += [x | (y << 21) | (z << 42) for y in range(1, z) for x in range(y)]
        else:
            out += [a | (z << 32) for a in range(z)]
        z += 1
    return out[:k]


@pytest.mark.parametrize("path,tile", [(M, 4), (M, 0), (V, 0), (P, 0)])
def test_three_way_all_sets_tied(path, tile):
    """256 monomorphic SNPs: 2,763,520 triplets with the same exact score,
    more than the default 1M candidate slots.  The winner is (0, 1, 2) with
    the oracle's FP64 score, for K2 and MI."""
    n, Nc, Nt = 256, 1000, 1200
    cases, controls = monomorphic(n, Nc, Nt)
    with episdet.Dataset(cases, controls, Nc, Nt) as ds:
        for score in (episdet.K2, episdet.MI):
            say(f"3-way all tied: path={path} tile={tile} score={score}")
            r = ds.run(3, score, path=path, tile=tile)
            o = O.search(cases, controls, Nc, Nt, 3, score)
            assert o["packed"] == colex_first(3, 1)[0]
            assert r["packed"] == o["packed"] and r["best_score"] == o["score"], (r, o)
            if path != P:
                assert r["fallback"] == 1 and r["candidates"] == episdet.n_choose_k(n, 3)


@pytest.mark.parametrize("path,tile", [(M, 1), (M, 0), (V, 0), (P, 0)])
def test_two_way_all_sets_tied(path, tile):
    """1500 monomorphic SNPs: 1,124,250 tied pairs, more than 1M slots."""
    n, Nc, Nt = 1500, 700, 900
    cases, controls = monomorphic(n, Nc, Nt)
    with episdet.Dataset(cases, controls, Nc, Nt) as ds:
        for score in (episdet.K2, episdet.MI):
            say(f"2-way all tied: path={path} tile={tile} score={score}")
            r = ds.run(2, score, path=path, tile=tile)
            o = O.search(cases, controls, Nc, Nt, 2, score)
            assert r["packed"] == o["packed"] == 1 << 32 and r["best_score"] == o["score"], (r, o)
            if path == P:
                assert r["rounds_run"] == episdet.reference_rounds(2, n)
            else:
                assert r["fallback"] == 1


def planted_with_copies(n, Nc, Nt, seed, copies, order=3):
    """A planted set (triplet, or pair with order=2) whose SNPs are each
    copied `copies` times further up: every combination of copies ties with
    the planted set exactly, so (copies + 1)^order sets share the best
    score."""
    syn = episdet.generate(n, Nc, Nt, seed=seed, planted_sets=1, planted_order=order)
    cases, controls = syn.cases.copy(), syn.controls.copy()
    planted = syn.planted[0]
    free = [v for v in range(n) if v not in planted][::-1]
    for src in planted:
        for _ in range(copies):
            d = free.pop(0)
            cases[d], controls[d] = cases[src], controls[src]
    return cases, controls


@pytest.mark.parametrize("path,tile", [(M, 4), (M, 5), (M, 0), (V, 0)])
def test_small_candidate_cap_overflows_into_windows(path, tile):
    """Planted data with 17^3 = 4913 exact ties at the top and
    candidate_cap=64 (raised to one work unit's sets, at most 4096 here):
    the first launch overflows, the windowed rerun returns the oracle's
    winner and FP64 score."""
    Nc, Nt = 1500, 1300
    cases, controls = planted_with_copies(220, Nc, Nt, seed=41, copies=16)
    with episdet.Dataset(cases, controls, Nc, Nt) as ds:
        for score in (episdet.K2, episdet.MI):
            say(f"small cap: path={path} tile={tile} score={score}")
            r = ds.run(3, score, path=path, tile=tile, candidate_cap=64)
            o = O.search(cases, controls, Nc, Nt, 3, score)
            assert r["packed"] == o["packed"] and r["best_score"] == o["score"], (r, o)
            assert r["fallback"] == 1
            full = ds.run(3, score, path=path, tile=tile)
            assert full["packed"] == r["packed"] and full["best_score"] == r["best_score"]


def test_topk_windows_keep_the_ranking():
    """Top-k with a small candidate_cap on data with heavy ties at the top:
    the windowed rerun seeds each window with the k-th best so far and still
    returns the oracle's full ranking prefix."""
    Nc, Nt = 1200, 1100
    cases, controls = planted_with_copies(100, Nc, Nt, seed=43, copies=9)
    with episdet.Dataset(cases, controls, Nc, Nt) as ds:
        for score in (episdet.K2, episdet.MI):
            ref_p, ref_s = O.ranked(cases, controls, Nc, Nt, 3, score)
            for path, tile in ((M, 4), (V, 0)):
                for k in (1, 100, 1000):
                    say(f"topk windows: path={path} tile={tile} score={score} k={k}")
                    p, s, r = ds.topk(3, k, score=score, path=path, tile=tile, candidate_cap=2048)
                    assert p == [int(x) for x in ref_p[:k]]
                    assert s == [float(x) for x in ref_s[:k]]
                    if k > 1:                      # k = 1 is the single-best search
                        assert r["fallback"] == 1


@pytest.mark.parametrize("tile", [0, 1])
def test_two_way_windows_in_xcd_blocks(tile):
    """2-way on 1100 SNPs (18 tiles of 64, so both MFMA pair kernels take
    their units in XCD blocks from per-XCD queues) with 101^2 = 10201 exact
    K2 ties at the top, more than the 8192 candidate slots asked for: single
    best (candidate_cap=64, raised to one unit's sets) and top-k
    (candidate_cap=8192) rerun in windows of units, and return the oracle's
    winner and ranking prefix, for K2 and for MI."""
    Nc, Nt = 1200, 1100
    cases, controls = planted_with_copies(1100, Nc, Nt, seed=47, copies=100, order=2)
    with episdet.Dataset(cases, controls, Nc, Nt) as ds:
        for score in (episdet.K2, episdet.MI):
            ref_p, ref_s = O.ranked(cases, controls, Nc, Nt, 2, score)
            say(f"2-way windows: tile={tile} score={score} single best")
            r = ds.run(2, score, path=M, tile=tile, candidate_cap=64)
            assert r["packed"] == int(ref_p[0]) and r["best_score"] == float(ref_s[0]), r
            if score == episdet.K2:            # the 10201 ties lead the K2 ranking
                assert r["fallback"] == 1
            for k in (100, 1000):
                say(f"2-way windows: tile={tile} score={score} k={k}")
                pk, sc, rk = ds.topk(2, k, score=score, path=M, tile=tile, candidate_cap=8192)
                assert pk == [int(x) for x in ref_p[:k]]
                assert sc == [float(x) for x in ref_s[:k]]
                if score == episdet.K2:
                    assert rk["fallback"] == 1


def test_topk_all_sets_tied():
    """400 monomorphic SNPs (10,586,800 tied triplets, more than top-k's
    default slots): the k best are the k lowest packed indices, all with
    the oracle's score."""
    n, Nc, Nt = 400, 600, 500
    cases, controls = monomorphic(n, Nc, Nt)
    o = O.search(cases, controls, Nc, Nt, 3, episdet.K2)
    with episdet.Dataset(cases, controls, Nc, Nt) as ds:
        say("topk all tied, k=1000")
        p, s, r = ds.topk(3, 1000, score=episdet.K2)
        assert p == colex_first(3, 1000)
        assert s == [o["score"]] * 1000
        assert r["fallback"] == 1


def test_plain_two_way_matches_oracle():
    """The 2-way plain path: its own round loop over block pairs
    (search-pairs.cu:137-211), count tables through HBM, FP64 objective."""
    syn = episdet.generate(700, 1500, 1800, seed=44, planted_sets=1, planted_order=2)
    with episdet.Dataset.from_synthetic(syn) as ds:
        for score in (episdet.K2, episdet.MI):
            say(f"plain 2-way score={score}")
            r = ds.run(2, score, path=P)
            o = O.search(syn.cases, syn.controls, 1500, 1800, 2, score)
            assert r["packed"] == o["packed"] and r["best_score"] == o["score"]
            assert r["rounds_run"] == episdet.reference_rounds(2, 700) and r["path"] == P
        assert tuple(r["idx"][:2]) == syn.planted[0]
