"""GPU parity: the gfx950 kernels against the CPU oracle and the golden
fixtures, through the C ABI.

Bar: every contingency table bit-exact; the winning set identical (ties to
the lowest packed index); the winner's FP64 score bit-identical to the
oracle's FP64 score (same tables, same operation sequence); the FP32 scores
of the fused kernel within 1e-5 relative of FP64 for K2.
"""
import itertools
import math
import sys

import numpy as np
import pytest

import episdet
import oracle as O

pytestmark = pytest.mark.gpu

# the FP32 K2 and MI of every re-scored candidate against its FP64 score
FP32_REL_TOL = 1e-5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def packed_of(order, s):
    return s[0] | (s[1] << 21) | (s[2] << 42) if order == 3 else s[0] | (s[1] << 32)


def check_best(ds, syn_cases, syn_controls, Nc, Nt, order, score, path=episdet.PATH_FUSED_VALU, tile=0):
    say(f"  best: order={order} score={score} path={path} tile={tile} n={syn_cases.shape[0]} Nc={Nc} Nt={Nt}")
    r = ds.run(order, score, path=path, tile=tile)
    o = O.search(syn_cases, syn_controls, Nc, Nt, order, score)
    assert r["packed"] == o["packed"], (r, o)
    assert r["best_score"] == o["score"], (r["best_score"], o["score"])
    return r


# ------------------------------------------------------------- golden ----

def test_golden_tables_and_best(golden, golden_planes):
    cases, controls, Nc, Nt = golden_planes
    with episdet.Dataset(cases, controls, Nc, Nt) as ds:
        for order in (2, 3):
            say(f"golden dump order={order}")
            dump = ds.dump_tables(order)
            sets = golden[f"order{order}"]["sets"]
            assert dump.shape[0] == len(sets)
            for s in sets:
                snps = s["snps"]
                r = sum(math.comb(v, i + 1) for i, v in enumerate(snps))
                assert list(dump[r, 0]) == s["cases"] and list(dump[r, 1]) == s["controls"], snps
            t = ds.tables(order, [s["snps"] for s in sets])
            assert (t[:, 0] == np.array([s["cases"] for s in sets])).all()
            assert (t[:, 1] == np.array([s["controls"] for s in sets])).all()
            for score, key, sign in ((episdet.K2, "k2", 1), (episdet.MI, "mi", -1)):
                r = ds.run(order, score)
                best = min(sign * s[key] for s in sets)
                tied = [s["packed"] for s in sets if abs(sign * s[key] - best) <= 1e-9 * max(1.0, abs(best))]
                assert r["packed"] == min(tied)
                assert sign * r["best_score"] == pytest.approx(best, rel=1e-12, abs=1e-12)


# -------------------------------------------------- random vs oracle ----

CASES = [
    # n, Nc, Nt, words_c, words_t, seed
    (24, 45, 70, 0, 0, 1),
    (67, 1000, 1100, 0, 0, 2),
    (130, 2048, 2048, 0, 0, 3),
    (90, 3001, 2500, 101, 79, 4),       # arbitrary padding (W % 128 != 0)
    (33, 1, 1, 0, 0, 5),                # one sample per class
]


@pytest.mark.parametrize("n,Nc,Nt,wc,wt,seed", CASES)
def test_fused_tables_bit_exact(n, Nc, Nt, wc, wt, seed):
    syn = episdet.generate(n, Nc, Nt, seed=seed, cases_words=wc, controls_words=wt)
    with episdet.Dataset.from_synthetic(syn) as ds:
        V, M = episdet.PATH_FUSED_VALU, episdet.PATH_FUSED_MFMA
        for order, path, tile in ((2, V, 0), (2, M, 0), (2, M, 1), (3, V, 0), (3, M, 0), (3, M, 1), (3, M, 2), (3, M, 3),
                                  (3, M, 4), (3, M, 5)):
            say(f"dump order={order} path={path} tile={tile} n={n} Nc={Nc} Nt={Nt}")
            dump = ds.dump_tables(order, path=path, tile=tile)
            ref, cs = O.all_tables(syn.cases, syn.controls, Nc, Nt, order)
            assert (dump == ref).all()
            say(f"checksum order={order} path={path} tile={tile}")
            assert ds.checksum(order, path=path, tile=tile) == cs


@pytest.mark.parametrize("n,Nc,Nt,wc,wt,seed", CASES)
def test_fused_best_matches_oracle(n, Nc, Nt, wc, wt, seed):
    syn = episdet.generate(n, Nc, Nt, seed=seed, cases_words=wc, controls_words=wt,
                           planted_sets=1 if n >= 60 else 0, planted_order=3)
    with episdet.Dataset.from_synthetic(syn) as ds:
        for order, path in ((2, episdet.PATH_FUSED_VALU), (2, episdet.PATH_FUSED_MFMA), (3, episdet.PATH_FUSED_VALU),
                            (3, episdet.PATH_FUSED_MFMA)):
            for score in (episdet.K2, episdet.MI):
                r = check_best(ds, syn.cases, syn.controls, Nc, Nt, order, score, path=path)
                if r["fallback"] == 0:
                    assert r["fp32_max_rel_err"] <= FP32_REL_TOL


def test_fp32_mi_relative_error_on_weak_sets():
    """MI on data with no association (MI of 1e-3 bits and less, where H(G) +
    H(Y) - H(G,Y) cancels three orders of magnitude): the FP32 MI of every candidate,
    the top-1000 included, is within 1e-5 relative of its FP64 score, on
    every path and order, and the winner is the oracle's."""
    syn = episdet.generate(160, 3000, 4500, seed=17)
    with episdet.Dataset.from_synthetic(syn) as ds:
        for order, path, tile in ((2, episdet.PATH_FUSED_MFMA, 1), (2, episdet.PATH_FUSED_VALU, 0),
                                  (3, episdet.PATH_FUSED_MFMA, 4), (3, episdet.PATH_FUSED_MFMA, 5),
                                  (3, episdet.PATH_FUSED_VALU, 0)):
            r = check_best(ds, syn.cases, syn.controls, 3000, 4500, order, episdet.MI, path=path, tile=tile)
            assert r["best_score"] < 0.02 and r["fp32_max_rel_err"] <= FP32_REL_TOL, r
            say(f"  top-1000 order={order} path={path} tile={tile}")
            _, sc, rk = ds.topk(order, 1000, score=episdet.MI, path=path, tile=tile)
            assert len(sc) == 1000 and rk["candidates"] >= 1000
            assert rk["fp32_max_rel_err"] <= FP32_REL_TOL, rk


@pytest.mark.parametrize("tile", [0, 1, 3])
def test_triplet_tiles_agree(tile):
    syn = episdet.generate(150, 1500, 1700, seed=11, planted_sets=1)
    with episdet.Dataset.from_synthetic(syn) as ds:
        check_best(ds, syn.cases, syn.controls, 1500, 1700, 3, episdet.K2, tile=tile)


@pytest.mark.parametrize("tile", [0, 1, 2, 3, 4, 5])
def test_mfma_versions_agree(tile):
    syn = episdet.generate(170, 2500, 2100, seed=19, planted_sets=1)
    with episdet.Dataset.from_synthetic(syn) as ds:
        check_best(ds, syn.cases, syn.controls, 2500, 2100, 3, episdet.K2, path=episdet.PATH_FUSED_MFMA, tile=tile)


def test_seeded_search_keeps_the_exact_winner():
    """3-way single-best searches start from a bound found through the 4 best
    pairs (seeded = 1); the winner and its FP64 score stay the oracle's, with a
    planted triplet and with no planted set at all.  Fewer than 4 SNPs: no
    seed."""
    for planted, seed in ((1, 71), (0, 72)):
        syn = episdet.generate(160, 2500, 2500, seed=seed, planted_sets=planted)
        with episdet.Dataset.from_synthetic(syn) as ds:
            for path in (episdet.PATH_FUSED_MFMA, episdet.PATH_FUSED_VALU):
                r = check_best(ds, syn.cases, syn.controls, 2500, 2500, 3, episdet.K2, path=path)
                assert r["seeded"] == 1
            if planted:
                assert tuple(r["idx"]) == syn.planted[0]
    syn = episdet.generate(3, 400, 300, seed=73)
    with episdet.Dataset.from_synthetic(syn) as ds:
        r = check_best(ds, syn.cases, syn.controls, 400, 300, 3, episdet.K2, path=episdet.PATH_FUSED_MFMA)
        assert r["seeded"] == 0


NULL_KERNELS = [(3, episdet.PATH_FUSED_MFMA, t) for t in range(6)] + [(3, episdet.PATH_FUSED_VALU, 0)] + \
    [(2, episdet.PATH_FUSED_MFMA, t) for t in range(2)] + [(2, episdet.PATH_FUSED_VALU, 0)]


_NULL_RANKED = {}


@pytest.mark.parametrize("order,path,tile", NULL_KERNELS)
def test_null_data_k2_bound_keeps_the_winner(order, path, tile):
    """No planted set: K2 winners are weak, and the score bound skips most
    sets only through its per-cell Stirling term.  Single best and top-k
    (k-th-best bound) stay the oracle's, FP64 scores bit-identical."""
    n, Nc, Nt = (150, 2100, 1900) if order == 3 else (600, 1700, 2300)
    if order not in _NULL_RANKED:              # the oracle's full ranking, once per order
        syn = episdet.generate(n, Nc, Nt, seed=81 + order, planted_sets=0)
        _NULL_RANKED[order] = (syn, O.ranked(syn.cases, syn.controls, Nc, Nt, order, episdet.K2))
    syn, (ref_p, ref_s) = _NULL_RANKED[order]
    with episdet.Dataset.from_synthetic(syn) as ds:
        check_best(ds, syn.cases, syn.controls, Nc, Nt, order, episdet.K2, path=path, tile=tile)
        for k in (10, 200):
            say(f"  null topk order={order} path={path} tile={tile} k={k}")
            p, sc, _ = ds.topk(order, k, score=episdet.K2, path=path, tile=tile)
            assert p == [int(x) for x in ref_p[:k]]
            assert sc == [float(x) for x in ref_s[:k]]


def test_more_than_2_24_samples_per_class():
    """Both classes above 2^24 samples (33.6 M in all): the FP4 path's FP32
    counts would not be exact, so AUTO runs the VALU path and a forced MFMA
    path is refused; the K2 bound runs without its Stirling term (N >= 2^25).
    Winners and FP64 scores equal the oracle's for 2-/3-way, K2 and MI."""
    Nc, Nt = (1 << 24) + 37, (1 << 24) + 1001
    syn = episdet.generate(5, Nc, Nt, seed=91, planted_sets=0)
    with episdet.Dataset.from_synthetic(syn) as ds:
        for order in (2, 3):
            for score in (episdet.K2, episdet.MI):
                r = check_best(ds, syn.cases, syn.controls, Nc, Nt, order, score, path=episdet.PATH_AUTO)
                assert r["path"] == episdet.PATH_FUSED_VALU
            say(f"  forced MFMA path at Nc={Nc}: refused")
            with pytest.raises(episdet.EpisdetError) as e:
                ds.run(order, episdet.K2, path=episdet.PATH_FUSED_MFMA)
            assert e.value.code == 1


def test_auto_runs_sequenced_classes_for_uneven_classes():
    """AUTO runs version 5 (tile 5, classes in sequence) when one class has
    under 65% of the other's words, version 4 (tile 4) otherwise; both give
    the oracle's winner."""
    for Nc, Nt, tile in ((1500, 4600, 5), (3000, 3100, 4)):
        syn = episdet.generate(90, Nc, Nt, seed=61, planted_sets=1)
        with episdet.Dataset.from_synthetic(syn) as ds:
            r = check_best(ds, syn.cases, syn.controls, Nc, Nt, 3, episdet.K2, path=episdet.PATH_AUTO)
            assert (r["path"], r["tile"]) == (episdet.PATH_FUSED_MFMA, tile)


@pytest.mark.parametrize("tile", [0, 1])
def test_mfma_pair_versions_agree(tile):
    """Both 2-way MFMA kernels on uneven classes (the batch and autotuner
    choose between them): winner and FP64 score equal to the oracle's."""
    syn = episdet.generate(700, 2600, 3900, seed=23, planted_sets=1, planted_order=2)
    with episdet.Dataset.from_synthetic(syn) as ds:
        for score in (episdet.K2, episdet.MI):
            check_best(ds, syn.cases, syn.controls, 2600, 3900, 2, score, path=episdet.PATH_FUSED_MFMA, tile=tile)


@pytest.mark.parametrize("tile", [0, 1])
def test_plain_path_matches_oracle(tile):
    """The reference-shaped round loop, its rounds counted on the VALU (tile
    0) or by the FP4 MFMA count kernels (tile 1), tables through HBM, FP64
    objective: the oracle's winner and score, every round run, 3-way and
    2-way; ragged classes, a round count that is not a multiple of the MFMA
    launch's 16."""
    for n, Nc, Nt, seed in ((300, 1024, 1024, 12), (281, 1500, 1733, 14)):
        syn = episdet.generate(n, Nc, Nt, seed=seed, planted_sets=1)
        with episdet.Dataset.from_synthetic(syn) as ds:
            for score in (episdet.K2, episdet.MI):
                r = check_best(ds, syn.cases, syn.controls, Nc, Nt, 3, score, path=episdet.PATH_PLAIN, tile=tile)
                assert r["rounds_run"] == episdet.reference_rounds(3, n) and r["tile"] == tile
            assert tuple(r["idx"]) == syn.planted[0]
            for score in (episdet.K2, episdet.MI):
                r = check_best(ds, syn.cases, syn.controls, Nc, Nt, 2, score, path=episdet.PATH_PLAIN, tile=tile)
                assert r["rounds_run"] == episdet.reference_rounds(2, n) and r["tile"] == tile
            if tile == 1:
                say("plain MFMA, first 37 rounds only")
                r = ds.run(3, episdet.K2, path=episdet.PATH_PLAIN, tile=1, plain_rounds=37)
                assert r["rounds_run"] == 37


def test_explicit_set_tables():
    syn = episdet.generate(200, 5000, 7000, seed=13)
    rng = np.random.default_rng(0)
    with episdet.Dataset.from_synthetic(syn) as ds:
        for order in (2, 3):
            sets = np.array([np.sort(rng.choice(200, order, replace=False)) for _ in range(300)], dtype=np.int32)
            say(f"tables order={order}")
            t = ds.tables(order, sets)
            for i, s in enumerate(sets):
                assert (t[i, 0] == O.table_bits(syn.cases, 5000, s)).all()
                assert (t[i, 1] == O.table_bits(syn.controls, 7000, s)).all()


# -------------------------------------------------------- edge cases ----

def test_minimum_sizes():
    for n, order in ((2, 2), (3, 3), (3, 2)):
        syn = episdet.generate(n, 100, 90, seed=n)
        with episdet.Dataset.from_synthetic(syn) as ds:
            check_best(ds, syn.cases, syn.controls, 100, 90, order, episdet.K2)


def test_mfma_paths_agree_with_valu_and_oracle():
    syn = episdet.generate(200, 5000, 3000, seed=17, planted_sets=1)
    with episdet.Dataset.from_synthetic(syn) as ds:
        for score in (episdet.K2, episdet.MI):
            a = check_best(ds, syn.cases, syn.controls, 5000, 3000, 3, score, path=episdet.PATH_FUSED_MFMA)
            b = check_best(ds, syn.cases, syn.controls, 5000, 3000, 3, score, path=episdet.PATH_FUSED_VALU)
            assert a["packed"] == b["packed"] and a["best_score"] == b["best_score"]
        assert ds.checksum(3, path=episdet.PATH_FUSED_MFMA) == ds.checksum(3, path=episdet.PATH_FUSED_VALU)


def test_degenerate_data_all_homozygous_and_duplicates():
    # every SNP monomorphic (all genotype 0): every table has one full cell,
    # all scores tie exactly, the lowest packed index must win.
    n, Nc, Nt = 40, 3000, 2000
    cases = np.zeros((n, 2, episdet.padded_words(Nc)), dtype=np.uint32)
    controls = np.zeros((n, 2, episdet.padded_words(Nt)), dtype=np.uint32)
    for arr, N in ((cases, Nc), (controls, Nt)):
        full, rem = divmod(N, 32)
        arr[:, 0, :full] = 0xFFFFFFFF
        if rem:
            arr[:, 0, full] = (0xFFFFFFFF << (32 - rem)) & 0xFFFFFFFF
    with episdet.Dataset(cases, controls, Nc, Nt) as ds:
        for order, path in ((2, episdet.PATH_FUSED_VALU), (3, episdet.PATH_FUSED_VALU), (3, episdet.PATH_FUSED_MFMA)):
            dump = ds.dump_tables(order, path=path)
            ref, _ = O.all_tables(cases, controls, Nc, Nt, order)
            assert (dump == ref).all()
            for score in (episdet.K2, episdet.MI):
                r = check_best(ds, cases, controls, Nc, Nt, order, score, path=path)
                assert r["packed"] == packed_of(order, tuple(range(order)))


def test_duplicate_snps_tie_rule():
    syn = episdet.generate(80, 2000, 2000, seed=21)
    syn.cases[50], syn.controls[50] = syn.cases[10], syn.controls[10]
    syn.cases[30] = syn.cases[10]       # strong pairs (10,30) and its twin (30,50)
    with episdet.Dataset.from_synthetic(syn) as ds:
        for order in (2, 3):
            for score in (episdet.K2, episdet.MI):
                check_best(ds, syn.cases, syn.controls, 2000, 2000, order, score)


# --------------------------------------------- reference shapes (small) --

def test_512_snps_2048_samples_all_pair_tables():
    """512 SNPs, 1024 cases + 1024 controls: all 130,816 pair tables
    bit-exact and the best pair identical."""
    syn = episdet.generate(512, 1024, 1024, seed=1)
    with episdet.Dataset.from_synthetic(syn) as ds:
        ref, cs = O.all_tables(syn.cases, syn.controls, 1024, 1024, 2)
        for path, tile in ((episdet.PATH_FUSED_VALU, 0), (episdet.PATH_FUSED_MFMA, 0), (episdet.PATH_FUSED_MFMA, 1)):
            say(f"512 x 2048 dump path={path} tile={tile}")
            dump = ds.dump_tables(2, path=path, tile=tile)
            assert dump.shape[0] == 130_816
            assert (dump == ref).all()
            assert ds.checksum(2, path=path, tile=tile) == cs
            for score in (episdet.K2, episdet.MI):
                check_best(ds, syn.cases, syn.controls, 1024, 1024, 2, score, path=path, tile=tile)


def test_full_sample_axis_planted_triplet():
    """The benchmark's sample axis (131072 + 131072) on 192 SNPs: winner
    identical to the oracle, the planted triplet recovered, FP32 within tol."""
    syn = episdet.generate(192, 131072, 131072, seed=2, planted_sets=1, planted_order=3)
    with episdet.Dataset.from_synthetic(syn) as ds:
        for path, tile in ((episdet.PATH_FUSED_VALU, 0), (episdet.PATH_FUSED_MFMA, 0), (episdet.PATH_FUSED_MFMA, 2),
                           (episdet.PATH_FUSED_MFMA, 3), (episdet.PATH_FUSED_MFMA, 4), (episdet.PATH_FUSED_MFMA, 5)):
            r = check_best(ds, syn.cases, syn.controls, 131072, 131072, 3, episdet.K2, path=path, tile=tile)
            assert tuple(r["idx"]) == syn.planted[0]
            assert r["fp32_max_rel_err"] <= FP32_REL_TOL
        say("checksums at the full sample axis: MFMA tiles vs VALU")
        cv = ds.checksum(3, path=episdet.PATH_FUSED_VALU)
        for tile in (0, 1, 2, 3, 4, 5):
            assert ds.checksum(3, path=episdet.PATH_FUSED_MFMA, tile=tile) == cv
        say("checksum at full sample axis vs plain tables")
        sets = np.array(list(itertools.islice(itertools.combinations(range(192), 3), 0, 4000, 7)), dtype=np.int32)
        t = ds.tables(3, sets)
        for i in range(0, len(sets), 50):
            assert (t[i, 0] == O.table_bits(syn.cases, 131072, sets[i])).all()
            assert (t[i, 1] == O.table_bits(syn.controls, 131072, sets[i])).all()


def test_full_sample_axis_checksums_agree():
    """The benchmark's sample axis (131072 + 131072) on 1024 SNPs (178,433,024
    triplets, 1/64 of the benchmark's): the order-independent checksum of
    every (set, 54-cell table) from the default kernel equals that of two
    other tilings (tiles 0 and 5) and of the integer AND+popcount VALU path.
    (At 4096 SNPs each checksum pass would hash 1.1e10 tables, minutes.)"""
    syn = episdet.generate(1024, 131072, 131072, seed=1, planted_sets=1, planted_order=3)
    with episdet.Dataset.from_synthetic(syn) as ds:
        say("1024 x 262144 checksum: MFMA tile 4")
        c4 = ds.checksum(3, path=episdet.PATH_FUSED_MFMA, tile=4)
        for tile in (0, 5):
            say(f"1024 x 262144 checksum: MFMA tile {tile}")
            assert ds.checksum(3, path=episdet.PATH_FUSED_MFMA, tile=tile) == c4
        say("1024 x 262144 checksum: VALU")
        assert ds.checksum(3, path=episdet.PATH_FUSED_VALU) == c4


def test_mi_pairs_uneven_classes_planted():
    """Config-3 style: 40/60 classes with padding that is not a 1024-bit
    multiple of the valid bits, MAF ~ Beta(0.5, 2), planted pair."""
    syn = episdet.generate(1024, 26214, 39322, seed=3, maf_dist=episdet.MAF_BETA, maf=(0.01, 0.5),
                           planted_sets=1, planted_order=2)
    with episdet.Dataset.from_synthetic(syn) as ds:
        for path in (episdet.PATH_FUSED_VALU, episdet.PATH_FUSED_MFMA):
            r = check_best(ds, syn.cases, syn.controls, 26214, 39322, 2, episdet.MI, path=path)
            assert tuple(r["idx"][:2]) == syn.planted[0]
        say("uneven classes: VALU and MFMA pair checksums")
        assert ds.checksum(2, path=episdet.PATH_FUSED_MFMA) == ds.checksum(2, path=episdet.PATH_FUSED_VALU)


# ------------------------------------------------------------- batches --

def test_batch_of_uneven_jobs_matches_oracle():
    """Mixed 2-/3-way, K2/MI jobs of uneven sizes and tight (non-1024)
    padding in one graph launch: every winner and FP64 score equals the
    oracle's."""
    rng = np.random.default_rng(4)
    jobs = []
    for j in range(12):
        order = 3 if j % 2 else 2
        n = int(rng.integers(40, 120 if order == 3 else 300))
        Nc, Nt = int(rng.integers(300, 5000)), int(rng.integers(300, 5000))
        syn = episdet.generate(n, Nc, Nt, seed=100 + j, cases_words=(Nc + 31) // 32 + (j % 3),
                               controls_words=(Nt + 31) // 32, planted_sets=1 if j % 3 == 0 else 0,
                               planted_order=order)
        jobs.append((order, episdet.MI if j % 4 >= 2 else episdet.K2, syn))
    say("batch of 12 via host buffers")
    res, ms = episdet.search_batch(jobs)
    assert ms > 0
    for (order, score, syn), r in zip(jobs, res):
        o = O.search(syn.cases, syn.controls, syn.num_cases, syn.num_controls, order, score)
        assert r["packed"] == o["packed"], (order, score, syn.num_snps, r, o)
        assert r["best_score"] == o["score"]
    say("same batch resident, twice (graph relaunch)")
    dss = [episdet.Dataset.from_synthetic(syn) for _, _, syn in jobs]
    for rep in range(2):
        res2, _ = episdet.run_batch(dss, [o for o, _, _ in jobs], [s for _, s, _ in jobs], rebuild=rep == 1)
        assert [r["packed"] for r in res2] == [r["packed"] for r in res]
    say("explicit MFMA tiles 0 and 1 side by side in one graph")
    M = episdet.PATH_FUSED_MFMA
    res4, _ = episdet.run_batch(dss, [o for o, _, _ in jobs], [s for _, s, _ in jobs], paths=[M] * len(jobs),
                                tiles=[(j // 2) % 6 if o == 3 else (j // 2) % 2 for j, (o, _, _) in enumerate(jobs)])
    assert [r["packed"] for r in res4] == [r["packed"] for r in res]
    assert [r["best_score"] for r in res4] == [r["best_score"] for r in res]
    say("one dataset in two jobs of one batch")
    res3, _ = episdet.run_batch([dss[1], dss[1], dss[0]], [3, 2, 2], [episdet.K2, episdet.MI, episdet.K2])
    assert res3[0]["packed"] == res[1]["packed"] and res3[2]["packed"] == res[0]["packed"]
    o = O.search(jobs[1][2].cases, jobs[1][2].controls, jobs[1][2].num_cases, jobs[1][2].num_controls, 2, episdet.MI)
    assert res3[1]["packed"] == o["packed"] and res3[1]["best_score"] == o["score"]
    for d in dss:
        d.close()


# ------------------------------------------------------------- top-k ----

TOPK_CASES = [
    # order, path, tile, n, Nc, Nt, seed
    (2, episdet.PATH_FUSED_VALU, 0, 150, 900, 1100, 31),
    (2, episdet.PATH_FUSED_MFMA, 0, 200, 1100, 700, 36),
    (2, episdet.PATH_FUSED_MFMA, 1, 230, 1300, 800, 39),
    (3, episdet.PATH_FUSED_VALU, 0, 45, 1500, 1300, 32),
    (3, episdet.PATH_FUSED_MFMA, 0, 100, 2000, 1800, 33),
    (3, episdet.PATH_FUSED_MFMA, 1, 60, 700, 2100, 34),
    (3, episdet.PATH_FUSED_MFMA, 3, 70, 1500, 900, 37),
    (3, episdet.PATH_FUSED_MFMA, 4, 80, 2500, 1200, 38),
    (3, episdet.PATH_FUSED_MFMA, 5, 75, 900, 2300, 40),
]


@pytest.mark.parametrize("order,path,tile,n,Nc,Nt,seed", TOPK_CASES)
def test_topk_matches_oracle_ranking(order, path, tile, n, Nc, Nt, seed):
    """The k best sets, best first, with FP64 scores bit-identical to the
    oracle's full ranking (ties by packed index), for k from 1 to 1000."""
    syn = episdet.generate(n, Nc, Nt, seed=seed, planted_sets=1, planted_order=order)
    cases, controls = syn.cases.copy(), syn.controls.copy()
    cases[7], controls[7] = cases[3], controls[3]          # exact ties in the ranking
    with episdet.Dataset(cases, controls, Nc, Nt) as ds:
        for score in (episdet.K2, episdet.MI):
            ref_p, ref_s = O.ranked(cases, controls, Nc, Nt, order, score)
            for k in (1, 2, 10, 100, 1000):
                say(f"topk order={order} path={path} tile={tile} n={n} score={score} k={k}")
                p, s, r = ds.topk(order, k, score=score, path=path, tile=tile)
                assert len(p) == k
                assert p == [int(x) for x in ref_p[:k]]
                assert s == [float(x) for x in ref_s[:k]]
                assert r["packed"] == p[0] and r["best_score"] == s[0]


def test_topk_more_than_all_sets_and_errors():
    syn = episdet.generate(6, 400, 500, seed=35)
    with episdet.Dataset.from_synthetic(syn) as ds:
        for order in (2, 3):
            ref_p, ref_s = O.ranked(syn.cases, syn.controls, 400, 500, order, episdet.K2)
            p, s, _ = ds.topk(order, 50)
            assert p == [int(x) for x in ref_p] and s == [float(x) for x in ref_s]
        with pytest.raises(episdet.EpisdetError):
            ds.topk(3, 10, path=episdet.PATH_PLAIN)
        with pytest.raises(episdet.EpisdetError):
            ds.topk(3, episdet.TOPK_MAX + 1)


def test_sample_in_both_planes_is_rejected():
    """A sample marked in both genotype planes of a SNP encodes no genotype
    (its genotype-2 count N - c0 - c1 would be off): dataset creation and the
    one-shot search refuse it with EPISDET_ERR_GENOTYPES.  Both bits set past
    the last sample are padding, cleared on upload, and accepted."""
    syn = episdet.generate(20, 300, 200, seed=3)
    bad = syn.cases.copy()
    bad[7, 0, 2] |= 0x00010000          # sample 2 * 32 + 15 of SNP 7, in both planes
    bad[7, 1, 2] |= 0x00010000
    say("overlapping planes")
    with pytest.raises(episdet.EpisdetError) as e:
        episdet.Dataset(bad, syn.controls, 300, 200)
    assert e.value.code == 7
    with pytest.raises(episdet.EpisdetError) as e:
        episdet.search(3, episdet.K2, bad, syn.controls, 300, 200)
    assert e.value.code == 7
    pad = syn.controls.copy()
    pad[3, 0, 6] |= 1                   # sample 6 * 32 + 31 = 223 >= 200: padding
    pad[3, 1, 6] |= 1
    with episdet.Dataset(syn.cases, pad, 300, 200) as ds:
        r = ds.run(3, episdet.K2)
    o = O.search(syn.cases, syn.controls, 300, 200, 3, O.K2)
    assert r["packed"] == o["packed"] and r["best_score"] == o["score"]


# ---------------------------------------------------------- autotuner ----

@pytest.mark.parametrize("order,n,Nc,Nt", [(3, 150, 3000, 2500), (2, 700, 5000, 4000)])
def test_autotune_then_auto_matches_oracle(order, n, Nc, Nt):
    """Every candidate configuration is timed on a window of its work units,
    the fastest is cached, and AUTO (single and batched) then runs it with
    the oracle's winner."""
    syn = episdet.generate(n, Nc, Nt, seed=50 + order, planted_sets=1, planted_order=order)
    with episdet.Dataset.from_synthetic(syn) as ds:
        say(f"autotune order={order} n={n}")
        table, chosen = ds.autotune(order, episdet.K2)
        expect = (9 if order == 3 else 4)    # 3-way: MFMA tiles 0-5, VALU tiles 0, 1, 3; 2-way: MFMA 0-1, VALU R=4, 2
        assert len(table) == expect and 0 <= chosen < expect
        assert all(e["sample_ms"] > 0 and e["sample_sets"] > 0 and e["est_ms"] > 0 for e in table)
        assert table[chosen]["est_ms"] == min(e["est_ms"] for e in table)
        r = check_best(ds, syn.cases, syn.controls, Nc, Nt, order, episdet.K2, path=episdet.PATH_AUTO)
        assert (r["path"], r["tile"]) == (table[chosen]["path"], table[chosen]["tile"])
        res, _ = episdet.run_batch([ds], [order], [episdet.K2])
        assert res[0]["packed"] == r["packed"] and res[0]["best_score"] == r["best_score"]
