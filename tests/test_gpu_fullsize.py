"""GPU parity at full size: the benchmark workloads against the oracle.

tests/golden/fullsize.json holds, for each workload, what the CPU oracle
computed over every set of it (tests/golden/make_fullsize.py regenerates the
same seeded datasets and runs oracle/episdet_oracle.c over all of them): the
winner (packed index, indices, FP64 score as an exact hex float) and the
order-independent checksum of every set's 2 x 3^k table (orc_table_hash
summed mod 2^64, the quantity episdet_checksum computes on the GPU).  So
these tests pin every contingency table (helper.cu:438-548) and the best set
(helper.cu:552-584, tensor-episdet.cu:181-184) of each workload to the
oracle, not to another GPU kernel:

  3-way K2, 4096 SNPs x 262144 samples, one planted triplet
  2-way MI, 16384 SNPs x 65536 samples (40/60), MAF ~ Beta(0.5, 2)
  3-way K2, 8192 SNPs x 65536 samples, two planted triplets: MFMA and VALU
  the 64 uneven jobs of the batched workload, in one batch launch
"""
import json
import os
import sys

import pytest

import episdet

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
# the FP32 K2 and MI of every re-scored candidate against its FP64 score
FP32_REL_TOL = 1e-5
M, V = episdet.PATH_FUSED_MFMA, episdet.PATH_FUSED_VALU


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def fixture():
    with open(os.path.join(HERE, "golden", "fullsize.json")) as f:
        return json.load(f)


def workload(name):
    ent = fixture()["workloads"].get(name)
    if ent is None or ent.get("result") is None:
        pytest.skip(f"oracle fixture for {name} not generated (tests/golden/make_fullsize.py --only {name})")
    p = ent["params"]
    syn = episdet.generate(p["n"], p["cases"], p["controls"], seed=p["seed"], planted_sets=p["planted_sets"],
                           planted_order=p["order"], maf=tuple(p["maf"]), maf_dist=p["maf_dist"])
    if "planted" in ent:
        assert [list(s) for s in syn.planted] == ent["planted"]
    return p, ent["result"], syn


def check_winner(r, res, score):
    assert r["packed"] == res["packed"], (r["idx"], res["idx"])
    assert r["best_score"] == float.fromhex(res["score"]), (r["best_score"], res["score"])
    rel = abs(r["best_score_f32"] - r["best_score"]) / abs(r["best_score"])
    say(f"  winner {r['idx']} FP64 {r['best_score']!r} FP32 rel err of the winner {rel:.2e}, "
        f"max over {r['candidates']} candidates {r['fp32_max_rel_err']:.2e}")
    assert r["fp32_max_rel_err"] <= FP32_REL_TOL


def test_three_way_benchmark_workload():
    """4096 SNPs x 262144 samples: the checksum of all 11,444,858,880 tables
    from the default kernel (tile 4) and from tile 0 (64 x 64 units) equals
    the oracle's; the single-best search returns the oracle's winner and FP64
    score, and the top-16 search agrees with it."""
    p, res, syn = workload("w2")
    with episdet.Dataset.from_synthetic(syn) as ds:
        for tile in (4, 0):
            say(f"3-way workload checksum: MFMA tile {tile} (about 20 s)")
            assert ds.checksum(3, path=M, tile=tile) == int(res["checksum"], 16)
        say("3-way workload: single best")
        r = ds.run(3, episdet.K2)
        check_winner(r, res, episdet.K2)
        assert r["seeded"] == 1
        say("3-way workload: top-16")
        pk, sc, _ = ds.topk(3, 16, score=episdet.K2)
        assert pk[0] == res["packed"] and sc[0] == float.fromhex(res["score"])
        assert sc == sorted(sc) and len(set(pk)) == 16


def test_two_way_mi_workload():
    """16384 SNPs x 65536 samples (26214 cases / 39322 controls): the checksum
    of all 134,209,536 pair tables from both MFMA pair kernels and the VALU
    path equals the oracle's, and so do the MI winner and its FP64 score on
    both fused paths."""
    p, res, syn = workload("w3")
    with episdet.Dataset.from_synthetic(syn) as ds:
        for path, tile in ((M, 1), (M, 0), (V, 0)):
            say(f"2-way workload checksum: path {path} tile {tile}")
            assert ds.checksum(2, path=path, tile=tile) == int(res["checksum"], 16)
        for path, tile in ((M, 1), (V, 0)):
            say(f"2-way workload MI search: path {path} tile {tile}")
            check_winner(ds.run(2, episdet.MI, path=path, tile=tile), res, episdet.MI)


def test_large_three_way_workload_mfma():
    """8192 SNPs x 65536 samples, two planted triplets (91,592,417,280
    triplets): the default kernel's table checksum, winner and FP64 score
    equal the oracle's."""
    p, res, syn = workload("w5")
    with episdet.Dataset.from_synthetic(syn) as ds:
        say("large 3-way checksum: MFMA tile 4 (about 15 s)")
        assert ds.checksum(3, path=M, tile=4) == int(res["checksum"], 16)
        say("large 3-way search: AUTO")
        check_winner(ds.run(3, episdet.K2), res, episdet.K2)


def test_large_three_way_workload_valu_checksum():
    """The same workload's checksum from the AND + v_bcnt VALU kernel (about
    two minutes)."""
    p, res, syn = workload("w5")
    with episdet.Dataset.from_synthetic(syn) as ds:
        say("large 3-way checksum: VALU")
        assert ds.checksum(3, path=V) == int(res["checksum"], 16)


def test_large_three_way_workload_valu_search():
    """The same workload searched on the VALU kernel: the oracle's winner
    and FP64 score (about two minutes)."""
    p, res, syn = workload("w5")
    with episdet.Dataset.from_synthetic(syn) as ds:
        say("large 3-way search: VALU")
        check_winner(ds.run(3, episdet.K2, path=V), res, episdet.K2)


def test_batched_workload():
    """The 64 uneven jobs (2-/3-way, K2/MI, 256 to 8192 SNPs, 8000 to 131072
    samples packed tight) in one batch launch on AUTO kernels, autotuned per
    shape as the benchmark does: every job's
    winner and FP64 score equal the oracle's; every job's table checksum
    from its AUTO kernel equals the oracle's."""
    fx = fixture()["batch"]
    if fx is None or len(fx["jobs"]) < 64:
        pytest.skip("oracle fixture of the batched workload not generated")
    spec = fx["spec"]
    syns, dss = [], []
    try:
        for j, (order, score, n, Nc, Nt) in enumerate(spec):
            syn = episdet.generate(n, Nc, Nt, seed=fx["job_seed0"] + j, cases_words=(Nc + 31) // 32,
                                   controls_words=(Nt + 31) // 32)
            syns.append(syn)
            dss.append(episdet.Dataset.from_synthetic(syn))
        say("batched workload: autotuning the 64 shapes")
        for ds, s in zip(dss, spec):
            ds.autotune(s[0], s[1])
        say("batched workload: 64 jobs in one batch")
        res, ms = episdet.run_batch(dss, [s[0] for s in spec], [s[1] for s in spec])
        say(f"  batch graph {ms:.1f} ms")
        for j, (r, s) in enumerate(zip(res, spec)):
            want = fx["jobs"][str(j)]
            assert r["packed"] == want["packed"], (j, s, r["idx"], want["idx"])
            assert r["best_score"] == float.fromhex(want["score"]), (j, s)
            assert r["fp32_max_rel_err"] <= FP32_REL_TOL, (j, s)
        say("batched workload: 64 table checksums")
        for j, (ds, s) in enumerate(zip(dss, spec)):
            assert res[j]["path"] == M, (j, s, res[j]["path"])
            assert ds.checksum(s[0], path=M, tile=res[j]["tile"]) == int(fx["jobs"][str(j)]["checksum"], 16), (j, s)
    finally:
        for ds in dss:
            ds.close()
