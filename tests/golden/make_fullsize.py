"""Full-size oracle fixtures: the benchmark workloads searched on the CPU.

Regenerates the exact synthetic datasets of the benchmark workloads with the
project's seeded generator (host code, libepisdet_data.so) and runs the CPU
oracle (oracle/episdet_oracle.c, the checker) over every set of each:

  w2  3-way K2, 4096 SNPs x 262144 samples (131072 / 131072), MAF ~ U[0.05,0.5],
      one planted triplet, seed 1                     -- 11,444,858,880 triplets
  w3  2-way MI, 16384 SNPs x 65536 samples (26214 / 39322), MAF ~ Beta(0.5,2)
      clipped to [0.01,0.5], one planted pair, seed 1 -- 134,209,536 pairs
  w5  3-way K2, 8192 SNPs x 65536 samples (32768 / 32768), MAF ~ U[0.05,0.5],
      two planted triplets, seed 1                    -- 91,592,417,280 triplets
  w4  the 64 uneven jobs of the batched workload (generator seed 1000 + j,
      tight padding), each searched whole

For every search it records the FP64 winner (packed index, indices, score as
an exact hex float) and the order-independent checksum of every set's
2 x 3^k table (orc_table_hash summed mod 2^64, the quantity
episdet_checksum computes on the GPU).  That pins the reference's full
output at these sizes -- every contingency table (helper.cu:438-548) and the
best set (helper.cu:552-584, tensor-episdet.cu:181-184) -- against the
oracle, not against another GPU kernel.

Large searches run as slices of the largest set index (equal set counts),
each recorded as it finishes, so an interrupted run resumes where it
stopped.  CPU cost on 8 cores: w3 under a minute, w4 about 20 minutes, w2
about an hour, w5 about three hours.

    python tests/golden/make_fullsize.py [--only w2,w3,w4,w5] [--threads N]
"""
import argparse
import json
import math
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tensor-episdet_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

OUT = os.path.join(HERE, "fullsize.json")

# Generator parameters of each single-search workload (episdet.generate
# keywords; maf_dist 0 = uniform, 1 = Beta(0.5, 2) clipped).
WORKLOADS = {
    "w2": dict(order=3, score=0, n=4096, cases=131072, controls=131072, seed=1, planted_sets=1, maf=[0.05, 0.5],
               maf_dist=0, slices=16),
    "w3": dict(order=2, score=1, n=16384, cases=26214, controls=39322, seed=1, planted_sets=1, maf=[0.01, 0.5],
               maf_dist=1, slices=1),
    "w5": dict(order=3, score=0, n=8192, cases=32768, controls=32768, seed=1, planted_sets=2, maf=[0.05, 0.5],
               maf_dist=0, slices=32),
}
BATCH_SEED = 1      # the batched workload's job-list seed
BATCH_JOB_SEED0 = 1000


def batch_spec(seed=BATCH_SEED):
    """The 64 jobs of the batched workload: order and score 50/50, SNPs
    log-uniform in [256, 8192] (3-way [256, 2048]), samples uniform in
    [8000, 131072] split 50/50.  Same draws as the benchmark's job list."""
    import numpy as np
    rng = np.random.default_rng(seed)
    jobs = []
    for _ in range(64):
        order = 3 if rng.random() < 0.5 else 2
        hi = 2048 if order == 3 else 8192
        n = int(round(np.exp(rng.uniform(np.log(256), np.log(hi)))))
        N = int(rng.integers(8000, 131073))
        score = 1 if rng.random() < 0.5 else 0
        jobs.append((order, score, n, N // 2, N - N // 2))
    return jobs


def generate(w):
    import episdet
    return episdet.generate(w["n"], w["cases"], w["controls"], seed=w["seed"], planted_sets=w["planted_sets"],
                            planted_order=w["order"], maf=tuple(w["maf"]), maf_dist=w["maf_dist"])


def batch_generate(j, spec):
    import episdet
    order, score, n, Nc, Nt = spec[j]
    return episdet.generate(n, Nc, Nt, seed=BATCH_JOB_SEED0 + j, cases_words=(Nc + 31) // 32,
                            controls_words=(Nt + 31) // 32)


def slice_bounds(n, order, k):
    """k contiguous ranges of the largest index with about equal set counts."""
    total = math.comb(n, order)
    bounds, acc, lo = [], 0, order - 1
    for z in range(order - 1, n):
        acc += math.comb(z, order - 1)
        if acc >= total * (len(bounds) + 1) / k or z == n - 1:
            bounds.append((lo, z + 1))
            lo = z + 1
    return bounds


def better(score, a, b):
    """a better than b: (fp64 score, packed); K2 lower, MI higher, ties to the
    lower packed index (the oracle's rule)."""
    if a["score"] != b["score"]:
        return a["score"] < b["score"] if score == 0 else a["score"] > b["score"]
    return a["packed"] < b["packed"]


def load():
    if os.path.exists(OUT):
        with open(OUT) as f:
            return json.load(f)
    return {"about": "oracle results over the full benchmark workloads; written by make_fullsize.py",
            "checksum_rule": "sum over sets of orc_table_hash(colex rank, cases cells, controls cells) mod 2^64",
            "workloads": {}, "batch": None}


def save(fx):
    tmp = OUT + ".tmp"
    with open(tmp, "w") as f:
        json.dump(fx, f, indent=1)
    os.replace(tmp, OUT)


def record(res, order):
    return {"packed": res["packed"], "idx": list(res["idx"][:order]), "score": res["score"].hex(),
            "score_decimal": repr(res["score"]), "sets": res["sets"], "checksum": f"0x{res['checksum']:016X}"}


def run_single(name, fx, threads):
    import oracle
    w = WORKLOADS[name]
    ent = fx["workloads"].get(name)
    params = {k: v for k, v in w.items() if k != "slices"}
    if ent is None or ent["params"] != params:
        ent = {"params": params, "slices": [], "result": None}
        fx["workloads"][name] = ent
    if ent["result"] is not None:
        print(f"{name}: done", flush=True)
        return
    t0 = time.time()
    syn = generate(w)
    ent["planted"] = [list(p) for p in syn.planted]
    print(f"{name}: generated in {time.time() - t0:.1f} s, planted {syn.planted}", flush=True)
    bounds = slice_bounds(w["n"], w["order"], w["slices"])
    have = {tuple(s["last"]) for s in ent["slices"]}
    for lo, hi in bounds:
        if (lo, hi) in have:
            continue
        t = time.time()
        r = oracle.search(syn.cases, syn.controls, w["cases"], w["controls"], w["order"], w["score"],
                          threads=threads, last=(lo, hi), checksum=True)
        rec = record(r, w["order"])
        rec["last"] = [lo, hi]
        rec["seconds"] = round(time.time() - t, 1)
        ent["slices"].append(rec)
        save(fx)
        done = sum(s["sets"] for s in ent["slices"])
        print(f"{name}: largest index [{lo},{hi}) {r['sets']} sets in {rec['seconds']} s "
              f"({done / math.comb(w['n'], w['order']):.1%} done)", flush=True)
    # combine the slices
    best, cs, sets = None, 0, 0
    for s in ent["slices"]:
        cand = {"score": float.fromhex(s["score"]), "packed": s["packed"]}
        if best is None or better(w["score"], cand, best):
            best = dict(cand, idx=s["idx"])
        cs = (cs + int(s["checksum"], 16)) % 2**64
        sets += s["sets"]
    assert sets == math.comb(w["n"], w["order"]), (sets, math.comb(w["n"], w["order"]))
    ent["result"] = {"packed": best["packed"], "idx": best["idx"], "score": best["score"].hex(),
                     "score_decimal": repr(best["score"]), "sets": sets, "checksum": f"0x{cs:016X}",
                     "cpu_seconds": round(sum(s["seconds"] for s in ent["slices"]), 1), "threads": threads}
    save(fx)
    print(f"{name}: winner {best['idx']} {best['score']!r} checksum 0x{cs:016X}", flush=True)


def run_batch(fx, threads):
    import oracle
    spec = batch_spec()
    b = fx.get("batch")
    if b is None or b["spec"] != [list(s) for s in spec]:
        b = {"spec": [list(s) for s in spec], "job_seed0": BATCH_JOB_SEED0, "padding": "tight (ceil(N/32) words)",
             "jobs": {}}
        fx["batch"] = b
    for j in sorted(range(64), key=lambda j: math.comb(spec[j][2], spec[j][0]) * (spec[j][3] + spec[j][4])):
        if str(j) in b["jobs"]:
            continue
        order, score, n, Nc, Nt = spec[j]
        syn = batch_generate(j, spec)
        t = time.time()
        r = oracle.search(syn.cases, syn.controls, Nc, Nt, order, score, threads=threads, checksum=True)
        rec = record(r, order)
        rec["seconds"] = round(time.time() - t, 1)
        b["jobs"][str(j)] = rec
        save(fx)
        print(f"batch job {j}: {order}-way {'K2' if score == 0 else 'MI'} {n} x {Nc + Nt} in {rec['seconds']} s "
              f"-> {rec['idx']}", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", default="w3,w4,w2,w5")
    ap.add_argument("--threads", type=int, default=0)
    a = ap.parse_args()
    threads = a.threads or len(os.sched_getaffinity(0))
    fx = load()
    for name in a.only.split(","):
        if name == "w4":
            run_batch(fx, threads)
        else:
            run_single(name, fx, threads)


if __name__ == "__main__":
    main()
