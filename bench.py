"""Benchmark: exhaustive 3-way K2 search, sets x samples per second.

Workload (default): 4096 SNPs x 262144 samples (131072 cases / 131072
controls), HWE genotypes with MAF ~ U[0.05, 0.5], one planted triplet with a
seeded 27-cell penetrance table (baseline 0.1, peak 0.4), seed 1 -- the shape
of the reference's published run (README.md:54-60).  A step is one complete
search over all C(4096,3) = 11,444,858,880 triplets with the data already
resident in HBM: pair-table build, the fused persistent kernel, FP64
re-scoring of candidates and the final reduction.

value      = C(n,3) * (cases + controls) / seconds per step  (the reference's
             metric, tensor-episdet.cu:198, unpadded samples); seconds per
             step = the K timed steps, bracketed by device synchronisations,
             over K
roofline   = the fused kernel alone, from HIP events on its stream:
             FP4-MFMA path (the default below 2^24 samples per class):
             achieved = 16 flop per 3-way set-sample (8 multiply-adds, one
             per AND-count cell; 8 flop for 2-way) x set-samples / kernel
             seconds, against MI355X's dense FP4 peak of 10.1 PFLOP/s;
             VALU path: 0.5 lane-ops per 3-way set-sample (one v_and_b32 and
             one v_bcnt_u32_b32 per cell per 32 samples, 8 cells) against
             256 CU x 64 lanes x 2.4 GHz
null_data  = the same search on the same shape with no planted set (the
             score-bound skip depends on how strong the best set is), one
             step, reported beside the headline
plain      = the reference-shaped round loop (count kernel -> int32 tables in
             HBM -> FP64 objective kernel), timed on its first R rounds and
             scaled by the round count (all rounds have identical work), with
             its rounds counted by the FP4 MFMA kernel (mfma_count: fused /
             plain is then the gain of fusion alone) and on the VALU
             (valu_count)
cpu        = the OpenMP AND+popcount oracle on a contiguous slice of the same
             search (sets whose largest index lies in a stated range), on
             every host CPU this process may use, at most the OMP_NUM_THREADS
             share (16 per GPU on the MI355X boxes)

python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 1|2|3|5] [--path mfma|valu]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tensor-episdet_amd"))

VALU_PEAK_OPS = 256 * 64 * 2.4e9        # 32-bit VALU lane-ops/s, MI355X
OPS_PER_SET_SAMPLE = {2: 0.25, 3: 0.5}   # VALU: AND + BCNT per cell per 32 samples
MFMA_FP4_PEAK = 10.1e15                  # dense FP4 MFMA flop/s, MI355X
FLOPS_PER_SET_SAMPLE = {2: 8.0, 3: 16.0} # MFMA: 2^k multiply-adds per set-sample


SCORE_NAME = {0: "K2", 1: "MI"}
# Single-search workload presets 1, 2, 3 and 5 (preset 4, the 64-job batch,
# is not a single search).  Preset 3's classes are 26214 / 39322 of 65536.
PRESETS = {
    1: dict(order=2, score=0, n=512, samples=2048, case_frac=0.5, planted=0, maf=(0.05, 0.5), maf_dist=0,
            maf_text="U[0.05,0.5]"),
    2: dict(order=3, score=0, n=4096, samples=262144, case_frac=0.5, planted=1, maf=(0.05, 0.5), maf_dist=0,
            maf_text="U[0.05,0.5]"),
    3: dict(order=2, score=1, n=16384, samples=65536, case_frac=26214 / 65536, planted=1, maf=(0.01, 0.5),
            maf_dist=1, maf_text="Beta(0.5,2) clipped to [0.01,0.5]"),
    5: dict(order=3, score=0, n=8192, samples=65536, case_frac=0.5, planted=2, maf=(0.05, 0.5), maf_dist=0,
            maf_text="U[0.05,0.5]"),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_threads():
    """CPU threads for the oracle baseline: every CPU this process may run on,
    capped by the OMP_NUM_THREADS share when one is set; and the host's
    counts, reported beside it."""
    aff = len(os.sched_getaffinity(0))
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(aff, share) if share > 0 else aff
    return threads, {"host_cpus": os.cpu_count(), "affinity_cpus": aff, "omp_num_threads": share or None}


def measured_traffic(n, Nc, Nt, order, score, path, tile):
    """HBM bytes per launch of the dominant kernel for this exact workload, as
    recorded from the FETCH_SIZE / WRITE_SIZE counter passes (calibrated
    against the copy kernel) in profiles/traffic.json; None if that
    workload was not profiled."""
    f = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic.json")
    if not os.path.exists(f):
        return None
    with open(f) as fh:
        for e in json.load(fh)["entries"]:
            if (e["snps"], e["cases"], e["controls"], e["order"], e["score"], e["path"], e["tile"]) == \
                    (n, Nc, Nt, order, SCORE_NAME[score], path, tile):
                return e["bytes"]
    return None


def batch_jobs(seed):
    """Preset 4, the 64-job batch: 64 independent jobs; order and score drawn 50/50,
    SNPs log-uniform in [256, 8192] (3-way in [256, 2048]), samples uniform in
    [8000, 131072] split 50/50, packed tight (ceil(N/32) words, not the
    1024-sample padding), MAF ~ U[0.05, 0.5]."""
    import numpy as np
    rng = np.random.default_rng(seed)
    jobs = []
    for j in range(64):
        order = 3 if rng.random() < 0.5 else 2
        hi = 2048 if order == 3 else 8192
        n = int(round(np.exp(rng.uniform(np.log(256), np.log(hi)))))
        N = int(rng.integers(8000, 131073))
        score = 1 if rng.random() < 0.5 else 0
        jobs.append((order, score, n, N // 2, N - N // 2))
    return jobs


def bench_batch(args, episdet):
    spec = batch_jobs(args.seed)
    t = time.time()
    syns, dss = [], []
    for j, (order, score, n, Nc, Nt) in enumerate(spec):
        syn = episdet.generate(n, Nc, Nt, seed=1000 + j, cases_words=(Nc + 31) // 32,
                               controls_words=(Nt + 31) // 32)
        dss.append(episdet.Dataset.from_synthetic(syn))
        syns.append(syn)
    log(f"generated and uploaded 64 jobs in {time.time() - t:.1f} s")
    # autotune every job's shape once (untimed setup; the choice is cached
    # for the process and the batch runs it)
    t = time.time()
    picks = {}
    for d, (order, score, n, Nc, Nt) in zip(dss, spec):
        table, chosen = d.autotune(order, score)
        e = table[chosen]
        key = f"{order}-way {'mfma' if e['path'] == episdet.PATH_FUSED_MFMA else 'valu'} tile {e['tile']}"
        picks[key] = picks.get(key, 0) + 1
    log(f"autotuned 64 shapes in {time.time() - t:.1f} s: {picks}")
    orders = [s[0] for s in spec]
    scores = [s[1] for s in spec]
    setsamp = sum(episdet.n_choose_k(n, o) * (Nc + Nt) for o, _, n, Nc, Nt in spec)
    tri_ss = sum(episdet.n_choose_k(n, o) * (Nc + Nt) for o, _, n, Nc, Nt in spec if o == 3)
    for i in range(args.warmup):
        res, ms = episdet.run_batch(dss, orders, scores, rebuild=True)
        log(f"warmup {i}: graph {ms:.1f} ms")
    times, gms = [], []
    episdet.synchronize()
    t_all = time.perf_counter()
    for i in range(args.steps):
        t0 = time.perf_counter()
        res, ms = episdet.run_batch(dss, orders, scores, rebuild=True)
        times.append(time.perf_counter() - t0)
        gms.append(ms)
    episdet.synchronize()
    sec = (time.perf_counter() - t_all) / args.steps
    for i, (dt, ms) in enumerate(zip(times, gms)):
        log(f"step {i}: {dt:.3f} s (graph {ms:.1f} ms)")
    g = sum(gms) / len(gms) / 1e3
    outside = sum(1 for r in res if r["kernel_ms"] != ms)    # jobs that ran after the graph
    if outside:
        log(f"{outside} jobs ran outside the graph; roofline over the whole step")
        g = sec
    achieved = (16.0 * tri_ss + 8.0 * (setsamp - tri_ss)) / g
    cpu = None
    if not args.no_cpu:
        # the smallest jobs, whole, one after another, until about --cpu-seconds
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        threads, hostinfo = host_threads()
        order_by_size = sorted(range(len(spec)), key=lambda j: episdet.n_choose_k(spec[j][2], spec[j][0]) *
                               (spec[j][3] + spec[j][4]))
        done, ss, t0 = [], 0, time.perf_counter()
        for j in order_by_size:
            o, sc_, n_, Nc_, Nt_ = spec[j]
            oracle.search(syns[j].cases, syns[j].controls, Nc_, Nt_, o, sc_, threads=threads)
            done.append(j)
            ss += episdet.n_choose_k(n_, o) * (Nc_ + Nt_)
            if time.perf_counter() - t0 > args.cpu_seconds:
                break
        dt = time.perf_counter() - t0
        cpu = {"value": ss / dt, "unit": "set-samples/s", "cores": threads, "kind": "port", **hostinfo,
               "sample": f"oracle (OpenMP AND+popcount, FP64 K2/MI): the {len(done)} smallest of the 64 jobs, "
                         f"whole, {dt:.1f} s"}
        log(f"cpu: {len(done)} jobs in {dt:.1f} s on {threads} threads")
    out = {
        "metric": "batched 2-/3-way K2/MI searches, total sets x samples/s (64 uneven jobs, one graph launch)",
        "value": setsamp / sec, "unit": "set-samples/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": sec * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp4",
        "dtype_detail": "fp4 e2m1 0/1 MFMA operands, fp32 exact-integer accumulation (2- and 3-way); FP32 scoring, "
                        "FP64 re-score",
        "data": "synthetic",
        "config": {"workload": "64 jobs: order/score 50/50, SNPs log-U[256,8192] (3-way log-U[256,2048]), samples "
                               "U[8000,131072] split 50/50, tight padding, MAF~U[0.05,0.5]",
                   "jobs": [{"order": o, "score": SCORE_NAME[sc], "snps": n, "cases": Nc, "controls": Nt}
                            for o, sc, n, Nc, Nt in spec],
                   "set_samples": setsamp, "autotuned": picks, "parallelism": "1 GPU"},
        "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": MFMA_FP4_PEAK / 1e12,
                     "unit": "TFLOP/s",
                     "work": "FP4 MFMA work (16 flop per 3-way, 8 per 2-way set-sample) over the graph's device time",
                     "frac": achieved / MFMA_FP4_PEAK, "traffic": None, "kernel_ms": g * 1e3},
        "cpu_baseline": cpu, "plain_path": None,
        "best": [{"idx": list(r["idx"][:o]), "score": r["best_score"], "path": r["path"]}
                 for r, o in zip(res, orders)],
    }
    print(json.dumps(out), flush=True)
    for d in dss:
        d.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--config", type=int, default=2, choices=(1, 2, 3, 4, 5),
                    help="workload preset; 2 (3-way K2, 4096 SNPs x 262144 samples) is the headline")
    ap.add_argument("--n", type=int, default=0, help="override the preset's SNP count")
    ap.add_argument("--samples", type=int, default=0, help="override the preset's sample count")
    ap.add_argument("--tile", type=int, default=0, help="kernel tile id (with --path mfma or valu)")
    ap.add_argument("--path", choices=("auto", "mfma", "valu"), default="auto",
                    help="auto: the library's choice for this shape (FP4 MFMA below 2^24 samples per class)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--plain-rounds", type=int, default=2048)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-null", action="store_true", help="skip the no-planted-set timing")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the search shards with no "
              "communication between GPUs (each GPU takes a contiguous range of work units)", file=sys.stderr)
        sys.exit(2)

    import episdet

    if args.config == 4:
        return bench_batch(args, episdet)
    pre = PRESETS[args.config]
    n = args.n or pre["n"]
    N = args.samples or pre["samples"]
    Nc = int(round(N * pre["case_frac"]))
    Nt = N - Nc
    order, score = pre["order"], pre["score"]
    t = time.time()
    syn = episdet.generate(n, Nc, Nt, seed=args.seed, planted_sets=pre["planted"], planted_order=order,
                           maf=pre["maf"], maf_dist=pre["maf_dist"])
    log(f"generated {n} SNPs x {N} samples in {time.time() - t:.1f} s; planted {syn.planted}")
    name, cu, clk = episdet.device_info()
    log(f"device {name}, {cu} CUs, {clk / 1e3:.0f} MHz")
    ds = episdet.Dataset.from_synthetic(syn)
    sets = episdet.n_choose_k(n, order)

    path = {"auto": episdet.PATH_AUTO, "mfma": episdet.PATH_FUSED_MFMA, "valu": episdet.PATH_FUSED_VALU}[args.path]

    def step():
        return ds.run(order, score, path=path, tile=args.tile, rebuild=True)

    for i in range(args.warmup):
        r = step()
        log(f"warmup {i}: {r['wall_s']:.3f} s, kernel {r['kernel_ms']:.1f} ms")
    times, kms, results = [], [], []
    episdet.synchronize()
    t_all = time.perf_counter()
    for i in range(args.steps):
        t0 = time.perf_counter()
        r = step()
        times.append(time.perf_counter() - t0)
        kms.append(r["kernel_ms"])
        results.append(r)
    episdet.synchronize()
    sec = (time.perf_counter() - t_all) / args.steps
    for i, (dt, r) in enumerate(zip(times, results)):
        log(f"step {i}: {dt:.3f} s (kernel {r['kernel_ms']:.1f} ms, pairs {r['prep_ms']:.1f} ms, rescore "
            f"{r['rescore_ms']:.2f} ms, {r['candidates']} candidates) best {r['idx'][:order]} {r['best_score']:.6f}")
    value = sets * N / sec
    kernel_s = sum(kms) / len(kms) / 1e3
    ran = "mfma" if results[-1]["path"] == episdet.PATH_FUSED_MFMA else "valu"
    if ran == "mfma":
        achieved = FLOPS_PER_SET_SAMPLE[order] * sets * N / kernel_s
        roof = {"bound": "mfma", "achieved": achieved / 1e12, "peak": MFMA_FP4_PEAK / 1e12,
                "unit": "TFLOP/s", "work": f"dense FP4 MFMA, {FLOPS_PER_SET_SAMPLE[order]:.0f} flop per set-sample",
                "frac": achieved / MFMA_FP4_PEAK}
    else:
        achieved = OPS_PER_SET_SAMPLE[order] * sets * N / kernel_s
        roof = {"bound": "valu", "achieved": achieved / 1e12, "peak": VALU_PEAK_OPS / 1e12,
                "unit": "T lane-ops/s", "work": f"32-bit VALU, {OPS_PER_SET_SAMPLE[order]} per set-sample",
                "frac": achieved / VALU_PEAK_OPS}
    r = results[-1]
    roof.update({"traffic": measured_traffic(n, Nc, Nt, order, score, "fused_" + ran, r["tile"]),
                 "kernel_ms": kernel_s * 1e3})
    planted_ok = (tuple(r["idx"][:order]) in syn.planted) if syn.planted else None

    plain = None
    if not args.no_plain:
        # the same round loop with each of its two count kernels: FP4 MFMA
        # (the fused path's counting engine, so fused / plain is the gain of
        # fusion alone) and AND + v_bcnt on the VALU
        plain = {}
        for name, tile in (("mfma_count", 1), ("valu_count", 0)):
            try:
                rr = ds.run(order, score, path=episdet.PATH_PLAIN, tile=tile, plain_rounds=args.plain_rounds)
            except episdet.EpisdetError:                # FP4 counts not exact at this N
                continue
            total_rounds = rr["rounds"]
            est = rr["kernel_ms"] / 1e3 * total_rounds / rr["rounds_run"]
            plain[name] = {"value": sets * N / est, "rounds_timed": rr["rounds_run"], "rounds_total": total_rounds,
                           "seconds_est": est, "dtype": "fp4 count" if tile == 1 else "u32 count",
                           "scoring": "f64", "speedup_fused": value / (sets * N / est)}
            log(f"plain ({name}): {rr['rounds_run']} rounds in {rr['kernel_ms']:.1f} ms -> {est:.1f} s estimated")

    null = None
    if not args.no_null and pre["planted"] and order == 3:
        # the same shape and generator seed with nothing planted: as many
        # warmups as the headline, then one timed step
        ds.close()
        nsyn = episdet.generate(n, Nc, Nt, seed=args.seed, planted_sets=0, maf=pre["maf"], maf_dist=pre["maf_dist"])
        ds = episdet.Dataset.from_synthetic(nsyn)
        for i in range(args.warmup):
            step()
        t0 = time.perf_counter()
        nr = step()
        ndt = time.perf_counter() - t0
        null = {"value": sets * N / ndt, "ms_per_step": ndt * 1e3, "kernel_ms": nr["kernel_ms"],
                "candidates": nr["candidates"], "best": list(nr["idx"][:order]), "best_score": nr["best_score"],
                "data": "same shape and generator seed, no planted set"}
        log(f"null data: {ndt:.3f} s (kernel {nr['kernel_ms']:.1f} ms) best {nr['idx'][:order]}")

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        threads, hostinfo = host_threads()
        # Calibrate a slice of the search (the sets whose largest index is in
        # [z0, z0+zn)) to about --cpu-seconds.  A set's cost is the same
        # everywhere, and C(Z, order-1) sets have largest index Z, so two
        # probes give the marginal cost per set (each probe also pays for its
        # own setup, which cancels out).  The oracle parallelises over the
        # largest index, so each probe spans a multiple of the thread count.
        def sets_in(lo, hi):
            return sum(math.comb(z, order - 1) for z in range(lo, hi))

        def span(lo, want):
            # smallest multiple of the thread count covering `want` sets
            hi, acc = lo, 0
            while hi < n and (acc < want or (hi - lo) % threads):
                acc += math.comb(hi, order - 1)
                hi += 1
            return hi

        if math.comb(n, order) <= 5_000_000:
            # the whole search is cheap: time it in full, repeated to about a second
            reps, t0 = 0, time.perf_counter()
            while reps == 0 or time.perf_counter() - t0 < 1.0:
                c = oracle.search(syn.cases, syn.controls, Nc, Nt, order, score, threads=threads)
                reps += 1
            dt = time.perf_counter() - t0
            cpu = {"value": reps * c["sets"] * N / dt, "unit": "set-samples/s", "cores": threads, "kind": "port",
                   **hostinfo,
                   "sample": f"oracle (OpenMP AND+popcount, FP64 {SCORE_NAME[score]}): the whole search "
                             f"({c['sets']} sets), {reps} times, {dt:.1f} s"}
            log(f"cpu: whole search x{reps} in {dt:.1f} s on {threads} threads")
        z0 = max(order, min(n // 4, 600))
        probe_t = []
        ends = (span(z0, 1_000_000), span(z0, 3_000_000)) if cpu is None else ()
        for hi in ends:
            t0 = time.perf_counter()
            oracle.search(syn.cases, syn.controls, Nc, Nt, order, score, threads=threads, last=(z0, hi))
            probe_t.append(time.perf_counter() - t0)
        if cpu is None:
            s1, s2 = sets_in(z0, ends[0]), sets_in(z0, ends[1])
            per_set = max((probe_t[1] - probe_t[0]) / max(s2 - s1, 1), probe_t[0] / max(s1, 1) / 4, 1e-12)
            target = args.cpu_seconds / per_set
            zn, acc = 0, 0
            while z0 + zn < n and (zn < threads or acc < target):
                acc += math.comb(z0 + zn, order - 1)
                zn += 1
            log(f"cpu: probes {probe_t[0]:.2f} s / {probe_t[1]:.2f} s, timing largest index in [{z0},{z0 + zn})")
            t0 = time.perf_counter()
            c = oracle.search(syn.cases, syn.controls, Nc, Nt, order, score, threads=threads, last=(z0, z0 + zn))
            dt = time.perf_counter() - t0
            cpu = {"value": c["sets"] * N / dt, "unit": "set-samples/s", "cores": threads, "kind": "port",
                   **hostinfo, "sample": f"oracle (OpenMP AND+popcount{', AVX-512 VPOPCNTDQ' if oracle.uses_avx512() else ''}, "
                             f"FP64 {SCORE_NAME[score]}) on the {c['sets']} sets whose largest index is in "
                             f"[{z0},{z0 + zn}), {dt:.1f} s"}
            log(f"cpu: {c['sets']} sets in {dt:.1f} s on {threads} threads")

    out = {
        "metric": f"{order}-way {SCORE_NAME[score]} sets x samples/s (fused FP32-scored search, end to end, "
                  "data resident in HBM)",
        "value": value,
        "unit": "set-samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp4" if ran == "mfma" else "u32",
        "dtype_detail": ("fp4 e2m1 0/1 operands, fp32 exact-integer accumulation" if ran == "mfma" else
                         "u32 bit-plane AND+popcount") + "; exact int counts, FP32 scoring, FP64 re-score",
        "data": "synthetic",
        "config": {"workload": f"{order}-way {SCORE_NAME[score]} exhaustive search, {n} SNPs x {N} samples "
                               f"({Nc}/{Nt}), HWE, MAF~{pre['maf_text']}, {pre['planted']} planted "
                               f"{'triplet' if order == 3 else 'pair'}(s), seed {args.seed}",
                   "preset": args.config,
                   "snps": n, "cases": Nc, "controls": Nt, "sets": sets, "path": "fused_" + ran,
                   "tile": r["tile"], "parallelism": "1 GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "plain_path": plain,
        "null_data": null,
        "best": {"idx": list(r["idx"][:order]), "score": r["best_score"], "planted_recovered": planted_ok,
                 "candidates": r["candidates"], "fp32_max_rel_err": r["fp32_max_rel_err"]},
    }
    print(json.dumps(out), flush=True)
    ds.close()


if __name__ == "__main__":
    main()
