"""Writes tests/golden/tiny.json: a tiny dataset with every pair and triplet
table and their K2 / MI scores, computed without any code of this project.

The reference (hiperbio/tensor-episdet) ships no tests, fixtures or golden
vectors, and cannot be built or run here (CUDA + CUTLASS 1.3, NVIDIA GPU), so
these vectors are made from first principles:

  * genotypes are drawn with numpy and packed by explicit Python bit
    arithmetic into the reference's plane layout: per class u32
    [snp][plane][W], plane 0 = genotype 0, plane 1 = genotype 1, sample s at
    bit 31 - s%32 of word s/32, W = 32 * ceil(N / 1024)
    (tensor-episdet.cu:94-95,111-148; helper.cu:72-73);
  * tables are counted per sample, cell index x*9+y*3+z (helper.cu:484);
  * K2 = sum over cells of ln((a+b+1)! / (a! b!)), evaluated exactly as
    ln prod_c (a_c+b_c+1) * C(a_c+b_c, a_c) on Python integers
    (the quantity of scoring.hpp:5-7);
  * MI = I(genotype cell; class) in bits from sklearn's mutual_info_score on
    the contingency table (the quantity of scoring.hpp:9-16,37);
  * best sets: K2 minimised, MI maximised, ties to the lowest packed index
    X | Y<<21 | Z<<42 (A | B<<32).

Run: python tests/golden/make_golden.py
"""
import itertools
import json
import math
import os

import numpy as np
from sklearn.metrics import mutual_info_score

SEED = 20240601
N_SNPS = 10
N_CASES, N_CONTROLS = 45, 70        # not multiples of 32: exercises the last-word mask


def pack(geno):
    """geno [snps, N] in {0,1,2} -> uint32 [snps, 2, W] (MSB-first)."""
    n, N = geno.shape
    W = 32 * ((N + 1023) // 1024)
    out = np.zeros((n, 2, W), dtype=np.uint64)
    for v in range(n):
        for s in range(N):
            g = int(geno[v, s])
            if g < 2:
                out[v, g, s // 32] |= 1 << (31 - s % 32)
    return out.astype(np.uint32)


def table(gc, gt, snps):
    nc = 3 ** len(snps)
    cas = [0] * nc
    ctl = [0] * nc
    for g, t in ((gc, cas), (gt, ctl)):
        for s in range(g.shape[1]):
            c = 0
            for v in snps:
                c = c * 3 + int(g[v, s])
            t[c] += 1
    return cas, ctl


def k2_exact(cas, ctl):
    prod = 1
    for a, b in zip(cas, ctl):
        prod *= (a + b + 1) * math.comb(a + b, a)
    return math.log(prod)


def mi_bits(cas, ctl):
    cont = np.array([cas, ctl], dtype=np.int64)
    cont = cont[:, cont.sum(axis=0) > 0]
    return mutual_info_score(None, None, contingency=cont) / math.log(2)


def pack_index(snps):
    if len(snps) == 3:
        return snps[0] | (snps[1] << 21) | (snps[2] << 42)
    return snps[0] | (snps[1] << 32)


def main():
    rng = np.random.default_rng(SEED)
    maf = rng.uniform(0.05, 0.5, N_SNPS)
    maf[3] = 0.0                     # a monomorphic SNP: every sample genotype 0
    probs = np.stack([(1 - maf) ** 2, 2 * maf * (1 - maf), maf ** 2], axis=1)
    gc = np.array([rng.choice(3, N_CASES, p=p) for p in probs], dtype=np.uint8)
    gt = np.array([rng.choice(3, N_CONTROLS, p=p) for p in probs], dtype=np.uint8)
    # an interacting pair/triplet: SNP 7 copies SNP 2 in cases (ties with it)
    gc[7] = gc[2]
    # a duplicated SNP in both classes makes exact score ties
    gc[9], gt[9] = gc[5], gt[5]

    out = {
        "seed": SEED, "num_snps": N_SNPS, "num_cases": N_CASES, "num_controls": N_CONTROLS,
        "cases": pack(gc).tolist(), "controls": pack(gt).tolist(),
        "genotypes_cases": gc.tolist(), "genotypes_controls": gt.tolist(),
    }
    for order in (2, 3):
        sets = []
        for snps in itertools.combinations(range(N_SNPS), order):
            cas, ctl = table(gc, gt, snps)
            sets.append({"snps": list(snps), "cases": cas, "controls": ctl,
                         "k2": k2_exact(cas, ctl), "mi": mi_bits(cas, ctl), "packed": pack_index(snps)})
        best_k2 = min(sets, key=lambda s: (s["k2"], s["packed"]))
        best_mi = min(sets, key=lambda s: (-s["mi"], s["packed"]))
        out[f"order{order}"] = {"sets": sets, "best_k2": best_k2["snps"], "best_mi": best_mi["snps"]}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tiny.json")
    with open(path, "w") as f:
        json.dump(out, f)
    print("wrote", path)


if __name__ == "__main__":
    main()
