"""The drop-in CLI binaries on the GPU: stdout lines of the reference's main
(tensor-episdet.cu:99-102,150,164,181-201) and the same winner as the oracle."""
import os
import re
import subprocess

import pytest

import episdet
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("order,score", [(3, "k2"), (3, "mi"), (2, "k2"), (2, "mi")])
def test_cli_output(tmp_path, order, score):
    syn = episdet.generate(300, 1500, 1400, seed=31, planted_sets=1, planted_order=order)
    txt = episdet.write_dataset(tmp_path, "db", syn)
    exe = os.path.join(episdet.BIN_DIR, f"tensor-episdet.{'triplets' if order == 3 else 'pairs'}.{score}.bin")
    print("cli", exe, flush=True)
    p = subprocess.run([exe, txt], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    lines = p.stdout.splitlines()
    assert lines[0] == "Num. SNPs: 300"
    assert lines[1] == "Num. Blocks of SNPs: 2"
    assert lines[2] == "Num. Cases: 1500"
    assert lines[3] == "Num. Controls: 1400"
    assert lines[4] == "-" * 31
    assert lines[5:7] == ["Outer loop iteration 0 out of 2", "Outer loop iteration 1 out of 2"]
    assert lines[7] == "-" * 31
    if order == 3:
        m = re.fullmatch(r"\{SNP_X_i, SNP_Y_i, SNP_Z_i\}: SCORE\t->\t\{(\d+), (\d+), (\d+)\}: (-?\d+\.\d{6})", lines[8])
    else:
        m = re.fullmatch(r"\{SNP_X_i, SNP_Y_i\}: SCORE\t->\t\{(\d+), (\d+)\}: (-?\d+\.\d{6})", lines[8])
    assert m, lines[8]
    idx = tuple(int(v) for v in m.groups()[:order])
    o = O.search(syn.cases, syn.controls, 1500, 1400, order, O.K2 if score == "k2" else O.MI)
    assert idx == tuple(o["idx"])
    # printed with 6 decimals (tensor-episdet.cu:182), from a float
    assert float(m.groups()[-1]) == pytest.approx(o["score"], rel=1e-6, abs=5e-7)
    rounds = episdet.reference_rounds(order, 300)
    assert lines[9] == f"Num. of rounds processed: {rounds}"
    assert re.fullmatch(r"Wall-clock time:\t\d+\.\d{3} seconds", lines[10])
    assert re.fullmatch(r"Tensor TOPS: \d+\.\d{3}", lines[11])
    assert re.fullmatch(r"Num\. unique sets per sec\. \(scaled to sample size\): \d+\.\d{3} × 10\^12", lines[12])
    assert lines[13] == f"Unique sets of SNPs evaluated (k={order}): {episdet.n_choose_k(300, order)}"
    assert re.fullmatch(r"Ratio of unique sets of SNPs: \d+\.\d{3}", lines[14])


def test_cli_errors(tmp_path):
    exe = os.path.join(episdet.BIN_DIR, "tensor-episdet.triplets.k2.bin")
    p = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert p.returncode == 1 and "Usage" in p.stderr
    p = subprocess.run([exe, str(tmp_path / "nope.txt")], capture_output=True, text=True, timeout=60)
    assert p.returncode == 1 and "does not exist" in p.stderr


SHIM_BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "shim", "bin")


@pytest.mark.parametrize("variant,order,score", [("triplets.k2", 3, O.K2), ("triplets.mi", 3, O.MI),
                                                  ("pairs.k2", 2, O.K2), ("pairs.mi", 2, O.MI)])
def test_reference_entry_point_through_shim(variant, order, score):
    """A C program written against EpistasisDetectionSearch
    (search-triplets.hpp:4), linked with tests/shim/episdet_shim.cpp (the
    binding INTEGRATION.md shows) instead of the reference's CUDA objects:
    the library prints the reference's "Outer loop iteration i out of nb"
    lines itself (search-triplets.cu:164, search-pairs.cu:139), before the
    program's own output, and the winner is the oracle's."""
    exe = os.path.join(SHIM_BIN, f"shim_main.{variant}")
    n, Nc, Nt, seed = 600, 1500, 1300, 33
    print("shim", exe, flush=True)
    p = subprocess.run([exe, str(n), str(Nc), str(Nt), str(seed)], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    lines = p.stdout.splitlines()
    assert lines[:3] == [f"Outer loop iteration {i} out of 3" for i in range(3)]
    syn = episdet.generate(n, Nc, Nt, seed=seed, planted_sets=1, planted_order=order)
    o = O.search(syn.cases, syn.controls, Nc, Nt, order, score)
    want = ", ".join(str(v) for v in o["idx"][:order])
    assert lines[3].split("\t->\t")[1].startswith("{" + want + "}: "), lines[3]
    assert float(lines[3].rsplit(": ", 1)[1]) == pytest.approx(o["score"], rel=1e-6, abs=5e-7)
    assert lines[4] == f"Num. of rounds processed: {episdet.reference_rounds(order, n)}"
    assert lines[5] == "planted: " + " ".join(str(v) for v in syn.planted[0])
