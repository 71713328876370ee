"""Dataset format (tensor-episdet.cu:58-148) and the seeded generator."""
import os

import numpy as np
import pytest

import episdet
import oracle as O


def test_generator_is_deterministic_and_thread_independent():
    a = episdet.generate(70, 1500, 1700, seed=42, planted_sets=1, threads=1)
    b = episdet.generate(70, 1500, 1700, seed=42, planted_sets=1, threads=8)
    assert (a.cases == b.cases).all() and (a.controls == b.controls).all()
    assert a.planted == b.planted
    c = episdet.generate(70, 1500, 1700, seed=43, planted_sets=1)
    assert not (a.cases == c.cases).all()


@pytest.mark.parametrize("Nc,Nt,wc,wt", [(1024, 1024, 0, 0), (26214, 39322, 0, 0), (45, 70, 0, 0),
                                         (9000, 8001, 290, 251)])
def test_padding_layout_and_validity(Nc, Nt, wc, wt):
    syn = episdet.generate(20, Nc, Nt, seed=1, cases_words=wc, controls_words=wt)
    assert syn.cases.shape == (20, 2, wc or episdet.padded_words(Nc))
    assert syn.controls.shape == (20, 2, wt or episdet.padded_words(Nt))
    assert O.check_planes(syn.cases, Nc) == 0
    assert O.check_planes(syn.controls, Nt) == 0
    s = O.singles(syn.cases, Nc)
    assert (s.sum(axis=1) == Nc).all()


def test_reference_padding_words():
    assert episdet.padded_words(131072) == 4096
    # ceil(26214/1024)*1024 = 26624 bits (tensor-episdet.cu:94)
    assert episdet.padded_words(26214) == 832
    assert episdet.padded_words(39322) == 1248       # 39936 bits
    assert episdet.padded_words(1) == 32


def test_hwe_genotype_frequencies():
    syn = episdet.generate(200, 20000, 20000, seed=5, maf=(0.05, 0.5))
    s = O.singles(syn.cases, 20000).astype(np.float64) + O.singles(syn.controls, 20000)
    f = s / 40000.0
    # allele frequency estimate, and the HWE heterozygote relation
    p = (2 * f[:, 2] + f[:, 1]) / 2
    assert p.min() > 0.03 and p.max() < 0.52
    assert np.abs(f[:, 1] - 2 * p * (1 - p)).max() < 0.02


def test_beta_maf_is_clipped():
    syn = episdet.generate(400, 4000, 4000, seed=2, maf_dist=episdet.MAF_BETA, beta=(0.5, 2.0), maf=(0.01, 0.5))
    s = O.singles(syn.cases, 4000).astype(np.float64) + O.singles(syn.controls, 4000)
    p = (2 * s[:, 2] + s[:, 1]) / (2 * 8000.0)
    assert p.max() < 0.53
    assert (p < 0.05).mean() > 0.2           # Beta(0.5, 2) puts much mass near 0


def test_planted_sets_and_penetrance():
    syn = episdet.generate(300, 5000, 5000, seed=9, planted_sets=2, planted_order=3)
    assert len(syn.planted) == 2
    flat = [v for s in syn.planted for v in s]
    assert len(set(flat)) == 6 and all(0 <= v < 300 for v in flat)
    assert all(list(s) == sorted(s) for s in syn.planted)
    pen = syn.penetrance
    assert pen.shape == (2, 27)
    assert np.isclose(pen.min(), 0.1) and np.isclose(pen.max(), 0.4)
    assert np.isclose(pen[:, 26], 0.4).all()


def test_file_round_trip(tmp_path):
    syn = episdet.generate(37, 900, 1100, seed=4)
    txt = episdet.write_dataset(tmp_path, "db", syn)
    lines = open(txt).read().split("\n")
    assert lines[:5] == ["37", "db.controls.bin", "1100", "db.cases.bin", "900"]
    assert os.path.getsize(tmp_path / "db.cases.bin") == 37 * 2 * 32 * 4
    back = episdet.read_dataset(txt)
    assert (back.cases == syn.cases).all() and (back.controls == syn.controls).all()
    assert (back.num_cases, back.num_controls) == (900, 1100)


def test_short_or_missing_bin_is_an_error(tmp_path):
    syn = episdet.generate(10, 100, 100, seed=4)
    txt = episdet.write_dataset(tmp_path, "db", syn)
    with open(tmp_path / "db.cases.bin", "r+b") as f:
        f.truncate(100)
    with pytest.raises(episdet.EpisdetError):
        episdet.read_dataset(txt)
    os.remove(tmp_path / "db.controls.bin")
    with pytest.raises(episdet.EpisdetError):
        episdet.read_dataset(txt)


def test_malformed_header_is_an_error(tmp_path):
    p = tmp_path / "bad.txt"
    p.write_text("12\nx.bin\n")
    with pytest.raises(episdet.EpisdetError):
        episdet.read_dataset(str(p))
