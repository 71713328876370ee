"""GPU: episdet_run_noalloc, the search with no device allocation, against
episdet_run on the same data, through the C ABI.

Its per-search device memory is the library's static scratch array, which
every call on a device shares; each check therefore runs two calls in a row,
the second on a smaller dataset, so that what the first left in the array
cannot leak into the second.  The answer must be episdet_run's: the same
winner (packed index) and the bit-identical FP64 score.
"""
import sys

import numpy as np
import pytest

import episdet

pytestmark = pytest.mark.gpu

M, V = episdet.PATH_FUSED_MFMA, episdet.PATH_FUSED_VALU
SCRATCH_ERR = 8          # EPISDET_ERR_SCRATCH


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def expected(syn, order, score, path, tile):
    """episdet_run on a dataset of its own, so the datasets under test keep
    no tables from it."""
    with episdet.Dataset.from_synthetic(syn) as ds:
        return ds.run(order, score, path=path, tile=tile)


def same(r, want):
    assert r["packed"] == want["packed"], (r["idx"], want["idx"])
    assert r["best_score"] == want["best_score"], (r["best_score"], want["best_score"])
    assert r["path"] == want["path"] and r["tile"] == want["tile"]
    assert r["seeded"] == 0


@pytest.mark.parametrize("order,score,path,tile", [
    (3, episdet.K2, episdet.PATH_AUTO, 0), (3, episdet.MI, M, 5), (3, episdet.K2, V, 0),
    (2, episdet.MI, episdet.PATH_AUTO, 0), (2, episdet.K2, M, 0), (2, episdet.K2, V, 1)])
def test_noalloc_matches_run_two_calls(order, score, path, tile):
    big = episdet.generate(300 if order == 3 else 1200, 3000, 2600, seed=61, planted_sets=1, planted_order=order)
    small = episdet.generate(90 if order == 3 else 400, 1100, 1500, seed=62, planted_sets=1, planted_order=order)
    want = [expected(s, order, score, path, tile) for s in (big, small)]
    with episdet.Dataset.from_synthetic(big) as d1, episdet.Dataset.from_synthetic(small) as d2:
        for rep in range(2):     # first: tables built in the scratch array; second: the same again
            for ds, w, name in ((d1, want[0], "larger"), (d2, want[1], "smaller")):
                say(f"noalloc order={order} score={score} path={path} tile={tile} {name} call {rep}")
                same(ds.run_noalloc(order, score, path=path, tile=tile), w)
        # with the datasets' own tables built by episdet_run, the call uses them
        for ds, w in ((d1, want[0]), (d2, want[1])):
            r = ds.run(order, score, path=path, tile=tile)
            assert r["packed"] == w["packed"] and r["best_score"] == w["best_score"]
            same(ds.run_noalloc(order, score, path=path, tile=tile), w)


def test_noalloc_overflow_windows():
    """All 2,763,520 triplets of 256 monomorphic SNPs tie: more than the
    scratch array's 2^18 candidate slots, so the call reruns in windows and
    still returns (0, 1, 2) with episdet_run's score; then a smaller tied
    dataset."""
    for n in (256, 120):
        cases = np.zeros((n, 2, episdet.padded_words(1000)), dtype=np.uint32)
        controls = np.zeros((n, 2, episdet.padded_words(1200)), dtype=np.uint32)
        cases[:, 0, :31] = 0xFFFFFFFF
        cases[:, 0, 31] = 0xFF000000          # 1000 = 31 * 32 + 8 samples
        controls[:, 0, :37] = 0xFFFFFFFF
        controls[:, 0, 37] = 0xFFFF0000       # 1200 = 37 * 32 + 16 samples
        with episdet.Dataset(cases, controls, 1000, 1200) as ref:
            want = ref.run(3, episdet.K2)
        with episdet.Dataset(cases, controls, 1000, 1200) as ds:
            say(f"noalloc all tied n={n}")
            r = ds.run_noalloc(3, episdet.K2)
            assert r["packed"] == want["packed"] == (1 << 21) | (2 << 42)     # (0, 1, 2)
            assert r["best_score"] == want["best_score"]
            if n == 256:
                assert r["fallback"] == 1


def test_noalloc_limits_and_refusals():
    """A 3-way dataset whose pair table does not fit the scratch array is
    refused before any device work (EPISDET_ERR_SCRATCH); once episdet_run
    has built the dataset's own tables it runs, with episdet_run's answer.
    PLAIN, rebuild and autotune are refused."""
    syn = episdet.generate(1500, 1000, 1000, seed=63, planted_sets=1, planted_order=3)
    with episdet.Dataset.from_synthetic(syn) as ds:
        say("noalloc too large for the scratch array")
        with pytest.raises(episdet.EpisdetError) as ei:
            ds.run_noalloc(3, episdet.K2)
        assert ei.value.code == SCRATCH_ERR
        want = ds.run(3, episdet.K2)
        say("noalloc with the dataset's own tables")
        same(ds.run_noalloc(3, episdet.K2), want)
        for kw in ({"path": episdet.PATH_PLAIN},):
            with pytest.raises(episdet.EpisdetError) as ei:
                ds.run_noalloc(3, episdet.K2, **kw)
            assert ei.value.code == 1
    lib = episdet.search_lib()
    for field in ("rebuild", "autotune"):
        o = episdet.Options(order=2, score=episdet.K2)
        setattr(o, field, 1)
        with episdet.Dataset.from_synthetic(episdet.generate(64, 500, 500, seed=64)) as ds:
            assert lib.episdet_run_noalloc(ds._h, episdet.C.byref(o), episdet.C.byref(episdet.Result())) == 1
