"""The score bounds with which the MFMA kernels skip FP32 scores that cannot
matter (bound_exceeds in device_common.hpp).  With pi = Nc / N,
S = sum_c (cas_c - pi n_c)^2 / n_c and K = max(-ln(1 - pi) / pi^2,
-ln(pi) / (1 - pi)^2), the largest KL(p || pi) / (p - pi)^2 on [0, 1]:

    K2 >= N H(pi) - K S          (nats; from C(n, k) >= e^{n H(k/n)} / (n + 1))
    MI <= K S / (N ln 2)         (bits; MI = sum_c (n_c / N) KL(p_c || pi))

K2 further adds each cell's Stirling term, ln((n+1)!/(a! b!)) >= n H(a/n) +
ln(n)/2 - c0 with c0 = ln(pi/2)/2 for n >= 1, so the kernels use

    K2 >= N H(pi) - K S + sum_c ln(max(n_c, 1))/2 - ncells c0.

Checked on the CPU against the oracle's exact FP64 K2 and MI over random,
skewed and extreme tables; the curvature constant is checked to be the
supremum it claims, and a halved constant is shown to break the bounds.
"""
import math

import numpy as np

import oracle as O


def curvature(pi):
    return max(-math.log1p(-pi) / pi ** 2, -math.log(pi) / (1 - pi) ** 2)


def chi_sum(cas, ctl):
    cas = np.asarray(cas, dtype=np.float64)
    ctl = np.asarray(ctl, dtype=np.float64)
    pi = cas.sum() / (cas.sum() + ctl.sum())
    n = cas + ctl
    m = n > 0
    return (((cas - pi * n) ** 2)[m] / n[m]).sum(), pi


def mi_upper(cas, ctl, k=None):
    s, pi = chi_sum(cas, ctl)
    if k is None:
        k = curvature(pi)
    N = float(np.sum(cas) + np.sum(ctl))
    return k * s / (N * math.log(2.0))


def bound(cas, ctl, k=None):
    cas = np.asarray(cas, dtype=np.float64)
    ctl = np.asarray(ctl, dtype=np.float64)
    Nc, Nt = cas.sum(), ctl.sum()
    N = Nc + Nt
    pi = Nc / N
    if k is None:
        k = curvature(pi)
    n = cas + ctl
    m = n > 0
    s = (((cas - pi * n) ** 2)[m] / n[m]).sum()
    nh = -N * (pi * math.log(pi) + (1 - pi) * math.log(1 - pi))
    return nh - k * s


C0 = 0.5 * math.log(math.pi / 2)


def stirling_bound(cas, ctl, k=None, c0=C0, w=0.5):
    n = np.asarray(cas, dtype=np.float64) + np.asarray(ctl, dtype=np.float64)
    return bound(cas, ctl, k) + w * np.log(np.maximum(n, 1.0)).sum() - len(n) * c0


def cell_gap(n):
    """ln((n+1)!/(a! b!)) - n H(a/n) - ln(n)/2 for a = 0..n (b = n - a), exactly
    as the oracle's K2 of a one-cell table."""
    from scipy.special import gammaln
    a = np.arange(0, n + 1, dtype=np.float64)
    b = n - a
    g = gammaln(n + 2) - gammaln(a + 1) - gammaln(b + 1)
    with np.errstate(divide="ignore", invalid="ignore"):
        nh = np.where(a > 0, -a * np.log(a / n), 0.0) + np.where(b > 0, -b * np.log(b / n), 0.0)
    return g - nh - 0.5 * math.log(n)


def random_table(rng, cells=27):
    Nc, Nt = (int(x) for x in rng.integers(1, 6000, 2))
    kind = rng.integers(0, 4)
    if kind == 0:                       # same cell distribution in both classes
        p = rng.dirichlet(np.full(cells, rng.uniform(0.05, 3.0)))
        q = p
    elif kind == 1:                     # independent distributions (strong signal)
        p = rng.dirichlet(np.full(cells, rng.uniform(0.05, 3.0)))
        q = rng.dirichlet(np.full(cells, rng.uniform(0.05, 3.0)))
    elif kind == 2:                     # a planted shift of one cell's mass
        p = rng.dirichlet(np.full(cells, 1.0))
        q = p.copy()
        i, j = rng.integers(0, cells, 2)
        q[i] += 0.3
        q[j] = max(q[j] - 0.3, 0.0)
        q /= q.sum()
    else:                               # few samples, many empty cells
        Nc, Nt = (int(x) for x in rng.integers(1, 12, 2))
        p = rng.dirichlet(np.full(cells, 0.2))
        q = rng.dirichlet(np.full(cells, 0.2))
    return rng.multinomial(Nc, p), rng.multinomial(Nt, q)


def test_curvature_is_the_supremum():
    for pi in (0.01, 0.1, 0.3, 0.5, 0.62, 0.9, 0.999):
        p = np.linspace(0.0, 1.0, 200001)
        p = p[np.abs(p - pi) > 1e-7]
        with np.errstate(divide="ignore", invalid="ignore"):
            kl = np.where(p > 0, p * np.log(p / pi), 0.0) + np.where(p < 1, (1 - p) * np.log((1 - p) / (1 - pi)), 0.0)
        assert (kl / (p - pi) ** 2).max() <= curvature(pi) * (1 + 1e-12)


def test_bound_below_exact_k2():
    rng = np.random.default_rng(7)
    for _ in range(2000):
        cells = 27 if rng.integers(0, 4) else 9
        cas, ctl = random_table(rng, cells)
        k2 = O.score_k2(cas, ctl)
        assert bound(cas, ctl) <= k2 + 1e-9 * max(1.0, abs(k2)), (cas, ctl)


def test_bound_is_tight_for_a_separating_table():
    # every cell pure cases or pure controls, 50/50: the bound is 0 while
    # K2 = sum ln(n_c + 1); a halved curvature constant breaks it
    cas = np.array([500] * 13 + [0] * 14, dtype=np.uint32)
    ctl = np.array([0] * 13 + [464] * 13 + [468], dtype=np.uint32)
    k2 = O.score_k2(cas, ctl)
    assert bound(cas, ctl) <= k2
    pi = cas.sum() / (cas.sum() + ctl.sum())
    assert bound(cas, ctl, k=curvature(pi) / 2) > k2


def test_mi_upper_bound():
    rng = np.random.default_rng(8)
    for _ in range(2000):
        cells = 27 if rng.integers(0, 4) else 9
        cas, ctl = random_table(rng, cells)
        mi = O.score_mi(cas, ctl)
        assert mi <= mi_upper(cas, ctl) + 1e-12, (cas, ctl)


def test_mi_bound_is_tight_for_a_separating_table():
    cas = np.array([500] * 13 + [0] * 14, dtype=np.uint32)
    ctl = np.array([0] * 13 + [464] * 13 + [468], dtype=np.uint32)
    mi = O.score_mi(cas, ctl)                # 1 bit: the table separates the classes
    assert abs(mi - 1.0) < 1e-12 and mi <= mi_upper(cas, ctl) + 1e-12
    assert mi_upper(cas, ctl, k=curvature(0.5) / 2) < mi


def test_cell_stirling_term_exhaustive():
    # every (a, b) with n = a + b <= 3000: the gap stays above -c0, and the
    # one-cell gap matches the oracle's K2 of that cell on a sample
    worst = min(cell_gap(n).min() for n in range(1, 3001))
    assert worst > -C0
    for a, b in ((0, 1), (1, 1), (3, 7), (250, 250), (1, 2999)):
        cas, ctl = np.array([a], dtype=np.uint32), np.array([b], dtype=np.uint32)
        n = a + b
        h = sum(-x * math.log(x / n) for x in (a, b) if x)
        assert abs(O.score_k2(cas, ctl) - h - 0.5 * math.log(n) - cell_gap(n)[a]) < 1e-9 * n


def test_cell_stirling_constant_is_the_infimum():
    # at a = b the gap is -c0 + about 0.75 / n: a constant 1e-3 smaller, or
    # a log weight 1% larger, fails for large balanced cells
    g = cell_gap(20000)[10000]
    assert -C0 < g < -C0 + 1e-4
    assert g < -(C0 - 1e-3)
    assert cell_gap(200000)[100000] - 0.005 * math.log(200000) < -C0


def test_stirling_bound_below_exact_k2():
    rng = np.random.default_rng(11)
    for _ in range(2000):
        cells = 27 if rng.integers(0, 4) else 9
        cas, ctl = random_table(rng, cells)
        k2 = O.score_k2(cas, ctl)
        assert stirling_bound(cas, ctl) <= k2 + 1e-9 * max(1.0, abs(k2)), (cas, ctl)


def test_stirling_bound_is_tight_for_balanced_null_tables():
    # 27 large cells with cases = controls: S = 0 and K2 exceeds the bound by
    # only sum_c 0.75 / n_c; a constant 1e-3 smaller breaks it
    cas = np.full(27, 5000, dtype=np.uint32)
    ctl = np.full(27, 5000, dtype=np.uint32)
    k2 = O.score_k2(cas, ctl)
    lb = stirling_bound(cas, ctl)
    assert lb <= k2 < lb + 27 * 1e-4
    assert stirling_bound(cas, ctl, c0=C0 - 1e-3) > k2


def test_stirling_term_lifts_the_bound_above_weak_winners():
    # null data: the plain bound of a typical set lies below the K2 of the
    # best of many random sets, so nothing could be skipped; the Stirling
    # term puts typical sets above it
    rng = np.random.default_rng(12)
    Nc = Nt = 8192
    tables = []
    for _ in range(400):
        p = rng.dirichlet(np.full(27, 2.0))
        tables.append((rng.multinomial(Nc, p), rng.multinomial(Nt, p)))
    k2 = np.array([O.score_k2(c, t) for c, t in tables])
    best = k2.min()
    plain = np.array([bound(c, t) for c, t in tables])
    lifted = np.array([stirling_bound(c, t) for c, t in tables])
    assert (plain < best).all()
    assert (lifted > best).mean() > 0.5
