"""ctypes wrapper of the CPU oracle (episdet_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by the CPU-baseline leg of bench.py, as the checker.  The product never
imports it.
"""
from __future__ import annotations

import ctypes as C
import itertools
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "lib", "libepisdet_oracle.so")

K2, MI = 0, 1

_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class OrcResult(C.Structure):
    _fields_ = [("score", C.c_double), ("packed", C.c_uint64), ("sets", C.c_int64), ("idx", C.c_int32 * 3)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} not built (make -C oracle)")
        L = C.CDLL(LIB)
        L.orc_n_choose_k.argtypes = [C.c_uint, C.c_uint]
        L.orc_n_choose_k.restype = C.c_uint64
        L.orc_decode.argtypes = [_u32p, C.c_int, C.c_int, C.c_int, _u8p]
        L.orc_check_planes.argtypes = [_u32p, C.c_int, C.c_int, C.c_int]
        L.orc_check_planes.restype = C.c_int64
        L.orc_singles.argtypes = [_u32p, C.c_int, C.c_int, C.c_int, _u32p]
        L.orc_table_direct.argtypes = [_u8p, C.c_int, C.c_int, _i32p, _u32p]
        L.orc_table_bits.argtypes = [_u32p, C.c_int, C.c_int, C.c_int, _i32p, _u32p]
        L.orc_lnfact_table.argtypes = [C.c_int, _f64p]
        L.orc_log2_table.argtypes = [C.c_int, _f64p]
        L.orc_score_k2.argtypes = [_u32p, _u32p, C.c_int, _f64p]
        L.orc_score_k2.restype = C.c_double
        L.orc_score_mi.argtypes = [_u32p, _u32p, C.c_int, C.c_int, C.c_int, _f64p]
        L.orc_score_mi.restype = C.c_double
        L.orc_search.argtypes = [_u32p, _u32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                 C.c_int, C.c_int, C.c_int, C.POINTER(OrcResult)]
        L.orc_search.restype = C.c_int
        L.orc_search_ex.argtypes = L.orc_search.argtypes + [C.POINTER(C.c_uint64)]
        L.orc_search_ex.restype = C.c_int
        L.orc_all_tables.argtypes = [_u32p, _u32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.c_void_p]
        L.orc_all_tables.restype = C.c_uint64
        L.orc_table_hash.argtypes = [C.c_uint64, _u32p, _u32p, C.c_int]
        L.orc_table_hash.restype = C.c_uint64
        L.orc_uses_avx512.restype = C.c_int
        _lib = L
    return _lib


def n_choose_k(n, k):
    return int(lib().orc_n_choose_k(n, k))


def decode(planes, N):
    planes = np.ascontiguousarray(planes, dtype=np.uint32)
    n, _, W = planes.shape
    g = np.zeros((n, N), dtype=np.uint8)
    lib().orc_decode(planes, n, N, W, g)
    return g


def check_planes(planes, N):
    planes = np.ascontiguousarray(planes, dtype=np.uint32)
    n, _, W = planes.shape
    return int(lib().orc_check_planes(planes, n, N, W))


def singles(planes, N):
    planes = np.ascontiguousarray(planes, dtype=np.uint32)
    n, _, W = planes.shape
    out = np.zeros((n, 3), dtype=np.uint32)
    lib().orc_singles(planes, n, N, W, out)
    return out


def table_direct(geno, N, snps):
    snps = np.ascontiguousarray(snps, dtype=np.int32)
    out = np.zeros(27 if len(snps) == 3 else 9, dtype=np.uint32)
    lib().orc_table_direct(np.ascontiguousarray(geno, dtype=np.uint8), N, len(snps), snps, out)
    return out


def table_bits(planes, N, snps):
    planes = np.ascontiguousarray(planes, dtype=np.uint32)
    snps = np.ascontiguousarray(snps, dtype=np.int32)
    out = np.zeros(27 if len(snps) == 3 else 9, dtype=np.uint32)
    lib().orc_table_bits(planes, N, planes.shape[2], len(snps), snps, out)
    return out


def lnfact_table(mmax):
    t = np.zeros(mmax + 1, dtype=np.float64)
    lib().orc_lnfact_table(mmax, t)
    return t


def log2_table(N):
    t = np.zeros(N + 1, dtype=np.float64)
    lib().orc_log2_table(N, t)
    return t


def score_k2(cas, ctl, table=None):
    cas = np.ascontiguousarray(cas, dtype=np.uint32)
    ctl = np.ascontiguousarray(ctl, dtype=np.uint32)
    if table is None:
        table = lnfact_table(int(cas.sum() + ctl.sum()) + 1)
    return float(lib().orc_score_k2(cas, ctl, len(cas), table))


def score_mi(cas, ctl, table=None):
    cas = np.ascontiguousarray(cas, dtype=np.uint32)
    ctl = np.ascontiguousarray(ctl, dtype=np.uint32)
    Nc, Nt = int(cas.sum()), int(ctl.sum())
    if table is None:
        table = log2_table(Nc + Nt)
    return float(lib().orc_score_mi(cas, ctl, len(cas), Nc, Nt, table))


def search(cases, controls, Nc, Nt, order, score, threads=0, last=None, checksum=False):
    """Exhaustive search; returns dict(score, packed, sets, idx), plus
    "checksum" (the orc_table_hash sum of every evaluated set's table, as
    all_tables computes it) when checksum=True.  `last` restricts the largest
    index of each set to [lo, hi)."""
    cases = np.ascontiguousarray(cases, dtype=np.uint32)
    controls = np.ascontiguousarray(controls, dtype=np.uint32)
    n = cases.shape[0]
    lo, hi = last if last is not None else (0, n)
    r = OrcResult()
    cs = C.c_uint64(0)
    rc = lib().orc_search_ex(cases, controls, n, Nc, Nt, cases.shape[2], controls.shape[2], order, score, threads,
                             lo, hi, C.byref(r), C.byref(cs) if checksum else None)
    if rc:
        raise RuntimeError("orc_search failed")
    out = {"score": r.score, "packed": int(r.packed), "sets": int(r.sets), "idx": tuple(r.idx)[:order]}
    if checksum:
        out["checksum"] = int(cs.value)
    return out


def all_tables(cases, controls, Nc, Nt, order, want_tables=True):
    """(tables [C(n,k), 2, 3^k] or None, checksum)."""
    cases = np.ascontiguousarray(cases, dtype=np.uint32)
    controls = np.ascontiguousarray(controls, dtype=np.uint32)
    n = cases.shape[0]
    nc = 27 if order == 3 else 9
    total = n_choose_k(n, order)
    out = np.zeros((total, 2, nc), dtype=np.uint32) if want_tables else None
    ptr = out.ctypes.data_as(C.c_void_p) if want_tables else None
    cs = lib().orc_all_tables(cases, controls, n, Nc, Nt, cases.shape[2], controls.shape[2], order, ptr)
    return out, int(cs)


def ranked(cases, controls, Nc, Nt, order, score):
    """Every set of a small dataset ranked best first by its FP64 score
    (K2 ascending, MI descending; ties to the lowest packed index, the tie
    rule of orc_search).  Returns (packed[], scores[]) as numpy arrays.  The
    reference keeps only the first entry (helper.cu:552-583)."""
    tabs, _ = all_tables(cases, controls, Nc, Nt, order)
    n = cases.shape[0]
    N = Nc + Nt
    if score == 0:
        t = lnfact_table(N + 1)
        sc = np.array([score_k2(c[0], c[1], t) for c in tabs])
    else:
        t = log2_table(N)
        sc = np.array([score_mi(c[0], c[1], t) for c in tabs])
    sets = list(itertools.combinations(range(n), order))   # colex rank order below
    sets.sort(key=lambda s: s[::-1])
    packed = np.array([s[0] | (s[1] << 21) | (s[2] << 42) if order == 3 else s[0] | (s[1] << 32) for s in sets],
                      dtype=np.uint64)
    key = sc if score == 0 else -sc
    o = np.lexsort((packed, key))
    return packed[o], sc[o]


def uses_avx512():
    return bool(lib().orc_uses_avx512())
