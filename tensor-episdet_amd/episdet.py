"""ctypes binding of the gfx950 epistasis search (include/episdet.h) and of
the host data library (include/episdet_data.h).

This is the Python-side mirror of the reference's operator boundary,
``EpistasisDetectionSearch`` (include/search-pairs.hpp:4,
include/search-triplets.hpp:4): ``search()`` takes the same twelve arguments'
worth of information and returns the same outputs.  ``Dataset`` exposes the
extended interface (device-resident data, path choice, table dumps).

The native libraries are loaded from ``lib/`` next to this file; nothing
falls back to Python or CPU code -- a missing library raises ``NativeMissing``.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")
BIN_DIR = os.path.join(_HERE, "bin")

K2, MI = 0, 1
PATH_AUTO, PATH_FUSED_VALU, PATH_FUSED_MFMA, PATH_PLAIN = 0, 1, 2, 3
TOPK_MAX = 1024   # EPISDET_TOPK_MAX
MAF_UNIFORM, MAF_BETA = 0, 1

STATUS = {0: "ok", 1: "invalid argument", 2: "no gfx950 device", 3: "HIP runtime error",
          4: "allocation failed", 5: "candidate buffer overflow", 6: "internal error",
          7: "a sample is marked in both genotype planes of a SNP",
          8: "the search needs more than the no-allocation entry point's scratch memory"}


class NativeMissing(RuntimeError):
    pass


class EpisdetError(RuntimeError):
    def __init__(self, code, what=""):
        super().__init__(f"{what}: {STATUS.get(code, code)} ({code})")
        self.code = code


def _load(name):
    path = os.path.join(LIB_DIR, name)
    if not os.path.exists(path):
        raise NativeMissing(f"{path} is not built; run __graft_entry__.build()")
    return C.CDLL(path)


class Options(C.Structure):
    _fields_ = [("order", C.c_int), ("score", C.c_int), ("path", C.c_int), ("tile", C.c_int),
                ("plain_rounds", C.c_int), ("candidate_cap", C.c_int), ("rebuild", C.c_int),
                ("topk", C.c_int), ("autotune", C.c_int), ("reserved", C.c_int * 5)]


class TuneEntry(C.Structure):
    _fields_ = [("path", C.c_int), ("tile", C.c_int), ("est_ms", C.c_double), ("sample_ms", C.c_double),
                ("sample_sets", C.c_longlong)]


class Result(C.Structure):
    _fields_ = [("best_score", C.c_double), ("best_score_f32", C.c_float), ("pad0", C.c_int),
                ("packed", C.c_ulonglong), ("idx", C.c_int * 3), ("path", C.c_int),
                ("sets", C.c_longlong), ("rounds", C.c_longlong), ("rounds_run", C.c_longlong),
                ("candidates", C.c_longlong), ("kernel_ms", C.c_double), ("rescore_ms", C.c_double), ("prep_ms", C.c_double),
                ("wall_s", C.c_double), ("fp32_max_abs_err", C.c_double), ("fp32_max_rel_err", C.c_double),
                ("margin", C.c_double), ("fallback", C.c_int), ("tile", C.c_int), ("seeded", C.c_int),
                ("reserved", C.c_int * 5)]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_ if f not in ("reserved", "pad0")}
        d["idx"] = tuple(self.idx)
        return d


class GenParams(C.Structure):
    _fields_ = [("num_snps", C.c_int), ("num_cases", C.c_int), ("num_controls", C.c_int),
                ("cases_words", C.c_int), ("controls_words", C.c_int), ("seed", C.c_ulonglong),
                ("maf_dist", C.c_int), ("maf_lo", C.c_double), ("maf_hi", C.c_double),
                ("beta_a", C.c_double), ("beta_b", C.c_double), ("planted_sets", C.c_int),
                ("planted_order", C.c_int), ("baseline_risk", C.c_double), ("peak_risk", C.c_double),
                ("planted_maf_lo", C.c_double), ("planted_maf_hi", C.c_double), ("threads", C.c_int)]


class Job(C.Structure):
    _fields_ = [("order", C.c_int), ("score", C.c_int), ("cases", C.c_void_p), ("controls", C.c_void_p),
                ("numSNPs", C.c_int), ("numCases", C.c_int), ("numControls", C.c_int),
                ("casesWords", C.c_int), ("controlsWords", C.c_int)]


_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")

# Every symbol include/episdet.h and include/episdet_data.h declare.
SEARCH_SYMBOLS = ("episdet_search", "episdet_dataset_create", "episdet_dataset_destroy", "episdet_run",
                  "episdet_run_noalloc", "episdet_run_topk", "episdet_autotune", "episdet_run_batch", "episdet_search_batch", "episdet_tables", "episdet_dump_tables",
                  "episdet_checksum", "episdet_n_choose_k", "episdet_reference_rounds", "episdet_status_string",
                  "episdet_device_info", "episdet_device_synchronize")
DATA_SYMBOLS = ("episdet_gen_defaults", "episdet_generate", "episdet_padded_words", "episdet_write_dataset",
                "episdet_read_header", "episdet_read_planes")

_data = None
_search = None


def data_lib():
    global _data
    if _data is None:
        lib = _load("libepisdet_data.so")
        lib.episdet_gen_defaults.argtypes = [C.POINTER(GenParams)]
        lib.episdet_generate.argtypes = [C.POINTER(GenParams), _u32p, _u32p, _i32p, _f64p]
        lib.episdet_generate.restype = C.c_int
        lib.episdet_padded_words.argtypes = [C.c_int]
        lib.episdet_padded_words.restype = C.c_int
        lib.episdet_write_dataset.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_int, _u32p, _u32p]
        lib.episdet_write_dataset.restype = C.c_int
        lib.episdet_read_header.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                            C.POINTER(C.c_int), C.c_char_p, C.c_char_p, C.c_int]
        lib.episdet_read_header.restype = C.c_int
        lib.episdet_read_planes.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, _u32p]
        lib.episdet_read_planes.restype = C.c_int
        _data = lib
    return _data


def search_lib():
    global _search
    if _search is None:
        lib = _load("libepisdet.so")
        lib.episdet_search.argtypes = [C.c_int, C.c_int, _u32p, _u32p, C.c_int, C.c_int, C.c_int, C.c_uint,
                                       C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                       C.POINTER(C.c_float), C.POINTER(C.c_ulonglong)]
        lib.episdet_search.restype = C.c_int
        lib.episdet_dataset_create.argtypes = [_u32p, _u32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                               C.POINTER(C.c_void_p)]
        lib.episdet_dataset_create.restype = C.c_int
        lib.episdet_dataset_destroy.argtypes = [C.c_void_p]
        lib.episdet_dataset_destroy.restype = None
        lib.episdet_run.argtypes = [C.c_void_p, C.POINTER(Options), C.POINTER(Result)]
        lib.episdet_run.restype = C.c_int
        lib.episdet_run_noalloc.argtypes = [C.c_void_p, C.POINTER(Options), C.POINTER(Result)]
        lib.episdet_run_noalloc.restype = C.c_int
        lib.episdet_run_topk.argtypes = [C.c_void_p, C.POINTER(Options), C.c_int, C.POINTER(C.c_ulonglong),
                                         C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(Result)]
        lib.episdet_run_topk.restype = C.c_int
        lib.episdet_autotune.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(TuneEntry), C.c_int,
                                         C.POINTER(C.c_int), C.POINTER(C.c_int)]
        lib.episdet_autotune.restype = C.c_int
        lib.episdet_run_batch.argtypes = [C.POINTER(C.c_void_p), C.POINTER(Options), C.c_int, C.POINTER(Result),
                                          C.POINTER(C.c_double)]
        lib.episdet_run_batch.restype = C.c_int
        lib.episdet_search_batch.argtypes = [C.POINTER(Job), C.c_int, C.POINTER(Options), C.POINTER(Result),
                                             C.POINTER(C.c_double)]
        lib.episdet_search_batch.restype = C.c_int
        lib.episdet_tables.argtypes = [C.c_void_p, C.c_int, _i32p, C.c_int, _u32p]
        lib.episdet_tables.restype = C.c_int
        lib.episdet_dump_tables.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, _u32p, C.c_size_t]
        lib.episdet_dump_tables.restype = C.c_int
        lib.episdet_checksum.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_ulonglong)]
        lib.episdet_checksum.restype = C.c_int
        lib.episdet_n_choose_k.argtypes = [C.c_uint, C.c_uint]
        lib.episdet_n_choose_k.restype = C.c_ulonglong
        lib.episdet_reference_rounds.argtypes = [C.c_int, C.c_int]
        lib.episdet_reference_rounds.restype = C.c_longlong
        lib.episdet_status_string.argtypes = [C.c_int]
        lib.episdet_status_string.restype = C.c_char_p
        lib.episdet_device_info.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        lib.episdet_device_info.restype = C.c_int
        lib.episdet_device_synchronize.argtypes = []
        lib.episdet_device_synchronize.restype = C.c_int
        _search = lib
    return _search


def padded_words(num_samples: int) -> int:
    return int(data_lib().episdet_padded_words(int(num_samples)))


@dataclass
class Synthetic:
    cases: np.ndarray        # [n, 2, Wc] uint32
    controls: np.ndarray     # [n, 2, Wt] uint32
    num_cases: int
    num_controls: int
    planted: list            # list of tuples of planted SNP indices
    penetrance: np.ndarray   # [sets, 27]

    @property
    def num_snps(self):
        return self.cases.shape[0]


def generate(num_snps, num_cases, num_controls, seed=1, maf=(0.05, 0.5), maf_dist=MAF_UNIFORM,
             beta=(0.5, 2.0), planted_sets=0, planted_order=3, baseline=0.1, peak=0.4,
             planted_maf=(0.25, 0.5), cases_words=0, controls_words=0, threads=0) -> Synthetic:
    """Seeded HWE dataset in the reference's plane layout (see episdet_data.h)."""
    lib = data_lib()
    gp = GenParams()
    lib.episdet_gen_defaults(C.byref(gp))
    gp.num_snps, gp.num_cases, gp.num_controls = num_snps, num_cases, num_controls
    gp.cases_words, gp.controls_words = cases_words, controls_words
    gp.seed = seed
    gp.maf_dist = maf_dist
    gp.maf_lo, gp.maf_hi = maf
    gp.beta_a, gp.beta_b = beta
    gp.planted_sets, gp.planted_order = planted_sets, planted_order
    gp.baseline_risk, gp.peak_risk = baseline, peak
    gp.planted_maf_lo, gp.planted_maf_hi = planted_maf
    gp.threads = threads
    wc = cases_words or padded_words(num_cases)
    wt = controls_words or padded_words(num_controls)
    cases = np.zeros((num_snps, 2, wc), dtype=np.uint32)
    controls = np.zeros((num_snps, 2, wt), dtype=np.uint32)
    planted = np.full(max(planted_sets, 1) * 3, -1, dtype=np.int32)
    pen = np.zeros(max(planted_sets, 1) * 27, dtype=np.float64)
    rc = lib.episdet_generate(C.byref(gp), cases, controls, planted, pen)
    if rc != 0:
        raise EpisdetError(1, "episdet_generate")
    sets = [tuple(int(v) for v in planted[3 * s:3 * s + planted_order]) for s in range(planted_sets)]
    return Synthetic(cases, controls, num_cases, num_controls, sets, pen.reshape(-1, 27)[:planted_sets])


def write_dataset(directory, name, syn: Synthetic):
    rc = data_lib().episdet_write_dataset(str(directory).encode(), name.encode(), syn.num_snps, syn.num_cases,
                                          syn.num_controls, syn.cases, syn.controls)
    if rc:
        raise EpisdetError(1, "episdet_write_dataset")
    return os.path.join(str(directory), name + ".txt")


def read_dataset(txt_path):
    lib = data_lib()
    n, nc, nt = C.c_int(), C.c_int(), C.c_int()
    pc, pt = C.create_string_buffer(4096), C.create_string_buffer(4096)
    if lib.episdet_read_header(str(txt_path).encode(), C.byref(n), C.byref(nc), C.byref(nt), pc, pt, 4096):
        raise EpisdetError(1, "episdet_read_header")
    cases = np.zeros((n.value, 2, padded_words(nc.value)), dtype=np.uint32)
    controls = np.zeros((n.value, 2, padded_words(nt.value)), dtype=np.uint32)
    if lib.episdet_read_planes(pc.value, n.value, nc.value, n.value, cases):
        raise EpisdetError(1, "episdet_read_planes(cases)")
    if lib.episdet_read_planes(pt.value, n.value, nt.value, n.value, controls):
        raise EpisdetError(1, "episdet_read_planes(controls)")
    return Synthetic(cases, controls, nc.value, nt.value, [], np.zeros((0, 27)))


def n_choose_k(n, k):
    return int(search_lib().episdet_n_choose_k(n, k))


def reference_rounds(order, n):
    return int(search_lib().episdet_reference_rounds(order, n))


def device_info():
    name = C.create_string_buffer(256)
    cu, clk = C.c_int(), C.c_int()
    rc = search_lib().episdet_device_info(name, 256, C.byref(cu), C.byref(clk))
    if rc:
        raise EpisdetError(rc, "episdet_device_info")
    return name.value.decode(), cu.value, clk.value


def synchronize():
    """Waits for all work on the current device."""
    rc = search_lib().episdet_device_synchronize()
    if rc:
        raise EpisdetError(rc, "episdet_device_synchronize")


def search(order, score, cases, controls, num_cases, num_controls):
    """Reference-compatible call: returns (packed, best_score_f32, rounds, seconds)."""
    lib = search_lib()
    n = cases.shape[0]
    nb = -(-n // 256)
    cases = np.ascontiguousarray(cases, dtype=np.uint32)
    controls = np.ascontiguousarray(controls, dtype=np.uint32)
    rounds, secs, best, packed = C.c_int(), C.c_double(), C.c_float(), C.c_ulonglong()
    rc = lib.episdet_search(order, score, cases, controls, n, num_cases, num_controls, nb * 256,
                            cases.shape[2] * 32, controls.shape[2] * 32, C.byref(rounds), C.byref(secs),
                            C.byref(best), C.byref(packed))
    if rc:
        raise EpisdetError(rc, "episdet_search")
    return packed.value, best.value, rounds.value, secs.value


def run_batch(datasets, orders, scores, rebuild=False, path=PATH_AUTO, autotune=False, paths=None, tiles=None):
    """Many resident searches in one graph launch; returns (results, batch_ms).
    paths / tiles: optional per-job path and tile (default: `path`, tile 0)."""
    n = len(datasets)
    hs = (C.c_void_p * n)(*[d._h for d in datasets])
    opts = (Options * n)()
    for j in range(n):
        opts[j].order, opts[j].score = orders[j], scores[j]
        opts[j].path = paths[j] if paths is not None else path
        opts[j].tile = tiles[j] if tiles is not None else 0
        opts[j].rebuild = 1 if rebuild else 0
        opts[j].autotune = 1 if autotune else 0
    res = (Result * n)()
    ms = C.c_double()
    rc = search_lib().episdet_run_batch(hs, opts, n, res, C.byref(ms))
    if rc:
        raise EpisdetError(rc, "episdet_run_batch")
    return [r.as_dict() for r in res], ms.value


def search_batch(jobs):
    """jobs: list of (order, score, Synthetic).  Host buffers, one call;
    returns (results, batch_ms)."""
    n = len(jobs)
    arr = (Job * n)()
    keep = []
    for j, (order, score, syn) in enumerate(jobs):
        cs = np.ascontiguousarray(syn.cases, dtype=np.uint32)
        ct = np.ascontiguousarray(syn.controls, dtype=np.uint32)
        keep += [cs, ct]
        arr[j] = Job(order, score, cs.ctypes.data, ct.ctypes.data, cs.shape[0], syn.num_cases, syn.num_controls,
                     cs.shape[2], ct.shape[2])
    res = (Result * n)()
    ms = C.c_double()
    rc = search_lib().episdet_search_batch(arr, n, None, res, C.byref(ms))
    if rc:
        raise EpisdetError(rc, "episdet_search_batch")
    return [r.as_dict() for r in res], ms.value


def unpack(order, packed):
    if order == 3:
        return (packed & 0x1FFFFF, (packed >> 21) & 0x1FFFFF, (packed >> 42) & 0x1FFFFF)
    return (packed & 0xFFFFFFFF, packed >> 32)


class Dataset:
    """Device-resident dataset (episdet_dataset_create ... _destroy)."""

    def __init__(self, cases, controls, num_cases, num_controls):
        self._lib = search_lib()
        self.cases = np.ascontiguousarray(cases, dtype=np.uint32)
        self.controls = np.ascontiguousarray(controls, dtype=np.uint32)
        self.n = self.cases.shape[0]
        self.num_cases, self.num_controls = num_cases, num_controls
        h = C.c_void_p()
        rc = self._lib.episdet_dataset_create(self.cases, self.controls, self.n, num_cases, num_controls,
                                              self.cases.shape[2], self.controls.shape[2], C.byref(h))
        if rc:
            raise EpisdetError(rc, "episdet_dataset_create")
        self._h = h

    @classmethod
    def from_synthetic(cls, syn: Synthetic):
        return cls(syn.cases, syn.controls, syn.num_cases, syn.num_controls)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.episdet_dataset_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, order, score=K2, path=PATH_AUTO, tile=0, plain_rounds=0, candidate_cap=0, rebuild=False,
            autotune=False) -> dict:
        o = Options(order=order, score=score, path=path, tile=tile, plain_rounds=plain_rounds,
                    candidate_cap=candidate_cap, rebuild=1 if rebuild else 0, autotune=1 if autotune else 0)
        r = Result()
        rc = self._lib.episdet_run(self._h, C.byref(o), C.byref(r))
        if rc:
            raise EpisdetError(rc, "episdet_run")
        return r.as_dict()

    def run_noalloc(self, order, score=K2, path=PATH_AUTO, tile=0, candidate_cap=0, topk=0) -> dict:
        """episdet_run_noalloc: the same search with no device allocation
        (per-search memory in the library's static scratch array)."""
        o = Options(order=order, score=score, path=path, tile=tile, candidate_cap=candidate_cap, topk=topk)
        r = Result()
        rc = self._lib.episdet_run_noalloc(self._h, C.byref(o), C.byref(r))
        if rc:
            raise EpisdetError(rc, "episdet_run_noalloc")
        return r.as_dict()

    def autotune(self, order, score=K2):
        """Times every kernel configuration for this shape and caches the
        fastest for AUTO: (list of dicts, index of the chosen one)."""
        cap = 32                                # more than the candidate kernels of any shape
        arr = (TuneEntry * cap)()
        cnt, chosen = C.c_int(), C.c_int()
        rc = self._lib.episdet_autotune(self._h, order, score, arr, cap, C.byref(cnt), C.byref(chosen))
        if rc:
            raise EpisdetError(rc, "episdet_autotune")
        out = [{"path": e.path, "tile": e.tile, "est_ms": e.est_ms, "sample_ms": e.sample_ms,
                "sample_sets": e.sample_sets} for e in arr[:min(cnt.value, cap)]]
        return out, chosen.value

    def topk(self, order, k, score=K2, path=PATH_AUTO, tile=0, rebuild=False, candidate_cap=0):
        """The k best sets by FP64 score (best first): (packed list, scores
        list, result dict)."""
        o = Options(order=order, score=score, path=path, tile=tile, rebuild=1 if rebuild else 0,
                    candidate_cap=candidate_cap)
        packed = (C.c_ulonglong * k)()
        scores = (C.c_double * k)()
        cnt = C.c_int()
        r = Result()
        rc = self._lib.episdet_run_topk(self._h, C.byref(o), k, packed, scores, C.byref(cnt), C.byref(r))
        if rc:
            raise EpisdetError(rc, "episdet_run_topk")
        m = cnt.value
        return [int(x) for x in packed[:m]], [float(x) for x in scores[:m]], r.as_dict()

    def tables(self, order, sets) -> np.ndarray:
        sets = np.ascontiguousarray(np.asarray(sets, dtype=np.int32).reshape(-1, order))
        nc = 27 if order == 3 else 9
        out = np.zeros((sets.shape[0], 2, nc), dtype=np.uint32)
        rc = self._lib.episdet_tables(self._h, order, sets, sets.shape[0], out)
        if rc:
            raise EpisdetError(rc, "episdet_tables")
        return out

    def dump_tables(self, order, path=PATH_FUSED_VALU, tile=0) -> np.ndarray:
        total = n_choose_k(self.n, order)
        nc = 27 if order == 3 else 9
        out = np.zeros((total, 2, nc), dtype=np.uint32)
        rc = self._lib.episdet_dump_tables(self._h, order, path, tile, out, out.size)
        if rc:
            raise EpisdetError(rc, "episdet_dump_tables")
        return out

    def checksum(self, order, path=PATH_FUSED_VALU, tile=0) -> int:
        v = C.c_ulonglong()
        rc = self._lib.episdet_checksum(self._h, order, path, tile, C.byref(v))
        if rc:
            raise EpisdetError(rc, "episdet_checksum")
        return v.value
