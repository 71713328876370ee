"""The C-ABI libraries load on a host without a GPU and export every symbol
include/*.h declares; host-only helpers agree with the reference's formulas.
No device compute is called here."""
import ctypes
import os
import re
import subprocess

import pytest

import episdet

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*[A-Za-z_][\w\s\*]*?\b(episdet_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def exported(lib_path):
    out = subprocess.run(["nm", "-D", "--defined-only", lib_path], capture_output=True, text=True, check=True)
    return {line.split()[-1] for line in out.stdout.splitlines() if line.strip()}


def test_search_library_exports_header():
    names = declared("episdet.h")
    assert len(names) >= 10
    syms = exported(os.path.join(episdet.LIB_DIR, "libepisdet.so"))
    missing = [n for n in names if n not in syms]
    assert not missing, missing
    assert sorted(episdet.SEARCH_SYMBOLS) == names


def test_data_library_exports_header():
    names = declared("episdet_data.h")
    syms = exported(os.path.join(episdet.LIB_DIR, "libepisdet_data.so"))
    missing = [n for n in names if n not in syms]
    assert not missing, missing
    assert sorted(episdet.DATA_SYMBOLS) == names


def test_libraries_load_without_device():
    lib = episdet.search_lib()
    for n in episdet.SEARCH_SYMBOLS:
        assert hasattr(lib, n)
    dlib = episdet.data_lib()
    for n in episdet.DATA_SYMBOLS:
        assert hasattr(dlib, n)


def test_host_helpers():
    assert episdet.n_choose_k(4096, 3) == 11_444_858_880
    # rounds of the reference's loops at BLOCK_SIZE = 256
    # (search-pairs.cu:137-141, search-triplets.cu:162-200)
    assert episdet.reference_rounds(3, 4096) == 208_896
    assert episdet.reference_rounds(3, 8192) == 1_531_904
    assert episdet.reference_rounds(2, 512) == 3
    assert episdet.reference_rounds(2, 16384) == 2080
    lib = episdet.search_lib()
    assert lib.episdet_status_string(0) == b"ok"
    assert lib.episdet_status_string(2) == b"no gfx950 device"
    assert lib.episdet_status_string(7) == episdet.STATUS[7].encode()


def test_cli_binaries_built():
    for order in ("pairs", "triplets"):
        for score in ("k2", "mi"):
            p = os.path.join(episdet.BIN_DIR, f"tensor-episdet.{order}.{score}.bin")
            assert os.access(p, os.X_OK), p


def test_invalid_arguments_rejected_before_device_use():
    import numpy as np
    lib = episdet.search_lib()
    h = ctypes.c_void_p()
    z = np.zeros((4, 2, 32), dtype=np.uint32)
    # numSNPs < 2 and words too small for the sample count are argument errors
    assert lib.episdet_dataset_create(z, z, 1, 10, 10, 32, 32, ctypes.byref(h)) == 1
    assert lib.episdet_dataset_create(z, z, 4, 2000, 10, 32, 32, ctypes.byref(h)) == 1
    # more than 2^30 samples in all (checked before the plane sizes)
    assert lib.episdet_dataset_create(z, z, 4, 1 << 29, 1 << 29, 1 << 24, 1 << 24, ctypes.byref(h)) == 1
    # the extended entry points check their arguments before touching a device
    res = episdet.Result()
    opt = episdet.Options(order=3, score=0)
    packed = (ctypes.c_ulonglong * 4)()
    scores = (ctypes.c_double * 4)()
    cnt = ctypes.c_int()
    assert lib.episdet_run(None, ctypes.byref(opt), ctypes.byref(res)) == 1
    assert lib.episdet_run_topk(None, ctypes.byref(opt), 4, packed, scores, ctypes.byref(cnt), ctypes.byref(res)) == 1
    assert lib.episdet_autotune(None, 3, 0, None, 0, ctypes.byref(cnt), ctypes.byref(cnt)) == 1
    assert lib.episdet_run_batch(None, ctypes.byref(opt), 0, ctypes.byref(res), None) == 1
    assert lib.episdet_search_batch(None, 0, None, ctypes.byref(res), None) == 1
    assert lib.episdet_dump_tables(None, 3, episdet.PATH_FUSED_MFMA, 0, np.zeros(4, dtype=np.uint32), 4) == 1
    assert lib.episdet_checksum(None, 3, episdet.PATH_FUSED_MFMA, 0, None) == 1
