"""The XCD-blocked work-unit orders of the MFMA pair kernels (no GPU needed).

kernels_valu.hpp and kernels_mfma_pairs.hpp hand the units of the 2-way
searches to the XCDs in blocks, through blocked_unrank (the 64 x 64 kernel)
and p2_blocked_unrank (the 128 x 64 kernel, which also builds the 3-way pair
table).  A unit skipped or taken twice would drop or double count pairs; this
test compiles tests/units/unit_orders.hip against those headers, runs it on
the host and
checks that each order is a bijection onto the units of the enumeration it
replaces, for 1-400 tiles and 2-20000 SNPs.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_blocked_unit_orders_cover_every_unit_once(tmp_path):
    exe = tmp_path / "unit_orders"
    subprocess.run([HIPCC, "-O2", "-std=c++17", "--offload-arch=gfx950", "-x", "hip",
                    "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "tensor-episdet_amd", "csrc"),
                    os.path.join(ROOT, "tests", "units", "unit_orders.hip"), "-o", str(exe)],
                   check=True, timeout=600)
    p = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    assert p.returncode == 0, p.stdout[-2000:]
    assert p.stdout.startswith("ok ")
