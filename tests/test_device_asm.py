"""Checks on the gfx950 assembly of the search library (no GPU needed).

The LDS-DMA helper in kernels_mfma.hpp is inline assembly that sets m0 (the
LDS destination of `buffer_load_dwordx4 ... lds`) without naming m0 as
clobbered, because the compiler treats m0 as reserved.  That is only sound
while no compiler-generated instruction in the same kernels keeps a value in
m0; this test proves it on the code the Makefile builds: every access to m0 in
the device assembly must be one of those `s_mov_b32 m0, sN` lines directly
followed by its LDS-DMA load.
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "tensor-episdet_amd")


@pytest.mark.skipif(shutil.which("hipcc") is None and not os.path.exists("/opt/rocm/bin/hipcc"),
                    reason="hipcc not installed")
def test_m0_written_only_by_lds_dma(tmp_path):
    out = tmp_path / "search_gfx950.s"
    subprocess.run(["make", "-s", "-C", PKG, "device-asm", f"ASM_OUT={out}"], check=True,
                   stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, timeout=600)
    lines = out.read_text().split("\n")
    fn = None
    ours = {}
    others = []
    for i, line in enumerate(lines):
        code = line.split(";", 1)[0].strip()
        m = re.match(r"^(_Z\S+):$", code)
        if m:
            fn = m.group(1)
            continue
        if not re.search(r"\bm0\b", code):
            continue
        nxt = lines[i + 1].split(";", 1)[0].strip() if i + 1 < len(lines) else ""
        if re.fullmatch(r"s_mov_b32 m0, s\d+", code) and nxt.startswith("buffer_load_dwordx4") \
                and nxt.endswith(" lds"):
            ours[fn] = ours.get(fn, 0) + 1
        else:
            others.append((fn, code))
    assert not others, f"m0 used outside the LDS-DMA helper: {others[:8]}"
    # the MFMA kernels (3-way versions 2-4, 2-way) all go through the helper
    kernels = {k for k in ours}
    for name in ("k_triplets_mfma2", "k_triplets_mfma3", "k_triplets_mfma4", "k_pairs_mfma"):
        assert any(name in k for k in kernels), (name, sorted(kernels))
