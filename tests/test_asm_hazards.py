"""Static check of the FP64 kernels' scalar column loads (CPU only: compiles to gfx950
assembly with hipcc).  Each column is read by an asm statement of back-to-back scalar
loads from one base address; a later load of the group must not read a base register
that an earlier load of the group writes (its data returns asynchronously)."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "gems_amd", "csrc")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

LOAD = re.compile(r"\s*s_load_dwordx(?:16|8|4)\s+s\[(\d+):(\d+)\],\s*s\[(\d+):(\d+)\],\s*(0x[0-9a-f]+)\s*$")


def groups(asm):
    lines = asm.split("\n")
    out = []
    i = 0
    while i < len(lines):
        m = LOAD.match(lines[i])
        if m and m.group(5) == "0x0":
            grp = [m]
            j = i + 1
            while j < len(lines):
                n = LOAD.match(lines[j])
                if not n or n.group(5) == "0x0" or n.group(3, 4) != m.group(3, 4):
                    break
                grp.append(n)
                j += 1
            if len(grp) > 1:
                out.append(grp)
            i = j
        else:
            i += 1
    return out


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
@pytest.mark.parametrize("src", ["kernels_f64.hip", "kernels_batch.hip"])
def test_column_loads_do_not_overwrite_their_base(src, tmp_path):
    out = tmp_path / "k.s"
    cmd = [HIPCC, "-O3", "-std=c++17", "-I" + os.path.join(ROOT, "include"), "-I" + SRC, "--offload-arch=gfx950",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "--offload-device-only", "-S", "-c",
           os.path.join(SRC, src), "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True, timeout=600)
    gs = groups(out.read_text())
    assert len(gs) > 0
    for g in gs:
        b0, b1 = int(g[0].group(3)), int(g[0].group(4))
        for ld in g[:-1]:  # loads followed by another load of the same base
            d0, d1 = int(ld.group(1)), int(ld.group(2))
            assert b1 < d0 or b0 > d1, "scalar load group overwrites its own base address"
