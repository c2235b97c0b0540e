"""Host-side check of the cell-record layout (CPU only).

Builds ``tests/layout_check.cpp`` against ``miniamr_amd/csrc/amr_internal.h``
with hipcc (host compilation only) and runs it: every canonical cell of
the reference's (nx+2)(ny+2)(nz+2) block array (ref/block.h:52) has exactly
one home in a record, the z-plane copies and the second interior overlap
nothing, and every part starts on a 128-byte line.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or ("/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else None)


@pytest.mark.skipif(HIPCC is None, reason="hipcc is not installed")
def test_record_layout_is_injective_and_line_aligned(tmp_path):
    exe = tmp_path / "layout_check"
    subprocess.run(
        [HIPCC, "-x", "hip", "--offload-host-only", "-std=c++17", "-O1",
         "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "miniamr_amd", "csrc"),
         os.path.join(ROOT, "tests", "layout_check.cpp"), "-o", str(exe)],
        check=True, capture_output=True, timeout=300)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "48 layouts checked, 0 failures" in out.stdout
