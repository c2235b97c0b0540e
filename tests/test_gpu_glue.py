"""The reference's own program on the device library (GPU).

oracle/_ref/miniAMR_ref_glue.x is the reference's unmodified objects
linked with miniamr_amd/glue/dev_glue.c and libminiamr_dev.so: its driver,
refinement, report and fatal paths are the reference's; its exchange,
stencils and checksums are the device library's, reached through the
reference's own call sites (ref/driver.c -> comm, stencil_driver,
check_sum; cells uploaded after init and around refine).  Every golden
case below must come out as the reference computed it: exit code,
messages, block counts, Summary fields, checksums within 1e-12 relative
(the device sums per block, then in a tree) and every cell within 4 ulp,
ghosts included.
"""
import os

import pytest

import amrcheck as A
from conftest import require

pytestmark = pytest.mark.gpu

GLUE_X = os.path.join(A.REPO, "oracle", "_ref", "miniAMR_ref_glue.x")
CASES = ["t7_a", "t27_a", "t27_perm", "t7_r3", "t27_cv", "t27_uniform", "c1_t7", "c2_t27",
         "s0_a", "s0_b", "err_tol", "cfg4"]


@pytest.mark.parametrize("name", CASES)
def test_reference_driver_on_the_device_library(name):
    require(GLUE_X, "make -C oracle glue (builds from the reference sources)")
    g = A.golden(name)
    print("case: reference driver + device library, %s: %s" % (name, g["args"]), flush=True)
    kind = "cells" if "cells" in g else ("bsums" if "bsums" in g else None)
    r = A.run(GLUE_X, g["args"], dump=kind, timeout=110)
    check_cells = not any("difference too large" in m for m in g["messages"])
    rep = A.compare(r, g, sum_rtol=1e-12, max_ulp=4, check_cells=check_cells)
    print(name, rep, flush=True)
