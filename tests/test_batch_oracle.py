"""The batch driver on the CPU oracle (no GPU).

miniAMR_oracle_batch.x is the product's batch driver (host/batch_run.c)
linked with the CPU restatement, whose amr_batch steps each member alone.
Running several different command lines as one batch must give each run
exactly what it gives alone: this pins the batch driver's lockstep logic
(per-run refinement, checksum rows, runs that finish early, --seed) before
any GPU kernel is involved.
"""
import pytest

import amrcheck as A
from conftest import require

OBJ = "--num_objects 1 --object 2 0 %s 0 0 0"
RUNS27 = [
    "--nx 4 --ny 4 --nz 4 --num_vars 2 --num_refine 2 --num_tsteps 4 --refine_freq 2 "
    "--max_blocks 400 " + OBJ % "0.3 0.3 0.3 0.05 0.05 0.05 0.2 0.2 0.2" + " --seed 7",
    "--init_x 2 --nx 6 --ny 6 --nz 6 --num_vars 3 --num_refine 1 --num_tsteps 3 --refine_freq 2 "
    "--max_blocks 400 " + OBJ % "0.6 0.4 0.4 -0.05 0.02 0.05 0.15 0.15 0.15",
    "--init_x 2 --init_y 2 --init_z 2 --nx 8 --ny 8 --nz 8 --num_vars 10 --num_refine 2 "
    "--num_tsteps 4 --refine_freq 1 --max_blocks 600 "
    + OBJ % "0.5 0.5 0.5 0.03 -0.02 0.01 0.3 0.3 0.3" + " --seed 3",
]
COMMON27 = " --stencil 27 --error_tol -10 --stages_per_ts 6 --checksum_freq 2"
COMMON7 = " --stencil 7 --permute --stages_per_ts 6 --checksum_freq 3"


@pytest.mark.parametrize("common", [COMMON27, COMMON7], ids=["27pt", "7pt_permute"])
def test_oracle_batch_equals_single_runs(common):
    require(A.ORACLE_BATCH_X, "run __graft_entry__.build()")
    runs = [r + common for r in RUNS27]
    b = A.run_batch(A.ORACLE_BATCH_X, runs, dump="cells", timeout=300)
    assert b["rc"] == 0, b["stdout"][-2000:]
    for i, args in enumerate(runs):
        s = A.run(A.ORACLE_X, args, dump="cells", timeout=300)
        assert s["rc"] == 0
        A.same_run(b["runs"][i], s, sum_rtol=0.0, max_ulp=0)


def test_seed_changes_only_the_values():
    require(A.ORACLE_X, "run __graft_entry__.build()")
    base = RUNS27[1] + COMMON27
    a = A.run(A.ORACLE_X, base + " --seed 1", dump="cells")
    d = A.run(A.ORACLE_X, base, dump="cells")
    c = A.run(A.ORACLE_X, base + " --seed 2", dump="cells")
    A.same_run(a, d)                     # --seed 1 is the reference's default
    assert a["summary"] == c["summary"]  # same mesh history
    assert a["sums"][0] != c["sums"][0]  # different initial values
