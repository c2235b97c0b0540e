"""Parity of the batching form (GPU).

miniAMR_batch.x steps several runs with different command lines (block
shapes of every stencil shape class, variable counts that leave partial
groups of 8 and partial stencil pairs, different refinement levels,
seeds, objects and run lengths) through one persistent kernel per stage
phase and one graph per timestep.  Each run must come out as it does
alone:
  FP64: against the CPU oracle run alone -- every cell bit-identical,
        checksums within 1e-12 relative (the device sums in a different
        association), same Summary line;
  FP32 / BF16: against the single-run GPU driver -- cells bit-identical;
  every launch shape (wavefront or workgroup items, persistent grids,
  variables per item): identical cells.
"""
import pytest

import amrcheck as A
from conftest import require

pytestmark = pytest.mark.gpu

OBJ = "--num_objects 1 --object %d 0 %s 0 0 0"
RUNS = [
    # 4^3: shape class 0, 2 variables
    "--nx 4 --ny 4 --nz 4 --num_vars 2 --num_refine 2 --num_tsteps 4 --refine_freq 2 "
    "--max_blocks 400 " + OBJ % (2, "0.3 0.3 0.3 0.05 0.05 0.05 0.2 0.2 0.2") + " --seed 7",
    # 8^3, 10 variables: an exchange group of 8 plus one of 2
    "--init_x 2 --init_y 2 --init_z 2 --nx 8 --ny 8 --nz 8 --num_vars 10 --num_refine 2 "
    "--num_tsteps 3 --refine_freq 1 --max_blocks 600 "
    + OBJ % (2, "0.5 0.5 0.5 0.03 -0.02 0.01 0.3 0.3 0.3") + " --seed 3",
    # 12^3: shape class 1, 5 variables (odd: a lone variable per stencil pair)
    "--nx 12 --ny 12 --nz 12 --num_vars 5 --num_refine 1 --num_tsteps 4 --refine_freq 2 "
    "--max_blocks 100 " + OBJ % (3, "0.6 0.4 0.4 -0.05 0.02 0.05 0.15 0.15 0.15"),
    # 16^3: shape class 2
    "--init_x 2 --nx 16 --ny 16 --nz 16 --num_vars 3 --num_refine 1 --num_tsteps 4 "
    "--refine_freq 2 --max_blocks 100 " + OBJ % (2, "0.4 0.5 0.5 0.04 0.01 0.0 0.2 0.2 0.2")
    + " --seed 11",
    # non-cubic 10 x 8 x 14: class 1 with different in-plane extents
    "--nx 10 --ny 8 --nz 14 --num_vars 4 --num_refine 2 --num_tsteps 2 --refine_freq 1 "
    "--max_blocks 300 " + OBJ % (0, "0.45 0.45 0.45 0.02 0.02 0.02 0.25 0.25 0.25") + " --seed 5",
]
COMMON27 = " --stencil 27 --error_tol -10 --stages_per_ts 6 --checksum_freq 2"
COMMON7 = " --stencil 7 --permute --stages_per_ts 6 --checksum_freq 3"
CODE1 = " --stencil 7 --code 1 --stages_per_ts 6 --checksum_freq 3"


@pytest.fixture(scope="module")
def exes():
    require(A.PRODUCT_BATCH_X, "run __graft_entry__.build()")
    require(A.PRODUCT_X, "run __graft_entry__.build()")
    require(A.ORACLE_X, "run __graft_entry__.build()")
    return A.PRODUCT_BATCH_X


# launch shapes: vpw, stencil grid, exchange grid, workgroup-mode classes,
# variables and threads per exchange item
SHAPES = {"wave": "1,0,0,0,8,64", "wg_all_persistent": "2,512,256,7,4,256",
          "mixed": "1,1024,0,5,16,128", "tuned": None}


@pytest.mark.parametrize("common", [COMMON27, COMMON7, CODE1],
                         ids=["27pt", "7pt_permute", "7pt_code1"])
@pytest.mark.parametrize("shape", sorted(SHAPES))
def test_batch_f64_matches_oracle(exes, common, shape):
    print("case: batch f64 %s shape=%s" % (common.split()[1], shape), flush=True)
    runs = [r + common for r in RUNS]
    b = A.run_batch(exes, runs, dump="cells", timeout=300,
                    flags=("--shape", SHAPES[shape]) if SHAPES[shape] else ())
    assert b["rc"] == 0, b["stdout"][-2000:]
    for i, args in enumerate(runs):
        s = A.run(A.ORACLE_X, args, dump="cells", timeout=300)
        assert s["rc"] == 0
        rep = A.same_run(b["runs"][i], s, sum_rtol=1e-12, max_ulp=0)
        print("run %d: %s" % (i, rep), flush=True)


@pytest.mark.parametrize("prec", ["f32", "bf16"])
def test_batch_narrow_matches_single_runs(exes, prec):
    print("case: batch %s" % prec, flush=True)
    runs = [r + COMMON27 + " --precision " + prec for r in RUNS]
    b = A.run_batch(exes, runs, dump="cells", timeout=300)
    assert b["rc"] == 0, b["stdout"][-2000:]
    for i, args in enumerate(runs):
        s = A.run(A.PRODUCT_X, args, dump="cells", timeout=300)
        assert s["rc"] == 0
        rep = A.same_run(b["runs"][i], s, sum_rtol=1e-12, max_ulp=0)
        print("run %d: %s" % (i, rep), flush=True)


@pytest.mark.parametrize("shape", ["wave", "wg_all_persistent", "mixed"])
def test_batch_shapes_do_not_change_results(exes, shape):
    """Cells of the FP32 batch do not depend on the launch shape (the same
    item bodies run in every shape); checksums only through the association
    of the per-block partial sums (workgroup vs wavefront items)."""
    print("case: batch f32 tuned vs %s" % shape, flush=True)
    runs = [r + COMMON27 + " --precision f32" for r in RUNS]
    a = A.run_batch(exes, runs, dump="cells", timeout=300)
    c = A.run_batch(exes, runs, dump="cells", timeout=300, flags=("--shape", SHAPES[shape]))
    assert a["rc"] == 0 and c["rc"] == 0
    for i in range(len(runs)):
        A.same_run(a["runs"][i], c["runs"][i], sum_rtol=1e-13, max_ulp=0)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_batch_fused_and_plain_stages_agree(exes, prec):
    """27-point timesteps of an even number of stages run fused (same-level
    faces read in place, as amr_dev_stage_fused) unless MINIAMR_BATCH_FUSED=0;
    both give the same cells, ghosts included (the stale ghost planes of
    in-place faces are brought up to date before each refinement and before
    the dumps), and checksums up to the association of the partial sums."""
    print("case: batch %s fused vs plain" % prec, flush=True)
    runs = [r + COMMON27 + " --precision " + prec for r in RUNS]
    a = A.run_batch(exes, runs, dump="cells", timeout=300)
    c = A.run_batch(exes, runs, dump="cells", timeout=300, env={"MINIAMR_BATCH_FUSED": "0"})
    assert a["rc"] == 0 and c["rc"] == 0
    for i in range(len(runs)):
        A.same_run(a["runs"][i], c["runs"][i], sum_rtol=1e-13, max_ulp=0)
