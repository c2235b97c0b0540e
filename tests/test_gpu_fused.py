"""Stages that read same-level faces in place (amr_dev_stage_fused).

The drop-in driver runs every stage of a timestep as a fused stage but the
last one before a refinement or the end of the run (by default when a stage
has >= 1e8 cell updates; MINIAMR_FUSED=2 at every size, as here): the
stencil reads same-level neighbours'
interiors instead of ghost copies, writing to the second interior copy of
each record.  Every checksum and every final cell (ghosts included) must
be bit-identical to the plain stages' run, on every golden mesh (7- and
27-point, permuted exchanges -- where the fused stage is the plain one --,
variable groups, refinement with stale ghosts), and the golden records of
the reference must still hold.  By default the fused stage also drops the
27-point edge extensions of the faces it reads in place (the stencil and
the remaining jobs read those edge cells where the extension would have
copied them from); the plan that keeps them as jobs is tested too.  A
context also refuses the calls that need
written ghost planes until a plain exchange has run.
"""
import os

import numpy as np
import pytest

import amrcheck as A
import meshgen
import miniamr_amd as M
from conftest import require

CASES = sorted(f[:-5] for f in os.listdir(A.GOLDEN)
               if f.endswith(".json") and not f.startswith("report_"))
# "2": fused at every size (by default only stages of >= 1e8 cell updates
# are); "0": the plain stages
FUSED = {"MINIAMR_FUSED": "2"}

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", CASES)
def test_fused_stages_are_identical(name):
    require(M.EXE_PATH, "run __graft_entry__.build()")
    g = A.golden(name)
    print("case: fused %s: %s" % (name, g["args"]), flush=True)
    kind = "cells" if "cells" in g else ("bsums" if "bsums" in g else None)
    if kind == "bsums" and "vars" in g["bsums"]:
        kind = "bsums_only"
    sample = "bsums" in g and "sample_in" in g["bsums"]
    a = A.run(M.EXE_PATH, g["args"], dump=kind, timeout=110, env={"MINIAMR_FUSED": "0"},
              sample=sample)
    b = A.run(M.EXE_PATH, g["args"], dump=kind, timeout=110, env=FUSED, sample=sample)
    assert_same(a, b)
    check_cells = not any("difference too large" in m for m in g["messages"])
    A.compare(b, g, sum_rtol=1e-12, max_ulp=4, check_cells=check_cells)


def assert_same(a, b):
    assert a["rc"] == b["rc"]
    assert a["sums"] == b["sums"]
    assert a["counts"] == b["counts"]
    for k in ("cells", "bsums", "sample"):
        if k in a:
            for key in a[k]:
                assert np.array_equal(a[k][key], b[k][key]), (k, key)


# The narrow storage forms: fused and plain stages must agree bit for bit
# in every precision (the ghosts are exact copies), with variable counts
# that leave partial stencil groups (2, 3, 5 variables against 4 per
# workgroup) and partial 16-variable exchange items, the bandwidth-scale
# workload itself (40 variables, fused by default), and the launch-shape
# overrides.
NARROW = [("t27_a", None), ("t27_cv", None), ("t7_r3", None), ("cfg4", None), ("cfg1", None),
          ("t27_cv", {"MINIAMR_STENCIL_VPW": "1", "MINIAMR_EXCHANGE": "16,256"}),
          ("cfg4", {"MINIAMR_STENCIL_VPW": "2", "MINIAMR_EXCHANGE": "4,128"})]


@pytest.mark.parametrize("prec", ["f32", "bf16"])
@pytest.mark.parametrize("name,extra", NARROW,
                         ids=["%s%s" % (n, "-override" if e else "") for n, e in NARROW])
def test_fused_stages_are_identical_in_narrow_storage(name, extra, prec):
    require(M.EXE_PATH, "run __graft_entry__.build()")
    g = A.golden(name)
    print("case: fused %s %s %s" % (prec, name, extra or ""), flush=True)
    kind = "cells" if "cells" in g else "bsums_only"
    env = dict(extra or {}, MINIAMR_PRECISION=prec)
    a = A.run(M.EXE_PATH, g["args"], dump=kind, timeout=110, env=dict(env, MINIAMR_FUSED="0"))
    b = A.run(M.EXE_PATH, g["args"], dump=kind, timeout=110, env=dict(env, **FUSED))
    assert len(a["sums"]) > 0
    assert_same(a, b)


# Without edge folding (MINIAMR_FOLD_EDGES=0) the 27-point edge extensions
# of the faces read in place stay exchange jobs; that plan must agree too.
@pytest.mark.parametrize("name", ["t27_a", "t27_cv", "cfg4", "cfg1"])
def test_fused_stages_without_edge_folding(name):
    require(M.EXE_PATH, "run __graft_entry__.build()")
    g = A.golden(name)
    print("case: fused, extensions as jobs %s" % name, flush=True)
    kind = "cells" if "cells" in g else "bsums_only"
    a = A.run(M.EXE_PATH, g["args"], dump=kind, timeout=110, env={"MINIAMR_FUSED": "0"})
    b = A.run(M.EXE_PATH, g["args"], dump=kind, timeout=110,
              env=dict(FUSED, MINIAMR_FOLD_EDGES="0"))
    assert_same(a, b)


# Meshes of 16 x 16-face blocks, where the stencil is the vector form:
# refinement around moving objects gives coarse/fine faces in all three
# directions, boundary blocks and stale ghosts.  Fused and plain stages must
# agree bit for bit, cells and ghosts included, in every precision, with
# edge folding switched off, with every job in its own direction's launch, and
# with a plain stage at the end of every timestep (by default only before a
# refinement and at the end of the run).
VEC_MESHES = [
    "--nx 16 --ny 16 --nz 16 --init_x 2 --init_y 2 --init_z 2 --num_vars 6 --stencil 27 "
    "--num_tsteps 4 --stages_per_ts 6 --checksum_freq 2 --num_refine 2 --refine_freq 2 "
    "--max_blocks 800 --error_tol -10 --num_objects 1 "
    "--object 2 0 0.3 0.3 0.3 0.05 0.04 0.03 0.2 0.2 0.2 0 0 0",
    "--nx 8 --ny 16 --nz 16 --init_x 2 --init_y 1 --init_z 2 --num_vars 5 --stencil 27 "
    "--num_tsteps 3 --stages_per_ts 5 --checksum_freq 1 --num_refine 3 --refine_freq 1 "
    "--max_blocks 1200 --error_tol -10 --num_objects 2 "
    "--object 2 0 0.6 0.4 0.5 -0.04 0.03 0.02 0.15 0.25 0.2 0 0 0 "
    "--object 0 0 0.2 0.7 0.3 0.03 -0.02 0.04 0.1 0.1 0.1 0 0 0",
]
VEC_ENVS = [{}, {"MINIAMR_PRECISION": "f32"}, {"MINIAMR_PRECISION": "bf16"},
            {"MINIAMR_FOLD_EDGES": "0"}, {"MINIAMR_HOIST": "0"}, {"MINIAMR_PLAIN_LAST": "1"}]


@pytest.mark.parametrize("env", VEC_ENVS,
                         ids=["f64", "f32", "bf16", "no-edge-fold", "no-hoist", "plain-last"])
@pytest.mark.parametrize("mesh", range(len(VEC_MESHES)))
def test_fused_stages_vector_stencil(mesh, env):
    require(M.EXE_PATH, "run __graft_entry__.build()")
    args = VEC_MESHES[mesh]
    print("case: fused vector mesh %d %s" % (mesh, env), flush=True)
    a = A.run(M.EXE_PATH, args, dump="cells", timeout=110, env=dict(env, MINIAMR_FUSED="0"))
    b = A.run(M.EXE_PATH, args, dump="cells", timeout=110, env=dict(env, **FUSED))
    assert a["rc"] == 0, a["stdout"][-2000:]
    assert len(a["sums"]) > 0 and len(a["counts"]) > 0
    assert_same(a, b)


def test_stale_ghosts_are_brought_up_to_date(monkeypatch):
    """After fused stages, the calls that look at ghost planes first give
    them the values the plain stages would have written (from the interiors
    the last fused stage read): downloads after fused and after plain stages
    agree bit for bit, ghosts included, and set_mesh accepts the mesh."""
    require(M.LIB_PATH, "run __graft_entry__.build()")
    monkeypatch.setenv("MINIAMR_FUSED", "2")
    for st in (27, 7):
        rng = np.random.default_rng(5)
        cells = [rng.random((3, 8, 6, 10)) for _ in range(11)]
        plan = meshgen.random_plan(rng, 5, 6)
        order = list(range(11))
        devs = []
        for fused in (True, False):
            d = M.Device(nx=6, ny=4, nz=8, num_vars=3, max_blocks=14, stencil=st)
            for s_, c in enumerate(cells):
                d.upload(s_, c)
            d.set_mesh(order, plan)
            for k in range(3):
                (d.stage_fused if fused else d.stage)(0, 3, stage=k)
            devs.append(d)
        for s_ in order:
            a, b = devs[0].download(s_), devs[1].download(s_)
            assert np.array_equal(a, b), (st, s_)
        devs[0].set_mesh(order, plan)
        for d in devs:
            d.close()
