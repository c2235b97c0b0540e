"""The batching form's own workload on the GPU: the 64 runs of
miniamr_amd/configs/batch64.txt (8^3/12^3/16^3 blocks, 8/20/40 variables,
2-4 refinement levels, seeds 1-64, 27-point, FP32 storage) stepped as one
batch by miniAMR_batch.x, checked run by run against FP64 fixtures of the
same command lines from the CPU oracle (tests/golden/batch64_f64.npz,
written by tests/golden/make_batch_golden.py):

* the Summary's blocks per timestep and max_blocks equal the FP64 run's
  (refinement depends on the objects only), as does the final active
  block count;
* every checksum of every run (initial sums and one row per 5 stages)
  within 1e-6 relative of the FP64 checksum;
* the first 512 seeded sample cells of each run (ghosts included) within
  the FP32 cell bound of the precision form (test_gpu_precision.py):
  200 * 28 * 2^-24 * max |cell|.
"""
import os

import numpy as np
import pytest

import amrcheck as A
from conftest import require

pytestmark = pytest.mark.gpu

TABLE = os.path.join(A.REPO, "miniamr_amd", "configs", "batch64.txt")
FIXTURE = os.path.join(A.GOLDEN, "batch64_f64.npz")


def table():
    with open(TABLE) as f:
        return [l.strip() for l in f if l.strip() and not l.startswith("#")]


def test_batch64_f32_matches_f64_fixtures():
    require(A.PRODUCT_BATCH_X, "run __graft_entry__.build()")
    gold = np.load(FIXTURE)
    runs = table()
    assert len(runs) == 64 and all("--precision f32" in r for r in runs)
    print("case: batch64 f32 (64 runs, one batch)", flush=True)
    b = A.run_batch(A.PRODUCT_BATCH_X, runs, timeout=300, sample=True)
    assert b["rc"] == 0, b["stdout"][-2000:]
    off = gold["sums_off"]
    worst_sum = worst_cell = 0.0
    shapes = set()
    for i, r in enumerate(b["runs"]):
        ref_sums = gold["sums"][off[i]:off[i + 1]]
        assert r["summary"] is not None, "run %d: no Summary line" % i
        assert r["summary"]["max_blocks"] == gold["max_blocks"][i], i
        assert abs(r["summary"]["blocks_per_ts"] - gold["blocks_per_ts"][i]) <= 1e-9 * \
            gold["blocks_per_ts"][i], i
        assert len(r["sums"]) == len(ref_sums), i
        e = A.rel_err(r["sums"], ref_sums).max()
        assert e <= 1e-6, "run %d: checksum rel err %.3e" % (i, e)
        smp = r["sample"]
        assert smp["n_active"] == gold["n_active"][i], i
        k = gold["sample_val"].shape[1]
        assert np.array_equal(smp["in"][:k], gold["sample_in"][i])
        assert np.array_equal(smp["var"][:k], gold["sample_var"][i])
        assert np.array_equal(smp["ijk"][:k], gold["sample_idx"][i])
        ref = gold["sample_val"][i]
        c = np.abs(smp["val"][:k] - ref).max()
        bound = 200 * 28 * 2.0 ** -24 * np.abs(ref).max()
        assert c <= bound, "run %d: cell err %.3e > %.3e" % (i, c, bound)
        worst_sum, worst_cell = max(worst_sum, e), max(worst_cell, c)
        a = table()[i].split()
        shapes.add((int(a[a.index("--nx") + 1]), int(a[a.index("--num_vars") + 1])))
    assert {n for n, _ in shapes} == {8, 12, 16} and {v for _, v in shapes} == {8, 20, 40}
    print("batch64 f32 vs FP64 fixtures: max checksum rel err %.3e, max sampled cell abs err "
          "%.3e over 64 runs" % (worst_sum, worst_cell), flush=True)
