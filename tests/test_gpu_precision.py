"""The precision form on the bandwidth-scale mesh (GPU): the 4x4x4 x 16^3,
40-variable, 27-point, 4-level workload of tests/golden/cfg1 with FP32 and
with BF16 storage (sums in FP32, checksums in FP64), against the FP64
reference's golden vectors.

* Block counts after every refinement equal the reference's (refinement
  looks at object positions only, never at cell values).
* Checksums: all 1,640 (40 initial, 40 per event at every 5th of 200
  stages).  FP32 storage must stay within 1e-6 relative of the reference at
  every event; BF16 is reported and held to a loose bound.
* Cells: the 16,384 seeded sample cells (ghosts included) against the
  reference's FP64 values.  Bound: each stage rounds every cell at most
  ~27 times in FP32 (26 adds and a divide, each <= 2^-24 of the running
  magnitude) plus once into storage (2^-24 in FP32, 2^-9 in BF16); the
  27-point averages and the restriction sums propagate an error at most
  with the magnitudes they propagate values, so after 200 stages the
  absolute error is bounded by 200 * (27 * 2^-24 + u_storage) * M, M the
  largest magnitude in the FP64 sample.
"""
import numpy as np
import pytest

import amrcheck as A
import miniamr_amd as M
from conftest import require

pytestmark = pytest.mark.gpu

U_STORAGE = {"f32": 2.0 ** -24, "bf16": 2.0 ** -9}
STAGES = 200


@pytest.mark.parametrize("prec", ["f32", "bf16"])
def test_precision_form_on_the_bandwidth_scale_mesh(prec):
    require(M.EXE_PATH, "run __graft_entry__.build()")
    g = A.golden("cfg1")
    assert "sample_in" in g["bsums"], "cfg1 golden without the cell sample"
    print("case: precision %s on cfg1" % prec, flush=True)
    r = A.run(M.EXE_PATH, g["args"] + " --precision " + prec, timeout=110, sample=True)
    assert r["rc"] == 0, r["stdout"][-2000:]
    assert r["counts"] == g["counts"], "block counts differ from the reference"
    assert len(r["sums"]) == len(g["sums"]) == 1640
    e = A.rel_err(r["sums"], g["sums"])
    per_event = e[40:].reshape(40, 40).max(axis=1)
    smp, b = r["sample"], g["bsums"]
    assert np.array_equal(smp["in"], b["sample_in"]) and np.array_equal(smp["ijk"],
                                                                         b["sample_idx"])
    cell_err = np.abs(smp["val"] - b["sample_val"])
    mag = np.abs(b["sample_val"]).max()
    bound = STAGES * (27 * 2.0 ** -24 + U_STORAGE[prec]) * mag
    print("precision %s: checksums max rel err %.3e (per event: first %.3e, last %.3e); "
          "16384 sampled cells: max abs err %.3e, mean %.3e, bound %.3e (max |cell| %.3f)"
          % (prec, e.max(), per_event[0], per_event[-1], cell_err.max(), cell_err.mean(),
             bound, mag), flush=True)
    assert cell_err.max() <= bound
    if prec == "f32":
        assert e.max() <= 1e-6, "FP32 checksum rel err %.3e at event %d" % (
            e.max(), int(e.argmax()))
    else:
        assert e.max() <= 2e-2


def test_stage_stepping_and_cell_comparison(tmp_path):
    """amr_run_stages (chunks ending at the checksum stages) gives the
    cells and checksums of amr_run_step; amr_dev_compare_cells is 0
    between those two FP64 runs and, against an FP32 run of the same
    command line, positive and within the FP32 cell bound at every
    checksum stage."""
    import ctypes
    import os
    g = A.golden("t27_perm")          # 27-point, permuted directions, 2 levels
    print("case: stage stepping + compare_cells on t27_perm", flush=True)
    host, dev = M.load_host_library(), M.load_library()
    runs = []
    for tag, extra in (("step", ""), ("stages", ""), ("f32", " --precision f32")):
        os.environ["AMR_DUMP_SUMS"] = str(tmp_path / ("sums_" + tag))
        runs.append(M.Run(g["args"] + " --report_perf 0" + extra, host=host))
    os.environ.pop("AMR_DUMP_SUMS")
    a, b, c = runs
    nv = a.stats().num_vars
    out = (ctypes.c_double * nv)()

    def compare(x, y):
        assert dev.amr_dev_compare_cells(x.device(), y.device(), 0, nv, out) == 0
        return max(out)

    spt, cf = 3, 1                    # --stages_per_ts 3 --checksum_freq 1
    assert "--stages_per_ts 3" in g["args"] and "--checksum_freq 1" in g["args"]
    worst = 0.0
    rc = 0
    while rc == 0:
        rc = a.step()
        for s in range(spt):
            rb = host.amr_run_stages(b.h, 1)
            rcc = host.amr_run_stages(c.h, 1)
            assert rb >= 0 and rcc >= 0
            if s < spt - 1:
                assert rb == 0 and rcc == 0
                e = compare(b, c)
                assert 0.0 < e <= STAGES * (27 * 2.0 ** -24 + U_STORAGE["f32"]) * 4
                worst = max(worst, e)
        assert rb == rc and rcc == rc
        assert compare(a, b) == 0.0
    for r in runs:
        r.finish()
        r.close()
    sa, sb = [open(tmp_path / ("sums_" + t)).read() for t in ("step", "stages")]
    assert sa == sb and len(sa.split()) == len(g["sums"])
    print("stage stepping identical; FP32 vs FP64 max cell err %.3e" % worst, flush=True)
