"""End-of-run reports (ref/profile.c:37-1300) against the reference's.

The golden reports (tests/golden/report_*.json, from
tests/golden/make_report_golden.py) are the reference build's stdout
report, results.txt and results.yaml for four runs -- 7- and 27-point,
permuted exchanges and variable groups, uniform refinement, simulated time,
five object types -- with the numbers that measure time or memory masked.
Every other line must be identical: configuration, object motion, faces
exchanged per direction, refinement steps, block counts, floating-point
operations, and the per-timestep block counts printed before the report.
The CPU oracle's driver is the product's host code, so these tests pin the
report on the CPU; the GPU test runs the drop-in binary itself.
"""
import difflib
import json
import os

import pytest

from amrcheck import GOLDEN, ORACLE_X, PRODUCT_X
import reportmask

CASES = ["sphere7", "mixed27", "uniform", "timed", "stencil0"]


def check(exe, name):
    with open(os.path.join(GOLDEN, "report_%s.json" % name)) as f:
        gold = json.load(f)
    ours = reportmask.run_reports(exe, gold["args"])
    for part in ("stdout", "txt", "yaml", "progress"):
        diff = list(difflib.unified_diff(gold[part], ours[part], lineterm="", n=1))
        assert not diff, "%s %s differs:\n%s" % (name, part, "\n".join(diff[:40]))


@pytest.mark.parametrize("name", CASES)
def test_report_matches_reference(name):
    if not os.path.exists(ORACLE_X):
        pytest.skip("oracle not built (make -C oracle oracle)")
    check(ORACLE_X, name)


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_gpu_report_matches_reference(name):
    print("case: report %s" % name, flush=True)
    check(PRODUCT_X, name)
