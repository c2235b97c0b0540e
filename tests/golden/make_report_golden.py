#!/usr/bin/env python3
"""Golden end-of-run reports from the reference miniAMR (test data only).

Runs ``oracle/_ref/miniAMR_ref.x`` (the unmodified ``ref/`` sources, built
with ``make -C oracle ref``) with ``--report_perf 15`` on each case, in a
scratch directory, and writes ``report_<case>.json`` here: the command
line and the three reports -- stdout's report (``ref/profile.c:880-1298``),
results.txt (``:472-877``) and results.yaml (``:52-470``) -- as lines, with
the numbers that measure time or memory masked (tests/reportmask.py).

Only this container has ``/root/reference``; the fixtures are what travels.
Usage: ``python tests/golden/make_report_golden.py [case ...]``.
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))
import reportmask  # noqa: E402

REF = os.path.join(REPO, "oracle", "_ref", "miniAMR_ref.x")

CASES = {
    # 7-point, moving sphere surface, two levels
    "sphere7": "--num_refine 2 --nx 8 --ny 8 --nz 8 --num_vars 4 --num_tsteps 6 "
               "--refine_freq 2 --num_objects 1 "
               "--object 2 0 0.3 0.3 0.3 0.02 0.02 0.02 0.2 0.2 0.2 0 0 0",
    # 27-point, permuted exchanges, variable groups, three levels, two objects
    # (rectangle surface bouncing, cylinder volume growing), 7 checksum stages
    "mixed27": "--init_x 2 --init_y 1 --init_z 1 --nx 6 --ny 4 --nz 8 --num_vars 5 "
               "--comm_vars 2 --stencil 27 --permute --error_tol -10 --num_refine 3 "
               "--max_blocks 2000 --num_tsteps 5 --stages_per_ts 7 --checksum_freq 3 "
               "--refine_freq 2 --block_change 2 --num_objects 2 "
               "--object 0 1 0.4 0.4 0.4 0.06 0.03 0.02 0.2 0.1 0.15 0 0 0 "
               "--object 21 2 0.5 0.2 0.7 -0.05 0.04 0 0.3 0.1 0.1 0.01 0.01 0.01",
    # uniform refinement, no objects
    "uniform": "--init_x 2 --init_y 2 --init_z 1 --nx 4 --ny 4 --nz 4 --num_vars 2 "
               "--uniform_refine 1 --num_refine 1 --num_tsteps 3 --stages_per_ts 4 "
               "--checksum_freq 2",
    # the variable-work stencils (--stencil 0) with --code 2: the operation
    # counts include stencil_check's data-dependent ones
    "stencil0": "--nx 4 --ny 4 --nz 4 --num_vars 9 --stencil 0 --code 2 --num_refine 1 "
                "--num_tsteps 3 --stages_per_ts 6 --checksum_freq 2 --refine_freq 1 "
                "--num_objects 1 --object 2 0 0.3 0.3 0.3 0.05 0.05 0.05 0.2 0.2 0.2 0 0 0",
    # simulated time instead of a timestep count, a hemispheroid that bounces
    "timed": "--nx 4 --ny 4 --nz 4 --num_vars 3 --num_refine 2 --time 0.5 "
             "--stages_per_ts 3 --checksum_freq 1 --refine_freq 1 --num_objects 1 "
             "--object 5 2 0.5 0.5 0.5 0.3 -0.2 0.1 0.2 0.2 0.2 0 0 0",
}


def main(names):
    if not os.path.exists(REF):
        sys.exit("build the reference first: make -C oracle ref")
    for name in names or CASES:
        rec = {"args": CASES[name]}
        rec.update(reportmask.run_reports(REF, CASES[name]))
        with open(os.path.join(HERE, "report_%s.json" % name), "w") as f:
            json.dump(rec, f, indent=0)
        print(name, len(rec["stdout"]), len(rec["txt"]), len(rec["yaml"]))


if __name__ == "__main__":
    main(sys.argv[1:])
