#!/usr/bin/env python3
"""FP64 fixtures for the batching form's workload (test data only).

Runs every command line of ``miniamr_amd/configs/batch64.txt`` alone on the
CPU oracle (``oracle/_build/miniAMR_oracle.x``: the product's host driver
linked with the C restatement of the hot path, pinned against the
reference build by the other fixtures here) in FP64 -- the table's
``--precision f32`` replaced by ``f64`` -- and writes
``batch64_f64.npz``:

* ``sums`` / ``sums_off``: every checksum of run i (initial sums first, in
  call order) is ``sums[sums_off[i]:sums_off[i+1]]``;
* ``blocks_per_ts``, ``max_blocks``: the Summary line's fields per run;
* ``counts`` / ``counts_off``: the per-level block counts after every
  refinement, rows of (timestep, level, count);
* ``n_active``: active blocks at the end of each run;
* ``sample_in``, ``sample_var``, ``sample_idx``, ``sample_val``: the first
  ``PREFIX`` cells of each run's seeded AMR_DUMP_SAMPLE draws (ghosts
  included), shape (runs, PREFIX[, 3]).

The reference build has no ``--seed`` flag, so these runs cannot come from
it; ``--seed s`` seeds the reference's generator as ``srand(s)`` would.
Usage: ``python tests/golden/make_batch_golden.py [workers]`` (~5 minutes
with 6 workers).
"""
import os
import sys
from concurrent.futures import ProcessPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "tests"))

import amrcheck as A  # noqa: E402

TABLE = os.path.join(REPO, "miniamr_amd", "configs", "batch64.txt")
PREFIX = 512


def table():
    with open(TABLE) as f:
        return [l.strip() for l in f if l.strip() and not l.startswith("#")]


def f64(args):
    return args.replace("--precision f32", "--precision f64")


def one(args):
    r = A.run(A.ORACLE_X, f64(args), timeout=3600, sample=True)
    if r["rc"] != 0:
        raise RuntimeError("oracle failed on %s:\n%s" % (args, r["stdout"][-2000:]))
    return r


def main(workers):
    runs = table()
    with ProcessPoolExecutor(workers) as ex:
        recs = list(ex.map(one, runs))
    off = np.cumsum([0] + [len(r["sums"]) for r in recs]).astype(np.int64)
    np.savez_compressed(
        os.path.join(HERE, "batch64_f64.npz"),
        sums=np.concatenate([np.array(r["sums"]) for r in recs]), sums_off=off,
        blocks_per_ts=np.array([r["summary"]["blocks_per_ts"] for r in recs]),
        max_blocks=np.array([r["summary"]["max_blocks"] for r in recs], dtype=np.int32),
        counts=np.concatenate([np.array(r["counts"], dtype=np.int32).reshape(-1, 3)
                               for r in recs]),
        counts_off=np.cumsum([0] + [len(r["counts"]) for r in recs]).astype(np.int64),
        n_active=np.array([r["sample"]["n_active"] for r in recs], dtype=np.int32),
        sample_in=np.stack([r["sample"]["in"][:PREFIX] for r in recs]).astype(np.int32),
        sample_var=np.stack([r["sample"]["var"][:PREFIX] for r in recs]).astype(np.int32),
        sample_idx=np.stack([r["sample"]["ijk"][:PREFIX] for r in recs]).astype(np.int32),
        sample_val=np.stack([r["sample"]["val"][:PREFIX] for r in recs]))
    print("%d runs, %d checksums" % (len(recs), off[-1]))


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 6)
