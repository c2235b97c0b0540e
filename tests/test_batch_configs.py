"""The 64-run workload of the batching form (miniamr_amd/configs/batch64.txt).

Checks on the CPU that the committed table is what make_batch64.py
generates, that it cycles block sizes, variable counts and refinement
levels as documented, and -- by running each command line's mesh history
through the CPU oracle with one variable and one stage per timestep (the
mesh does not depend on either) -- that every run keeps between 50 and
2000 active blocks and fits in its --max_blocks.
"""
import os
import re
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

import amrcheck as A
from conftest import require

CFG_DIR = os.path.join(A.REPO, "miniamr_amd", "configs")
TABLE = os.path.join(CFG_DIR, "batch64.txt")


def table():
    with open(TABLE) as f:
        return [l.strip() for l in f if l.strip() and not l.startswith("#")]


def opt(args, flag):
    a = args.split()
    return a[a.index(flag) + 1]


def test_table_is_generated():
    sys.path.insert(0, CFG_DIR)
    try:
        import make_batch64
    finally:
        sys.path.pop(0)
    assert make_batch64.runs() == table()


def test_table_shape():
    runs = table()
    assert len(runs) == 64
    for i, r in enumerate(runs):
        assert int(opt(r, "--nx")) == (8, 12, 16)[i % 3]
        assert int(opt(r, "--num_vars")) == (8, 20, 40)[(i // 3) % 3]
        assert int(opt(r, "--num_refine")) == (2, 3, 4)[(i // 9) % 3]
        assert int(opt(r, "--seed")) == i + 1
        assert opt(r, "--precision") == "f32" and opt(r, "--stencil") == "27"


def topology(args):
    a = args.split()
    for flag, val in (("--num_vars", "1"), ("--stages_per_ts", "1"), ("--checksum_freq", "1"),
                      ("--precision", "f64")):
        a[a.index(flag) + 1] = val
    out = subprocess.run([A.ORACLE_X] + a + ["--report_perf", "12"], stdout=subprocess.PIPE,
                         timeout=300).stdout.decode()
    tot = [int(x) for x in re.findall(r"Total number of blocks at timestep \d+ is (\d+)", out)]
    m = re.search(r"max_blocks (\d+)", out)
    return tot, int(m.group(1)) if m else None, out


def test_block_counts_in_range():
    require(A.ORACLE_X, "run __graft_entry__.build()")
    runs = table()
    with ThreadPoolExecutor(8) as ex:
        res = list(ex.map(topology, runs))
    counts = []
    for r, (tot, hw, out) in zip(runs, res):
        assert hw is not None, out[-500:]
        assert len(tot) == 3, tot          # refinement at timesteps 0, 4, 8
        assert 50 <= min(tot) and max(tot) <= 2000, (r, tot)
        assert hw <= int(opt(r, "--max_blocks")), (r, hw)
        counts += tot
    assert max(counts) / min(counts) > 10   # uneven on purpose


def test_fp64_fixtures_reproduce():
    """tests/golden/batch64_f64.npz is what the CPU oracle gives for the
    table's command lines in FP64 (make_batch_golden.py): re-run the four
    smallest runs and compare checksums, block counts and samples exactly."""
    import numpy as np
    require(A.ORACLE_X, "run __graft_entry__.build()")
    gold = np.load(os.path.join(A.GOLDEN, "batch64_f64.npz"))
    runs = table()
    small = sorted(range(64), key=lambda i: gold["blocks_per_ts"][i] *
                   int(opt(runs[i], "--nx")) ** 3 * int(opt(runs[i], "--num_vars")))[:4]

    def one(i):
        return i, A.run(A.ORACLE_X, runs[i].replace("--precision f32", "--precision f64"),
                        timeout=300, sample=True)
    with ThreadPoolExecutor(4) as ex:
        res = list(ex.map(one, small))
    off, coff = gold["sums_off"], gold["counts_off"]
    for i, r in res:
        assert r["rc"] == 0
        assert r["sums"] == list(gold["sums"][off[i]:off[i + 1]]), i
        assert r["counts"] == gold["counts"][coff[i]:coff[i + 1]].tolist(), i
        assert r["summary"]["max_blocks"] == gold["max_blocks"][i]
        k = gold["sample_val"].shape[1]
        assert np.array_equal(r["sample"]["val"][:k], gold["sample_val"][i]), i
