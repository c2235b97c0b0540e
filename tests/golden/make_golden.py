#!/usr/bin/env python3
"""Generate golden fixtures from the reference miniAMR (test data only).

Runs ``oracle/_ref/miniAMR_ref_dump.x`` (the unmodified ``ref/`` sources
linked with the project's PMPI harness ``oracle/ref_dump.c``; build it with
``make -C oracle ref``) on a list of cases and writes, per case, into this
directory:

* ``<case>.json``  -- command line, exit code, every checksum the run
  computed (``ref/check_sum.c:57``, in call order: the ``num_vars`` initial
  sums of ``ref/init.c:681-682`` first), the per-level block counts printed
  by ``refine`` (``ref/refine.c:178-189``), the driver's conservation
  message (``ref/driver.c:98``), error lines and the ``Summary:`` fields
  (``ref/profile.c:1298``) that do not depend on wall-clock time;
* ``<case>.cells.npz`` -- for tiny meshes, every active block's full cell
  array (ghosts included) at the end of the run;
* ``<case>.bsums.npz`` -- for larger meshes, per (block, variable) interior
  and ghost sums at the end of the run, plus a seeded sample of cells (for
  the bandwidth-scale workload: all 40 variables' sums and the 16,384-cell
  AMR_DUMP_SAMPLE draws, ghosts included, which the product's driver
  reproduces).

Only this container has ``/root/reference``; the fixtures are what travels.
Usage: ``python tests/golden/make_golden.py [case ...]``.
"""
import json
import os
import re
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF_DUMP = os.path.join(REPO, "oracle", "_ref", "miniAMR_ref_dump.x")

SPHERE0 = "--num_objects 1 --object 2 0 0.3 0.3 0.3 0.05 0.05 0.05 0.2 0.2 0.2 0 0 0"

# name -> (argument string, dump kind: "cells" | "bsums" | None)
CASES = {
    # 7-point, one level of refinement re-evaluated every step.
    "t7_a": ("--init_x 2 --init_y 1 --init_z 1 --nx 4 --ny 4 --nz 4 --num_vars 2 "
             "--num_refine 1 --num_tsteps 3 --stages_per_ts 4 --checksum_freq 2 "
             "--refine_freq 1 " + SPHERE0, "cells"),
    # Same mesh, 27-point (edges/corners, stale ghosts at level jumps).
    "t27_a": ("--init_x 2 --init_y 1 --init_z 1 --nx 4 --ny 4 --nz 4 --num_vars 2 "
              "--stencil 27 --error_tol -10 --num_refine 1 --num_tsteps 3 "
              "--stages_per_ts 4 --checksum_freq 2 --refine_freq 1 " + SPHERE0, "cells"),
    # Non-cubic blocks, permuted exchange order, two levels, bouncing solid sphere.
    "t27_perm": ("--init_x 2 --init_y 2 --init_z 1 --nx 6 --ny 4 --nz 8 --num_vars 3 "
                 "--stencil 27 --permute --error_tol -10 --num_refine 2 --num_tsteps 4 "
                 "--stages_per_ts 3 --checksum_freq 1 --refine_freq 1 --num_objects 1 "
                 "--object 3 1 0.2 0.6 0.4 0.1 -0.08 0.06 0.15 0.2 0.1 0.01 0 0", "cells"),
    # 7-point, three levels, block_change 2, rectangle surface + cylinder surface.
    "t7_r3": ("--init_x 1 --init_y 2 --init_z 2 --nx 4 --ny 6 --nz 2 --num_vars 2 "
              "--num_refine 3 --max_blocks 3000 --num_tsteps 6 --stages_per_ts 2 --checksum_freq 1 "
              "--refine_freq 2 --block_change 2 --num_objects 2 "
              "--object 0 1 0.4 0.4 0.4 0.06 0.03 0.02 0.2 0.1 0.15 0 0 0 "
              "--object 20 2 0.5 0.2 0.7 -0.05 0.04 0 0.3 0.1 0.1 0 0 0", "cells"),
    # 27-point with variable groups (comm_vars) and refine_ghost.
    "t27_cv": ("--init_x 1 --init_y 1 --init_z 2 --nx 4 --ny 4 --nz 4 --num_vars 5 "
               "--comm_vars 2 --stencil 27 --error_tol -10 --refine_ghost --num_refine 2 "
               "--num_tsteps 4 --stages_per_ts 3 --checksum_freq 3 --refine_freq 1 "
               "--num_objects 1 --object 1 0 0.2 0.3 0.5 0.08 0.05 0.0 0.15 0.1 0.2 0 0 0",
               "cells"),
    # Uniform refinement (the driver never refines again).
    "t27_uniform": ("--init_x 1 --init_y 2 --init_z 1 --nx 4 --ny 4 --nz 4 --num_vars 2 "
                    "--stencil 27 --uniform_refine 1 --num_refine 2 --num_tsteps 2 "
                    "--stages_per_ts 3 --checksum_freq 1", "cells"),
    # --time mode: delta from object speed (ref/driver.c:171-214).
    "t7_time": ("--init_x 2 --init_y 2 --init_z 2 --nx 2 --ny 2 --nz 2 --num_vars 1 "
                "--num_refine 3 --max_blocks 3000 --time 6.0 --stages_per_ts 1 --checksum_freq 1 "
                "--refine_freq 1 --num_objects 1 "
                "--object 2 1 0.5 0.5 0.5 0.02 0.03 -0.01 0.25 0.25 0.25 0.002 0 0", "cells"),
    # --code 1: whole ghost planes with the 7-point stencil, two levels, a
    # moving sphere (prolongation/restriction of the ghost ring).
    "c1_t7": ("--init_x 2 --init_y 1 --init_z 1 --nx 4 --ny 4 --nz 4 --num_vars 2 "
              "--code 1 --num_refine 2 --num_tsteps 3 --stages_per_ts 3 --checksum_freq 1 "
              "--refine_freq 1 " + SPHERE0, "cells"),
    # --code 2, 27-point, non-cubic blocks, permuted directions, variable
    # groups, two levels of refinement, bouncing solid sphere.
    "c2_t27": ("--init_x 2 --init_y 2 --init_z 1 --nx 6 --ny 4 --nz 8 --num_vars 3 "
               "--comm_vars 2 --stencil 27 --permute --code 2 --error_tol -10 --num_refine 2 "
               "--max_blocks 3000 --num_tsteps 4 --stages_per_ts 3 --checksum_freq 1 "
               "--refine_freq 1 --num_objects 1 "
               "--object 3 1 0.2 0.6 0.4 0.1 -0.08 0.06 0.15 0.2 0.1 0.01 0 0", "cells"),
    # --stencil 0: the variable-work stencils, 13 variables (mat = 3: variable
    # 1, every class, and variable 12 on the 7-point average), 14 stages
    # (every calculation stage % 6 twice), one level of refinement.
    "s0_a": ("--init_x 2 --init_y 1 --init_z 1 --nx 4 --ny 4 --nz 4 --num_vars 13 "
             "--stencil 0 --num_refine 1 --num_tsteps 2 --stages_per_ts 7 --checksum_freq 1 "
             "--refine_freq 1 " + SPHERE0, "cells"),
    # --stencil 0 with non-cubic blocks, variable groups, permuted directions,
    # --code 1, one level of refinement, a bouncing solid sphere (mat = 2).
    "s0_b": ("--init_x 1 --init_y 2 --init_z 1 --nx 6 --ny 4 --nz 8 --num_vars 9 "
             "--comm_vars 4 --stencil 0 --permute --code 1 --num_refine 1 --num_tsteps 3 "
             "--stages_per_ts 5 --checksum_freq 2 --refine_freq 1 --num_objects 1 "
             "--object 3 1 0.2 0.6 0.4 0.1 -0.08 0.06 0.15 0.2 0.1 0.01 0 0", "cells"),
    # Differentiation-form mesh (static sphere, 40 stages, 4 vars).
    "cfg4": ("--init_x 2 --init_y 2 --init_z 2 --nx 8 --ny 8 --nz 8 --num_vars 4 "
             "--stencil 7 --num_refine 2 --num_tsteps 2 --stages_per_ts 20 "
             "--num_objects 1 --object 2 0 0.3 0.3 0.3 0 0 0 0.2 0.2 0.2 0 0 0", "bsums"),
    # The reference's CPU-runnable configuration.
    "cfg0": ("--init_x 2 --init_y 2 --init_z 2 --nx 10 --ny 10 --nz 10 --num_vars 40 "
             "--stencil 7 --num_tsteps 20 --stages_per_ts 20 --checksum_freq 5 "
             "--num_refine 2 --refine_freq 5 --num_objects 1 "
             "--object 2 0 0.3 0.3 0.3 0.01 0.007 0.005 0.15 0.15 0.15 0 0 0", "bsums"),
    # Failure paths.
    "err_blocks": ("--init_x 2 --init_y 2 --init_z 2 --nx 4 --ny 4 --nz 4 --num_vars 1 "
                   "--num_refine 2 --max_blocks 20 --num_tsteps 2 " + SPHERE0, None),
    "err_tol": ("--init_x 2 --init_y 2 --init_z 2 --nx 4 --ny 4 --nz 4 --num_vars 2 "
                "--stencil 27 --num_refine 2 --num_tsteps 3 --stages_per_ts 5 "
                "--refine_freq 1 " + SPHERE0, "cells"),
    "err_flag": ("--nx 4 --bogus 3", None),
    # The bandwidth-scale workload (27-point, 4 levels, two objects).  The
    # serial reference needs ~20 minutes for it, so it is only generated
    # when named explicitly; per-block interior and ghost sums of all 40
    # variables and the seeded 16,384-cell sample (ghosts included).
    "cfg1": ("--init_x 4 --init_y 4 --init_z 4 --nx 16 --ny 16 --nz 16 --num_vars 40 "
             "--stencil 27 --num_tsteps 10 --stages_per_ts 20 --checksum_freq 5 --num_refine 4 "
             "--refine_freq 4 --max_blocks 8000 --error_tol -10 --num_objects 2 "
             "--object 2 0 0.15 0.15 0.15 0.02 0.02 0.02 0.1 0.1 0.1 0 0 0 "
             "--object 0 0 0.5 0.5 0.5 0.02 0.02 0.02 0.05 0.05 0.05 0 0 0", "bsums_sample"),
    "err_odd": ("--nx 5 --num_tsteps 1", None),
}

# Object-geometry sweep: every object type of ref/move.c:143-623 on a small
# mesh; block counts and checksums pin check_block's decisions.
for _t in [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 20, 21, 22, 23, 24, 25]:
    CASES["obj%02d" % _t] = (
        "--init_x 2 --init_y 2 --init_z 1 --nx 2 --ny 2 --nz 2 --num_vars 1 "
        "--num_refine 3 --max_blocks 3000 --num_tsteps 6 --stages_per_ts 1 "
        "--checksum_freq 1 --refine_freq 1 --num_objects 2 "
        "--object %d 1 0.45 0.55 0.5 0.04 -0.03 0.02 0.22 0.17 0.3 -0.005 0.004 0 "
        "--object %d 2 0.15 0.8 0.3 0.03 0.05 0.04 0.1 0.12 0.08 0.003 0 0.002"
        % (_t, (_t + 2) % 26 if (_t + 2) % 26 in (0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12,
                                                  13, 14, 15, 20, 21, 22, 23, 24, 25) else 3),
        None)

SUMMARY_RE = re.compile(r"^Summary: ranks (\d+) ts (\d+) time \S+ calc \S+ comm \S+ red \S+ "
                        r"refine \S+ blocks/ts (\S+) max_blocks (\d+)")
COUNT_RE = re.compile(r"^Number of blocks at level (\d+) at timestep (\d+) is (\d+)")


def read_dump(path):
    with open(path, "rb") as f:
        data = f.read()
    magic, nx, ny, nz, nv, nact, nref, _ = struct.unpack_from("<8i", data, 0)
    off = 32
    blocks = []
    for _ in range(nact):
        number, level, slot = struct.unpack_from("<qii", data, off)
        off += 16
        if magic == 0x414d5231:
            cnt = nv * (nx + 2) * (ny + 2) * (nz + 2)
            arr = np.frombuffer(data, dtype="<f8", count=cnt, offset=off).reshape(
                nv, nx + 2, ny + 2, nz + 2)
        else:
            cnt = nv * 2
            arr = np.frombuffer(data, dtype="<f8", count=cnt, offset=off).reshape(nv, 2)
        off += 8 * cnt
        blocks.append((number, level, slot, arr.copy()))
    return dict(nx=nx, ny=ny, nz=nz, num_vars=nv, num_refine=nref, blocks=blocks)


def run_case(name, args, kind, outdir=HERE):
    with tempfile.TemporaryDirectory() as td:
        env = dict(os.environ)
        env["AMR_DUMP_SUMS"] = os.path.join(td, "sums.txt")
        if kind == "cells":
            env["AMR_DUMP_CELLS"] = os.path.join(td, "cells.bin")
        elif kind == "bsums":
            env["AMR_DUMP_CELLS"] = os.path.join(td, "cells.bin")
            env["AMR_DUMP_BSUMS"] = os.path.join(td, "bsums.bin")
        elif kind == "bsums_only":
            env["AMR_DUMP_BSUMS"] = os.path.join(td, "bsums.bin")
        elif kind == "bsums_sample":
            env["AMR_DUMP_BSUMS"] = os.path.join(td, "bsums.bin")
            env["AMR_DUMP_SAMPLE"] = os.path.join(td, "sample.bin")
        p = subprocess.run([REF_DUMP] + args.split(), cwd=td, env=env,
                           stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                           timeout=3600)
        out = p.stdout.decode(errors="replace")
        sums = []
        if os.path.exists(env["AMR_DUMP_SUMS"]):
            with open(env["AMR_DUMP_SUMS"]) as f:
                sums = [float(x) for x in f.read().split()]
        rec = {"args": args, "rc": p.returncode, "sums": sums,
               "counts": [], "messages": [], "summary": None}
        for line in out.splitlines():
            m = COUNT_RE.match(line)
            if m:
                rec["counts"].append([int(m.group(2)), int(m.group(1)), int(m.group(3))])
            elif ("difference too large" in line or "ERROR" in line
                  or ("Error" in line and "tolorance" not in line)
                  or "must be" in line or "Unknown input" in line or "Out of free" in line):
                rec["messages"].append(line.strip())
            m = SUMMARY_RE.match(line)
            if m:
                rec["summary"] = {"ranks": int(m.group(1)), "ts": int(m.group(2)),
                                  "blocks_per_ts": float(m.group(3)),
                                  "max_blocks": int(m.group(4))}
        if kind == "cells" and os.path.exists(env["AMR_DUMP_CELLS"]):
            d = read_dump(env["AMR_DUMP_CELLS"])
            np.savez_compressed(
                os.path.join(outdir, name + ".cells.npz"),
                number=np.array([b[0] for b in d["blocks"]], dtype=np.int64),
                level=np.array([b[1] for b in d["blocks"]], dtype=np.int32),
                slot=np.array([b[2] for b in d["blocks"]], dtype=np.int32),
                cells=np.stack([b[3] for b in d["blocks"]]))
        if kind == "bsums" and os.path.exists(env["AMR_DUMP_BSUMS"]):
            d = read_dump(env["AMR_DUMP_BSUMS"])
            c = read_dump(env["AMR_DUMP_CELLS"])
            rng = np.random.default_rng(12345)
            nb = len(c["blocks"])
            g = (c["nx"] + 2, c["ny"] + 2, c["nz"] + 2)
            ns = 4096
            bi = rng.integers(0, nb, ns)
            vi = rng.integers(0, c["num_vars"], ns)
            ii = rng.integers(0, g[0], ns)
            jj = rng.integers(0, g[1], ns)
            kk = rng.integers(0, g[2], ns)
            vals = np.array([c["blocks"][b][3][v, i, j, k]
                             for b, v, i, j, k in zip(bi, vi, ii, jj, kk)])
            np.savez_compressed(
                os.path.join(outdir, name + ".bsums.npz"),
                number=np.array([b[0] for b in d["blocks"]], dtype=np.int64),
                level=np.array([b[1] for b in d["blocks"]], dtype=np.int32),
                slot=np.array([b[2] for b in d["blocks"]], dtype=np.int32),
                sums=np.stack([b[3] for b in d["blocks"]]),
                sample_block=bi.astype(np.int32), sample_var=vi.astype(np.int32),
                sample_idx=np.stack([ii, jj, kk], axis=1).astype(np.int32),
                sample_val=vals)
        if kind == "bsums_only" and os.path.exists(env["AMR_DUMP_BSUMS"]):
            save_bsums_subset(name, read_dump(env["AMR_DUMP_BSUMS"]), outdir)
        if kind == "bsums_sample" and os.path.exists(env["AMR_DUMP_BSUMS"]):
            save_bsums_sample(name, read_dump(env["AMR_DUMP_BSUMS"]),
                              read_sample(env["AMR_DUMP_SAMPLE"]), outdir)
        with open(os.path.join(outdir, name + ".json"), "w") as f:
            json.dump(rec, f, indent=0)
        return rec


BSUM_VARS = [0, 19, 39]


def save_bsums_subset(name, d, outdir=HERE):
    np.savez_compressed(
        os.path.join(outdir, name + ".bsums.npz"),
        number=np.array([b[0] for b in d["blocks"]], dtype=np.int64),
        level=np.array([b[1] for b in d["blocks"]], dtype=np.int32),
        slot=np.array([b[2] for b in d["blocks"]], dtype=np.int32),
        vars=np.array(BSUM_VARS, dtype=np.int32),
        sums=np.stack([b[3][BSUM_VARS] for b in d["blocks"]]))


def read_sample(path):
    """AMR_DUMP_SAMPLE: 8 int32 header (count last), then per cell int32 in,
    var, i, j, k, pad and a double."""
    with open(path, "rb") as f:
        data = f.read()
    hdr = struct.unpack_from("<8i", data, 0)
    assert hdr[0] == 0x414d5333
    rec = np.frombuffer(data, dtype=np.dtype([("idx", "<i4", 6), ("val", "<f8")]),
                        count=hdr[7], offset=32)
    return rec


def save_bsums_sample(name, d, smp, outdir=HERE):
    np.savez_compressed(
        os.path.join(outdir, name + ".bsums.npz"),
        number=np.array([b[0] for b in d["blocks"]], dtype=np.int64),
        level=np.array([b[1] for b in d["blocks"]], dtype=np.int32),
        slot=np.array([b[2] for b in d["blocks"]], dtype=np.int32),
        vars=np.arange(d["num_vars"], dtype=np.int32),
        sums=np.stack([b[3] for b in d["blocks"]]),
        sample_in=smp["idx"][:, 0].astype(np.int32),
        sample_var=smp["idx"][:, 1].astype(np.int32),
        sample_idx=smp["idx"][:, 2:5].astype(np.int32),
        sample_val=smp["val"].copy())


def main(names):
    if not os.path.exists(REF_DUMP):
        sys.exit("build the reference first: make -C oracle ref")
    for name in names or [c for c in CASES if CASES[c][1] not in ("bsums_only", "bsums_sample")]:
        args, kind = CASES[name]
        rec = run_case(name, args, kind)
        print("%-12s rc=%d sums=%d counts=%d msgs=%s" % (
            name, rec["rc"], len(rec["sums"]), len(rec["counts"]), rec["messages"][:2]))


if __name__ == "__main__":
    main(sys.argv[1:])
