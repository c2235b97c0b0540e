"""A synthetic recorded run for the reverse-mode tests: uploads, mesh
plans (meshgen.random_plan: every exchange rule, both z-copy paths),
stages with and without --permute, separate exchanges and stencils, a
split and a consolidation with slot reuse -- in the tape's own format
(ops, plans, job lists) so that it can be applied to any amr_dev.h
implementation and restated in torch."""
import numpy as np

import meshgen
import miniamr_amd as M

OP_UPLOAD, OP_SET_MESH, OP_SPLIT, OP_MERGE, OP_COMM, OP_STENCIL, OP_STAGE = range(1, 8)


def build(seed, n=(4, 6, 8), V=3, stencil=27, ghosts=False):
    """ghosts: the plans' jobs move whole ghost planes (--code 1/2)."""
    rng = np.random.default_rng(seed)
    nparents = 2
    n_split = meshgen.split_slots_needed(nparents)
    n_src, n_dst = 5, 5
    slots = max(n_split, n_src + n_dst, 8 * 2 + 1)
    G = (V, n[0] + 2, n[1] + 2, n[2] + 2)
    uploads = list(range(slots))
    u0 = rng.random((slots,) + G)                 # whole blocks, ghosts included
    plan0 = meshgen.random_plan(rng, n_src, n_dst, ghosts=ghosts)
    plan1 = meshgen.random_plan(rng, n_src, n_dst, ghosts=ghosts)
    plans = {}
    for p, pl in enumerate((plan0, plan1)):
        jobs = np.concatenate(pl)
        plans[p] = (np.array(list(range(n_src, n_src + n_dst)), dtype=np.int32), jobs,
                    [len(x) for x in pl])
    split = meshgen.split_jobs_with_reuse(rng, nparents)
    merge = meshgen.merge_jobs_with_reuse(rng, 2)
    ops, jobs = [], {}
    for s in uploads:
        ops.append([OP_UPLOAD, s, 0, 0, 0, 0, 0, -1])
    ops.append([OP_SET_MESH, 0, 0, 0, 0, 0, 0, 0])
    ops.append([OP_STAGE, 0, V, 0, 0, 0, 0, 0])
    ops.append([OP_STAGE, 0, V, 1, 1, 0, 0, 0])
    ops.append([OP_STAGE, 1, V - 1, 2, 1, 0, 0, 0])
    ops.append([OP_COMM, 0, 2, 4, 1, 0, 0, 0])
    ops.append([OP_STENCIL, 1, 0, 0, 0, 0, 0, 0])
    jobs[len(ops)] = split
    ops.append([OP_SPLIT, 0, len(split), 0, 0, 0, 0, 0])
    ops.append([OP_SET_MESH, 1, 0, 0, 0, 0, 0, 1])
    ops.append([OP_STAGE, 0, V, 5, 1, 0, 0, 1])
    jobs[len(ops)] = merge
    ops.append([OP_MERGE, 0, len(merge), 0, 0, 0, 0, 1])
    ops.append([OP_STAGE, 0, V, 3, 0, 0, 0, 1])
    cfg = dict(nx=n[0], ny=n[1], nz=n[2], num_vars=V, max_blocks=slots, stencil=stencil)
    return {"ops": ops, "plans": plans, "jobs": jobs, "uploads": uploads, "u0": u0,
            "cfg": cfg, "dir_plans": {0: plan0, 1: plan1}}


def apply(dev, t):
    """Run the synthetic sequence on a Device."""
    for i, op in enumerate(t["ops"]):
        k = op[0]
        if k == OP_UPLOAD:
            dev.upload(op[1], t["u0"][t["uploads"].index(op[1])])
        elif k == OP_SET_MESH:
            dev.set_mesh(t["plans"][op[1]][0], t["dir_plans"][op[1]])
        elif k == OP_STAGE:
            dev.stage(op[1], op[2], stage=op[3], permute=bool(op[4]))
        elif k == OP_COMM:
            dev.comm(op[1], op[2], stage=op[3], permute=bool(op[4]))
        elif k == OP_STENCIL:
            dev.stencil(op[1])
        elif k == OP_SPLIT:
            dev.split(t["jobs"][i])
        elif k == OP_MERGE:
            dev.consolidate(t["jobs"][i])
