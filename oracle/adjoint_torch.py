"""Torch restatement of a recorded run, for FP64 autograd (TEST
INFRASTRUCTURE: used only by tests/ as the checker of the adjoint kernels;
the product never imports it).

Given the tape of a run (include/amr_adjoint.h: uploads, split/merge job
lists, mesh plans, stages), this rebuilds the same forward map with plain
torch tensor operations on the reference's canonical cell arrays --
state[var][slot][i][j][k] with one ghost layer, ghosts that nothing writes
keeping their values -- so that torch.autograd can differentiate it.  Each
step restates the reference's arithmetic:
  split         child interior = parent octant / 8        ref/block.c:161-173
  consolidate   merged octant = ordered sum of 8 cells    ref/block.c:411-431
  exchange      per direction, every ghost plane fill of the plan, read
                before any write:
                  same level: copy                        ref/comm.c:1473-1534
                  prolong: fine ghost = coarse / 4        ref/comm.c:1618-1625
                  restrict: ((f00 + f01) + f10) + f11     ref/comm.c:1626-1630
                  boundary: copy of the adjacent plane    ref/comm.c:1911-1965
                with the 27-point stencil's extended in-plane ranges
                (ref/comm.c:1496-1533, 1941-1964) and the direction order
                of --permute (ref/comm.c:45-55); --code 1/2 fills (kind |
                GHOSTS) cover whole ghost planes: prolongation over the ring
                too, restriction of a quarter plus its share of the ring
                (2 fine ring cells per ring edge cell, 1 at the corner)
                                                          ref/comm.c:1535-1579, 1689-1907;
  stencil       7-point (((((W+S)+D)+C)+U)+N)+E, / 7      ref/stencil.c:82-101
                27-point ((sb + sm) + sf) / 27            ref/stencil.c:102-144
Only what the tape holds is used; values come from the caller's tensor.
Checked by tests/test_gpu_adjoint.py (forward value against the GPU run,
gradient against the adjoint kernels and finite differences).
"""
import ctypes

import numpy as np
import torch

OP_UPLOAD, OP_SET_MESH, OP_SPLIT, OP_MERGE, OP_COMM, OP_STENCIL, OP_STAGE = range(1, 8)
SAME, PROLONG, RESTRICT, BC = 0, 1, 2, 3
GHOSTS = 4   # or'ed into a kind: whole ghost planes (--code 1/2)
PERM = [(0, 1, 2), (1, 2, 0), (2, 0, 1), (0, 2, 1), (1, 0, 2), (2, 1, 0)]


def read_tape(lib, dev, face_job, split_job, merge_job):
    """Ops, plans and job lists of the tape of `dev` (ctypes handles)."""
    n = ctypes.c_int()
    lib.amr_dev_tape_length(dev, ctypes.byref(n))
    ops, plans, jobs = [], {}, {}
    for i in range(n.value):
        op = (ctypes.c_int32 * 8)()
        lib.amr_dev_tape_op(dev, i, op)
        op = list(op)
        ops.append(op)
        if op[0] == OP_SET_MESH and op[1] not in plans:
            na, cnt = ctypes.c_int32(), (ctypes.c_int32 * 3)()
            lib.amr_dev_tape_plan(dev, op[1], ctypes.byref(na), None, cnt, None)
            sorted_ = np.zeros(max(1, na.value), dtype=np.int32)
            fj = np.zeros(max(1, sum(cnt)), dtype=face_job)
            lib.amr_dev_tape_plan(dev, op[1], None,
                                  sorted_.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), cnt,
                                  fj.ctypes.data_as(ctypes.c_void_p))
            plans[op[1]] = (sorted_[:na.value].copy(), fj[:sum(cnt)].copy(), list(cnt))
        elif op[0] in (OP_SPLIT, OP_MERGE):
            arr = np.zeros(op[2], dtype=split_job if op[0] == OP_SPLIT else merge_job)
            f = lib.amr_dev_tape_split_jobs if op[0] == OP_SPLIT else lib.amr_dev_tape_merge_jobs
            f(dev, i, arr.ctypes.data_as(ctypes.c_void_p))
            jobs[i] = arr
    return ops, plans, jobs


def _canon(d, p, a, b):
    return (p, a, b) if d == 0 else (a, p, b) if d == 1 else (a, b, p)


class Restatement:
    def __init__(self, ops, plans, jobs, n, stencil, num_vars, num_slots):
        self.ops, self.plans, self.jobs = ops, plans, jobs
        self.nx, self.ny, self.nz = n
        self.G = (self.nx + 2, self.ny + 2, self.nz + 2)
        self.stencil, self.V, self.S = stencil, num_vars, num_slots
        self._fills = {}

    def flat(self, slot, i, j, k):
        gx, gy, gz = self.G
        return ((slot * gx + i) * gy + j) * gz + k

    def fills(self, plan, d):
        """(kind groups) index arrays of direction d of a plan: for copies
        (dst, src), for prolong (dst, src) with /4, for restrict
        (dst, s0, s1, s2, s3)."""
        key = (plan, d)
        if key in self._fills:
            return self._fills[key]
        _, fj, cnt = self.plans[plan]
        off = [0, cnt[0], cnt[0] + cnt[1]][d]
        n = [self.nx, self.ny, self.nz][d]
        na = self.ny if d == 0 else self.nx
        nb = self.ny if d == 2 else self.nz
        cp_d, cp_s, pr_d, pr_s, rs_d, rs_s = [], [], [], [], [], []
        r1_d, r1_s, r2_d, r2_s = [], [], [], []   # ring cells of whole-plane restrictions
        for jb in fj[off:off + cnt[d]]:
            kind, face, dst, src, iq, jq = (int(x) for x in jb)
            whole = kind & GHOSTS
            kind &= 3
            hi = face & 1
            ghost = n + 1 if hi else 0
            plane = 1 if hi else n
            if kind in (SAME, BC):
                ea = eb = 0
                if whole:
                    ea = eb = 1
                elif self.stencil == 27:
                    if kind == BC:
                        ea = eb = 1
                    elif d == 1:
                        ea = 1
                    elif d == 2:
                        ea = eb = 1
                sp = (n if hi else 1) if kind == BC else plane
                s_slot = dst if kind == BC else src
                for a in range(1 - ea, na + 1 + ea):
                    for b in range(1 - eb, nb + 1 + eb):
                        cp_d.append(self.flat(dst, *_canon(d, ghost, a, b)))
                        cp_s.append(self.flat(s_slot, *_canon(d, sp, a, b)))
            elif kind == PROLONG:
                ao, bo = jq * (na // 2), iq * (nb // 2)
                w = 1 if whole else 0
                for a in range(1 - w, na + 1 + w):
                    for b in range(1 - w, nb + 1 + w):
                        ca, cb = (a + 1) // 2 + ao, (b + 1) // 2 + bo
                        pr_d.append(self.flat(dst, *_canon(d, ghost, a, b)))
                        pr_s.append(self.flat(src, *_canon(d, plane, ca, cb)))
            elif whole:
                ha, hb = na // 2, nb // 2
                ao, bo = jq * ha, iq * hb
                arange = range(ha + 1, na + 2) if jq else range(0, ha + 1)
                brange = range(hb + 1, nb + 2) if iq else range(0, hb + 1)
                for a in arange:
                    fa = [a] if a in (0, na + 1) else [2 * (a - ao) - 1, 2 * (a - ao)]
                    for b in brange:
                        fb = [b] if b in (0, nb + 1) else [2 * (b - bo) - 1, 2 * (b - bo)]
                        terms = [self.flat(src, *_canon(d, plane, x, y)) for x in fa for y in fb]
                        dcell = self.flat(dst, *_canon(d, ghost, a, b))
                        if len(terms) == 4:
                            rs_d.append(dcell)
                            rs_s.append(terms)
                        elif len(terms) == 2:
                            r2_d.append(dcell)
                            r2_s.append(terms)
                        else:
                            r1_d.append(dcell)
                            r1_s.append(terms[0])
            else:
                ha, hb = na // 2, nb // 2
                ao, bo = jq * ha, iq * hb
                for a in range(1, ha + 1):
                    for b in range(1, hb + 1):
                        rs_d.append(self.flat(dst, *_canon(d, ghost, a + ao, b + bo)))
                        rs_s.append([self.flat(src, *_canon(d, plane, 2 * a - 1 + (r >> 1),
                                                             2 * b - 1 + (r & 1)))
                                     for r in range(4)])
        t = lambda x: torch.tensor(x, dtype=torch.long)
        res = (t(cp_d), t(cp_s), t(pr_d), t(pr_s), t(rs_d),
               t(rs_s).reshape(-1, 4) if rs_s else torch.zeros((0, 4), dtype=torch.long),
               t(r1_d), t(r1_s), t(r2_d),
               t(r2_s).reshape(-1, 2) if r2_s else torch.zeros((0, 2), dtype=torch.long))
        self._fills[key] = res
        return res

    def exchange(self, u, plan, d, v0, v1):
        cp_d, cp_s, pr_d, pr_s, rs_d, rs_s, r1_d, r1_s, r2_d, r2_s = self.fills(plan, d)
        f = u.reshape(self.V, -1)
        rows = f[v0:v1]
        dst = torch.cat([cp_d, pr_d, rs_d, r1_d, r2_d])
        vals = torch.cat([rows[:, cp_s], rows[:, pr_s] / 4.0,
                          ((rows[:, rs_s[:, 0]] + rows[:, rs_s[:, 1]]) + rows[:, rs_s[:, 2]])
                          + rows[:, rs_s[:, 3]],
                          rows[:, r1_s], rows[:, r2_s[:, 0]] + rows[:, r2_s[:, 1]]], dim=1)
        rows = rows.index_copy(1, dst, vals)
        f = torch.cat([f[:v0], rows, f[v1:]])
        return f.reshape(u.shape)

    def stencil_op(self, u, plan, v0, v1):
        slots = torch.tensor(self.plans[plan][0], dtype=torch.long)
        a = u[v0:v1][:, slots]
        if self.stencil == 7:
            c = a[..., 1:-1, 1:-1, 1:-1]
            w = a[..., :-2, 1:-1, 1:-1] + a[..., 1:-1, :-2, 1:-1]
            w = w + a[..., 1:-1, 1:-1, :-2]
            w = w + c
            w = w + a[..., 1:-1, 1:-1, 2:]
            w = w + a[..., 1:-1, 2:, 1:-1]
            w = w + a[..., 2:, 1:-1, 1:-1]
            w = w / 7.0
        else:
            def s9(p):   # 9-term sum of plane slice p, (j, k) order
                t = None
                for dj in range(3):
                    for dk in range(3):
                        x = p[..., dj:dj + self.ny, dk:dk + self.nz]
                        t = x if t is None else t + x
                return t
            sb = s9(a[..., :-2, :, :])
            sm = s9(a[..., 1:-1, :, :])
            sf = s9(a[..., 2:, :, :])
            w = ((sb + sm) + sf) / 27.0
        out = u.clone()
        inner = out[v0:v1][:, slots].clone()
        inner[..., 1:-1, 1:-1, 1:-1] = w
        out[v0:v1, slots] = inner
        return out

    def split(self, u, jobs):
        hx, hy, hz = self.nx // 2, self.ny // 2, self.nz // 2
        snap = u.clone()
        out = u.clone()
        for jb in jobs:
            p, c, o = int(jb["parent"]), int(jb["child"]), int(jb["octant"])
            i1, j1, k1 = (o & 1) * hx, ((o >> 1) & 1) * hy, (o >> 2) * hz
            par = snap[:, p, 1 + i1:1 + i1 + hx, 1 + j1:1 + j1 + hy, 1 + k1:1 + k1 + hz]
            ch = par.repeat_interleave(2, 1).repeat_interleave(2, 2).repeat_interleave(2, 3) / 8.0
            out[:, c, 1:-1, 1:-1, 1:-1] = ch
        return out

    def merge(self, u, jobs):
        hx, hy, hz = self.nx // 2, self.ny // 2, self.nz // 2
        snap = u.clone()
        out = u.clone()
        for jb in jobs:
            dst, c, o = int(jb["dst"]), int(jb["child"]), int(jb["octant"])
            i1, j1, k1 = (o & 1) * hx, ((o >> 1) & 1) * hy, (o >> 2) * hz
            ch = snap[:, c, 1:-1, 1:-1, 1:-1]
            s = ch[:, 0::2, 0::2, 0::2]
            s = s + ch[:, 1::2, 0::2, 0::2]
            s = s + ch[:, 0::2, 1::2, 0::2]
            s = s + ch[:, 1::2, 1::2, 0::2]
            s = s + ch[:, 0::2, 0::2, 1::2]
            s = s + ch[:, 1::2, 0::2, 1::2]
            s = s + ch[:, 0::2, 1::2, 1::2]
            s = s + ch[:, 1::2, 1::2, 1::2]
            out[:, dst, 1 + i1:1 + i1 + hx, 1 + j1:1 + j1 + hy, 1 + k1:1 + k1 + hz] = s
        return out

    def run(self, u0, upload_slots, keep=()):
        """u0: (n_uploads, V, nx, ny, nz) float64 interiors (ghosts zero) or
        (n_uploads, V, nx+2, ny+2, nz+2) whole blocks, uploaded in that
        order.  Returns the final state (V, slots, gx, gy, gz) and the
        states before the operations listed in `keep` (operation indices;
        len(ops) for the final state)."""
        gx, gy, gz = self.G
        u = torch.zeros((self.V, self.S, gx, gy, gz), dtype=torch.float64)
        pad = u0 if u0.shape[-1] == gz else torch.nn.functional.pad(u0, (1, 1, 1, 1, 1, 1))
        idx = torch.tensor(upload_slots, dtype=torch.long)
        u = u.index_copy(1, idx, pad.transpose(0, 1))
        kept = {}
        for i, op in enumerate(self.ops):
            if i in keep:
                kept[i] = u
            kind, plan = op[0], op[7]
            if kind == OP_SPLIT:
                u = self.split(u, self.jobs[i])
            elif kind == OP_MERGE:
                u = self.merge(u, self.jobs[i])
            elif kind in (OP_COMM, OP_STAGE):
                v0, v1 = op[1], op[1] + op[2]
                for o in range(3):
                    d = PERM[op[3] % 6][o] if op[4] else o
                    u = self.exchange(u, plan, d, v0, v1)
                if kind == OP_STAGE:
                    u = self.stencil_op(u, plan, v0, v1)
            elif kind == OP_STENCIL:
                u = self.stencil_op(u, plan, op[1], op[1] + 1)
        if len(self.ops) in keep:
            kept[len(self.ops)] = u
        return u, kept
