"""Reverse mode of the hot path as torch.autograd Functions (ctypes).

A :class:`TapedRun` runs one reference-style command line forward on the
GPU with the device library's tape on (``--tape``), which records every
operation that changes cell values: the initial uploads, the split and
consolidate job lists, the mesh plans and the stages.  The tape then
defines a differentiable map from the initial block values to the cells,
built from three autograd Functions whose forward and backward are calls
into the HIP library (include/amr_adjoint.h):

* ``Init``    initial values -> whole cell pool (the uploads);
              backward reads the adjoint of each uploaded block;
* ``Segment`` pool -> pool over a range of recorded operations (replayed
              forward); backward runs the adjoint kernels over the same
              range, last operation first;
* ``Cells``   pool -> the interior cells of the final active blocks;
              backward places the incoming gradient on those cells.

:meth:`TapedRun.forward` chains them with one ``Segment`` per ``k`` stages
(the first also holds the initial refinement), so the pool state every k
stages is a tensor of its own: a checkpoint that any loss may use and from
which the reverse pass of the following segments restarts.  Every recorded
operation is linear, so the backward passes need no forward values.

Pool tensors (the checkpoints) are CPU tensors, one element per pool
element in the storage precision; the library copies them to and from its
device pool, and torch only differentiates on the host.  torch must be
imported before the library is loaded (this module checks): its bundled
HIP runtime then is the process's only one.
"""
import ctypes
import sys

from . import _LOADED_BEFORE_TORCH, AmrError, Run, load_library

if "torch" not in sys.modules and _LOADED_BEFORE_TORCH:
    raise AmrError("import torch (or miniamr_amd.autodiff) before loading %s: torch brings a "
                   "HIP runtime of its own, and loading it after the library would put two "
                   "runtimes in one process" % _LOADED_BEFORE_TORCH[0])

import numpy as np  # noqa: E402
import torch  # noqa: E402

OP_UPLOAD, OP_SET_MESH, OP_SPLIT, OP_MERGE, OP_COMM, OP_STENCIL, OP_STAGE = range(1, 8)
_DTYPE = {0: torch.float64, 1: torch.float32, 2: torch.bfloat16}


class TapedRun:
    """One forward run with a tape, and the autograd Functions over it."""

    def __init__(self, args, k=8, lib=None):
        self.lib = lib or load_library()
        self.run = Run(args + " --tape")
        while self.run.step() == 0:
            pass
        self.dev = ctypes.c_void_p(self.run.device())
        self._check(self.lib.amr_dev_tape(self.dev, 0), "tape off")
        self.lib.amr_dev_set_timing(self.dev, 0)
        ms = (ctypes.c_double * 4)()
        self.lib.amr_dev_read_timing(self.dev, ms)
        st = self.run.stats()
        self.nx, self.ny, self.nz, self.num_vars = st.nx, st.ny, st.nz, st.num_vars
        self.precision = st.precision
        self.dtype = _DTYPE[st.precision]
        n = ctypes.c_int()
        self.lib.amr_dev_tape_length(self.dev, ctypes.byref(n))
        self.ops = []
        for i in range(n.value):
            op = (ctypes.c_int32 * 8)()
            self._check(self.lib.amr_dev_tape_op(self.dev, i, op), "tape_op")
            self.ops.append(list(op))
        ups = [i for i, op in enumerate(self.ops) if op[0] == OP_UPLOAD]
        if ups != list(range(len(ups))):
            raise AmrError("uploads are expected only at the start of the tape")
        self.upload_slots = [self.ops[i][1] for i in ups]
        self.first = len(ups)
        nb = ctypes.c_size_t()
        self.lib.amr_dev_state_bytes(self.dev, ctypes.byref(nb))
        self.numel = nb.value // torch.tensor([], dtype=self.dtype).element_size()
        self.blocks = self.run.blocks()          # final active blocks
        h = ctypes.c_void_p()
        rc = self.lib.amr_adj_create(self.dev, ctypes.byref(h))
        self.adj = h
        if rc:
            raise AmrError("amr_adj_create failed (%d): %s" % (
                rc, self.lib.amr_adj_last_error(h).decode()))
        self.segments = self._segments(k)

    # -- helpers ---------------------------------------------------------
    def _check(self, rc, what):
        if rc:
            raise AmrError("%s failed (%d): %s" % (
                what, rc, self.lib.amr_dev_last_error(self.dev).decode()))

    def _acheck(self, rc, what):
        if rc:
            raise AmrError("%s failed (%d): %s" % (
                what, rc, self.lib.amr_adj_last_error(self.adj).decode()))

    def _segments(self, k):
        """Operation ranges holding k stages each (stage numbers of the
        recorded STAGE/COMM operations); the first starts after the uploads."""
        bounds = [self.first]
        for i, op in enumerate(self.ops):
            if op[0] in (OP_STAGE, OP_COMM) and op[3] > 0 and op[3] % k == 0:
                prev = [o for o in self.ops[:i] if o[0] in (OP_STAGE, OP_COMM)]
                if prev and prev[-1][3] != op[3] and i > bounds[-1]:
                    bounds.append(i)
        bounds.append(len(self.ops))
        return list(zip(bounds[:-1], bounds[1:]))

    def n_stages(self):
        return len({op[3] for op in self.ops if op[0] in (OP_STAGE, OP_COMM)})

    def new_state(self):
        return torch.empty(self.numel, dtype=self.dtype)

    @property
    def block_shape(self):
        return (self.num_vars, self.nx, self.ny, self.nz)

    def canonical(self, interior):
        g = np.zeros((self.num_vars, self.nx + 2, self.ny + 2, self.nz + 2))
        g[:, 1:-1, 1:-1, 1:-1] = interior
        return np.ascontiguousarray(g)

    def cell_centres(self, blk):
        """Cell centres of an active block (slot, level, lo, hi) as three
        broadcastable arrays (x, y, z)."""
        _, _, lo, hi = blk
        n = (self.nx, self.ny, self.nz)
        c = [lo[d] + (np.arange(n[d]) + 0.5) * (hi[d] - lo[d]) / n[d] for d in range(3)]
        return c[0][:, None, None], c[1][None, :, None], c[2][None, None, :]

    # -- the differentiable map -----------------------------------------
    def forward(self, u0):
        """u0: (n_initial_blocks, num_vars, nx, ny, nz) float64 tensor of
        initial interior values.  Returns (cells, checkpoints): the final
        active blocks' interior cells (n_active, num_vars, nx, ny, nz,
        float64, CPU) and the pool states after each segment."""
        state = Init.apply(u0, self)
        checkpoints = []
        for op0, op1 in self.segments:
            state = Segment.apply(state, self, op0, op1)
            checkpoints.append(state)
        return Cells.apply(state, self), checkpoints

    def close(self):
        if getattr(self, "adj", None) is not None and self.adj.value:
            self.lib.amr_adj_destroy(self.adj)
            self.adj = ctypes.c_void_p()
        if getattr(self, "run", None) is not None:
            self.run.finish()
            self.run.close()
            self.run = None


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


class Init(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u0, tr):
        ctx.tr = tr
        ctx.meta = (u0.dtype, u0.device)
        zero = torch.zeros(tr.numel, dtype=tr.dtype)
        tr._check(tr.lib.amr_dev_load_state(tr.dev, _ptr(zero)), "load_state")
        vals = u0.detach().to("cpu", torch.float64).numpy()
        for b, slot in enumerate(tr.upload_slots):
            cells = tr.canonical(vals[b])
            tr._check(tr.lib.amr_dev_upload_block(
                tr.dev, slot, cells.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), "upload")
        out = tr.new_state()
        tr._check(tr.lib.amr_dev_save_state(tr.dev, _ptr(out)), "save_state")
        return out

    @staticmethod
    def backward(ctx, g):
        tr = ctx.tr
        g = g.contiguous().to("cpu", tr.dtype)
        tr._acheck(tr.lib.amr_adj_load(tr.adj, _ptr(g)), "adj_load")
        grads = []
        buf = np.zeros((tr.num_vars, tr.nx + 2, tr.ny + 2, tr.nz + 2))
        for slot in tr.upload_slots:
            tr._acheck(tr.lib.amr_adj_get_block(
                tr.adj, slot, buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), "get_block")
            grads.append(buf[:, 1:-1, 1:-1, 1:-1].copy())
        dtype, device = ctx.meta
        return torch.from_numpy(np.stack(grads)).to(device=device, dtype=dtype), None


class Segment(torch.autograd.Function):
    @staticmethod
    def forward(ctx, state, tr, op0, op1):
        ctx.tr, ctx.ops = tr, (op0, op1)
        state = state.contiguous()
        tr._check(tr.lib.amr_dev_load_state(tr.dev, _ptr(state)), "load_state")
        tr._check(tr.lib.amr_dev_replay(tr.dev, op0, op1), "replay")
        out = tr.new_state()
        tr._check(tr.lib.amr_dev_save_state(tr.dev, _ptr(out)), "save_state")
        return out

    @staticmethod
    def backward(ctx, g):
        tr = ctx.tr
        g = g.contiguous().to("cpu", tr.dtype)
        tr._acheck(tr.lib.amr_adj_load(tr.adj, _ptr(g)), "adj_load")
        tr._acheck(tr.lib.amr_adj_run(tr.adj, *ctx.ops), "adj_run")
        out = tr.new_state()
        tr._acheck(tr.lib.amr_adj_save(tr.adj, _ptr(out)), "adj_save")
        return out, None, None, None


class Cells(torch.autograd.Function):
    @staticmethod
    def forward(ctx, state, tr):
        ctx.tr = tr
        state = state.contiguous()
        tr._check(tr.lib.amr_dev_load_state(tr.dev, _ptr(state)), "load_state")
        out = np.empty((len(tr.blocks),) + tr.block_shape)
        buf = np.zeros((tr.num_vars, tr.nx + 2, tr.ny + 2, tr.nz + 2))
        for b, (slot, _, _, _) in enumerate(tr.blocks):
            tr._check(tr.lib.amr_dev_download_block(
                tr.dev, slot, buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), "download")
            out[b] = buf[:, 1:-1, 1:-1, 1:-1]
        return torch.from_numpy(out)

    @staticmethod
    def backward(ctx, g):
        tr = ctx.tr
        gv = g.detach().to("cpu", torch.float64).numpy()
        tr._acheck(tr.lib.amr_adj_zero(tr.adj), "adj_zero")
        for b, (slot, _, _, _) in enumerate(tr.blocks):
            cells = tr.canonical(gv[b])
            tr._acheck(tr.lib.amr_adj_set_block(
                tr.adj, slot, cells.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), "set_block")
        out = tr.new_state()
        tr._acheck(tr.lib.amr_adj_save(tr.adj, _ptr(out)), "adj_save")
        return out, None
