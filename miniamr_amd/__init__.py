"""miniAMR hot path on MI355X (gfx950): Python view of the C ABI.

The product is two in-tree artefacts built by ``make -C miniamr_amd``:

* ``lib/libminiamr_dev.so`` -- HIP kernels behind the C ABI declared in
  ``include/amr_dev.h`` (stencil, ghost exchange, checksum, split and
  consolidate, all resident in device memory);
* ``bin/miniAMR.x`` -- the drop-in driver: the reference's command line,
  block/object model and refinement decisions in C, calling the library.

This module loads the library with ctypes (for tests, the benchmark and
Python callers) and runs the driver.  There is no CPU fallback: if the
library is missing or no gfx950 device is present, calls raise.
"""
import ctypes
import os
import subprocess

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
LIB_PATH = os.path.join(PKG, "lib", "libminiamr_dev.so")
EXE_PATH = os.path.join(PKG, "bin", "miniAMR.x")
INCLUDE = os.path.join(REPO, "include", "amr_dev.h")

AMR_OK = 0
PREC_F64, PREC_F32, PREC_BF16 = 0, 1, 2
FACE_SAME, FACE_PROLONG, FACE_RESTRICT, FACE_BC = 0, 1, 2, 3
FACE_GHOSTS = 4   # or'ed into a kind: whole ghost planes (--code 1/2)

# Every symbol include/amr_dev.h declares.
ABI_SYMBOLS = [
    "amr_dev_create", "amr_dev_destroy", "amr_dev_last_error", "amr_dev_upload_block",
    "amr_dev_download_block", "amr_dev_set_mesh", "amr_dev_comm", "amr_dev_stencil",
    "amr_dev_check_sum", "amr_dev_split", "amr_dev_consolidate", "amr_dev_reserve_sums",
    "amr_dev_read_sums", "amr_dev_stage", "amr_dev_sync", "amr_dev_set_timing",
    "amr_dev_read_timing", "amr_dev_read_exchange_timing", "amr_dev_memory",
    "amr_dev_stage_fused", "amr_dev_set_materials", "amr_dev_stencil_stage",
    "amr_dev_read_check_counts", "amr_dev_mark_refine", "amr_dev_split_noalloc",
    "amr_dev_face_msg_len", "amr_dev_pack_faces", "amr_dev_unpack_faces",
    "amr_dev_compare_cells", "amr_dev_copy_bandwidth",
]


class Config(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("num_vars", ctypes.c_int), ("max_blocks", ctypes.c_int),
                ("stencil", ctypes.c_int), ("precision", ctypes.c_int),
                ("device", ctypes.c_int)]


FACE_JOB = np.dtype([("kind", "<i4"), ("face", "<i4"), ("dst", "<i4"), ("src", "<i4"),
                     ("iq", "<i4"), ("jq", "<i4")])
SPLIT_JOB = np.dtype([("parent", "<i4"), ("child", "<i4"), ("octant", "<i4"), ("pad", "<i4")])
MERGE_JOB = np.dtype([("dst", "<i4"), ("child", "<i4"), ("octant", "<i4"), ("pad", "<i4")])
FACE_MSG = np.dtype([("slot", "<i4"), ("face_case", "<i4"), ("offset", "<i8")])


class BatchShape(ctypes.Structure):
    """amr_batch_shape of include/amr_batch.h."""
    _fields_ = [("vpw", ctypes.c_int), ("stencil_wgs", ctypes.c_int),
                ("exchange_wgs", ctypes.c_int), ("wg_classes", ctypes.c_int),
                ("exchange_vars", ctypes.c_int), ("exchange_threads", ctypes.c_int)]


class AmrError(RuntimeError):
    pass


# One HIP runtime per process.  torch ships its own libamdhip64 with the
# same soname as ROCm's: imported first, it is the runtime the libraries
# bind to; imported after a library was loaded, torch would load a second
# copy next to ROCm's.  _LOADED_BEFORE_TORCH records that second order so
# that miniamr_amd.autodiff (which needs torch) can refuse it.
_LOADED_BEFORE_TORCH = []


def _cdll(path):
    import sys
    if "torch" not in sys.modules:
        _LOADED_BEFORE_TORCH.append(path)
    return ctypes.CDLL(path)


def load_library(path=LIB_PATH):
    """Load an implementation of amr_dev.h (the HIP library by default)."""
    if not os.path.exists(path):
        raise AmrError("%s is not built: run `make -C miniamr_amd` (or "
                       "__graft_entry__.build())" % path)
    lib = _cdll(path)
    vp, ip, i32p, dp = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int32), \
        ctypes.POINTER(ctypes.c_double)
    sig = {
        "amr_dev_create": (ip, [ctypes.POINTER(Config), ctypes.POINTER(vp)]),
        "amr_dev_destroy": (ip, [vp]),
        "amr_dev_last_error": (ctypes.c_char_p, [vp]),
        "amr_dev_upload_block": (ip, [vp, ip, dp]),
        "amr_dev_download_block": (ip, [vp, ip, dp]),
        "amr_dev_set_mesh": (ip, [vp, i32p, ip, vp, i32p]),
        "amr_dev_comm": (ip, [vp, ip, ip, ip, ip]),
        "amr_dev_stencil": (ip, [vp, ip]),
        "amr_dev_check_sum": (ip, [vp, ip, dp]),
        "amr_dev_split": (ip, [vp, vp, ip]),
        "amr_dev_split_noalloc": (ip, [vp, vp, ip]),
        "amr_dev_face_msg_len": (ip, [vp, ip, ip, ip, ip]),
        "amr_dev_compare_cells": (ip, [vp, vp, ip, ip, dp]),
        "amr_dev_copy_bandwidth": (ip, [ctypes.c_size_t, ip, dp, dp]),
        "amr_dev_pack_faces": (ip, [vp, ip, ip, ip, ip, vp, ip, vp, ctypes.c_size_t]),
        "amr_dev_unpack_faces": (ip, [vp, ip, ip, ip, ip, vp, ip, vp, ctypes.c_size_t]),
        "amr_dev_consolidate": (ip, [vp, vp, ip]),
        "amr_dev_reserve_sums": (ip, [vp, ip]),
        "amr_dev_read_sums": (ip, [vp, ip, ip, dp]),
        "amr_dev_stage": (ip, [vp, ip, ip, ip, ip, ip]),
        "amr_dev_stage_fused": (ip, [vp, ip, ip, ip, ip, ip]),
        "amr_dev_sync": (ip, [vp]),
        "amr_dev_set_timing": (ip, [vp, ip]),
        "amr_dev_read_timing": (ip, [vp, dp]),
        "amr_dev_read_exchange_timing": (ip, [vp, dp]),
        "amr_dev_memory": (ip, [vp, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ip)]),
        "amr_dev_set_materials": (ip, [vp, ctypes.c_double, dp, ip]),
        "amr_dev_stencil_stage": (ip, [vp, ip, ip]),
        "amr_dev_read_check_counts": (ip, [vp, ctypes.POINTER(ctypes.c_int64)]),
        "amr_dev_mark_refine": (ip, [vp, vp, vp, ip, i32p, vp, ip, vp, vp, vp]),
        # include/amr_batch.h, device half
        "amr_batch_create": (ip, [ip, ip, ip, ctypes.POINTER(vp)]),
        "amr_batch_destroy": (ip, [vp]),
        "amr_batch_last_error": (ctypes.c_char_p, [vp]),
        "amr_batch_add": (ip, [vp, ctypes.POINTER(Config), ctypes.POINTER(vp)]),
        "amr_batch_set_active": (ip, [vp, vp, ip]),
        "amr_batch_set_shape": (ip, [vp, ctypes.POINTER(BatchShape)]),
        "amr_batch_get_shape": (ip, [vp, ctypes.POINTER(BatchShape)]),
        "amr_batch_stage": (ip, [vp, ip, ip, ip]),
        "amr_batch_timestep": (ip, [vp, ip, ip, ip, ip]),
        "amr_batch_autotune": (ip, [vp, dp]),
        "amr_batch_sync": (ip, [vp]),
        "amr_batch_set_timing": (ip, [vp, ip]),
        "amr_batch_read_timing": (ip, [vp, dp]),
        # include/amr_adjoint.h
        "amr_dev_tape": (ip, [vp, ip]),
        "amr_dev_tape_length": (ip, [vp, ctypes.POINTER(ctypes.c_int)]),
        "amr_dev_tape_op": (ip, [vp, ip, i32p]),
        "amr_dev_tape_plan": (ip, [vp, ip, i32p, i32p, i32p, vp]),
        "amr_dev_tape_split_jobs": (ip, [vp, ip, vp]),
        "amr_dev_tape_merge_jobs": (ip, [vp, ip, vp]),
        "amr_dev_replay": (ip, [vp, ip, ip]),
        "amr_dev_state_bytes": (ip, [vp, ctypes.POINTER(ctypes.c_size_t)]),
        "amr_dev_save_state": (ip, [vp, vp]),
        "amr_dev_load_state": (ip, [vp, vp]),
        "amr_dev_clear_block": (ip, [vp, ip]),
        "amr_adj_create": (ip, [vp, ctypes.POINTER(vp)]),
        "amr_adj_destroy": (ip, [vp]),
        "amr_adj_last_error": (ctypes.c_char_p, [vp]),
        "amr_adj_zero": (ip, [vp]),
        "amr_adj_load": (ip, [vp, vp]),
        "amr_adj_save": (ip, [vp, vp]),
        "amr_adj_get_block": (ip, [vp, ip, dp]),
        "amr_adj_set_block": (ip, [vp, ip, dp]),
        "amr_adj_run": (ip, [vp, ip, ip]),
        "amr_adj_sync": (ip, [vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name, None)
        if f is None:
            if name.startswith(("amr_adj", "amr_dev_tape_", "amr_dev_replay", "amr_dev_state",
                                "amr_dev_save", "amr_dev_load", "amr_dev_clear")):
                continue   # reverse mode: the HIP library only
            raise AmrError("%s does not export %s" % (path, name))
        f.restype = res
        f.argtypes = args
    return lib


class RunStats(ctypes.Structure):
    """amr_run_stats of include/amr_run.h."""
    _fields_ = [("timesteps_done", ctypes.c_int64), ("stages_done", ctypes.c_int64),
                ("block_stages", ctypes.c_int64), ("cell_var_updates", ctypes.c_double),
                ("active_blocks", ctypes.c_int64), ("max_slots", ctypes.c_int64),
                ("device_ms", ctypes.c_double * 4), ("stage_calls", ctypes.c_int64),
                ("nx", ctypes.c_int), ("ny", ctypes.c_int),
                ("nz", ctypes.c_int), ("num_vars", ctypes.c_int), ("stencil", ctypes.c_int),
                ("precision", ctypes.c_int)]


HOST_LIB_PATH = os.path.join(PKG, "lib", "libminiamr_host.so")


def load_host_library(path=HOST_LIB_PATH):
    """Load the driver library (include/amr_run.h and the driver half of
    include/amr_batch.h)."""
    if not os.path.exists(path):
        raise AmrError("%s is not built: run `make -C miniamr_amd`" % path)
    lib = _cdll(path)
    vp, ip, dp = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    argv_t = ctypes.POINTER(ctypes.c_char_p)
    sig = {
        "amr_run_create": (vp, [ip, argv_t]),
        "amr_run_step": (ip, [vp]),
        "amr_run_stages": (ip, [vp, ip]),
        "amr_run_finish": (ip, [vp]),
        "amr_run_stats_get": (ip, [vp, ctypes.POINTER(RunStats)]),
        "amr_run_report": (None, [vp]),
        "amr_run_device": (vp, [vp]),
        "amr_run_active": (ip, [vp]),
        "amr_run_block": (ip, [vp, ip, ctypes.POINTER(ctypes.c_int32),
                                ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(dp),
                                ctypes.POINTER(dp)]),
        "amr_run_destroy": (None, [vp]),
        "amr_run_last_error": (ctypes.c_char_p, []),
        "amr_batch_run_create": (vp, [ip, ctypes.POINTER(ctypes.c_int),
                                      ctypes.POINTER(argv_t), ip]),
        "amr_batch_run_step": (ip, [vp]),
        "amr_batch_run_finish": (ip, [vp]),
        "amr_batch_run_instances": (ip, [vp]),
        "amr_batch_run_stats_get": (ip, [vp, ip, ctypes.POINTER(RunStats)]),
        "amr_batch_run_get": (vp, [vp, ip]),
        "amr_batch_run_batch": (vp, [vp]),
        "amr_batch_run_tune_ms": (dp, [vp]),
        "amr_batch_run_device_ms": (dp, [vp]),
        "amr_batch_run_destroy": (None, [vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


def _host_error(lib):
    return (lib.amr_run_last_error() or b"").decode(errors="replace").strip()


def _argv(args, prog=b"miniAMR.x"):
    toks = [prog] + [x.encode() for x in (args.split() if isinstance(args, str) else args)]
    arr = (ctypes.c_char_p * (len(toks) + 1))(*toks, None)
    return len(toks), arr


class Run:
    """One drop-in driver run (include/amr_run.h), stepped from Python."""

    def __init__(self, args, host=None):
        self.lib = host or load_host_library()
        self._argc, self._argv = _argv(args)
        self.h = self.lib.amr_run_create(self._argc, self._argv)
        if not self.h:
            raise AmrError("amr_run_create: " + _host_error(self.lib))

    def step(self):
        """0: more timesteps follow, 1: done, 2: stopped by the conservation
        check; raises AmrError on a fatal error."""
        rc = self.lib.amr_run_step(self.h)
        if rc < 0:
            raise AmrError("amr_run_step: " + _host_error(self.lib))
        return rc

    def stats(self):
        st = RunStats()
        self.lib.amr_run_stats_get(self.h, ctypes.byref(st))
        return st

    def finish(self):
        rc = self.lib.amr_run_finish(self.h)
        if rc < 0:
            raise AmrError("amr_run_finish: " + _host_error(self.lib))
        return rc

    def device(self):
        return self.lib.amr_run_device(self.h)

    def blocks(self):
        """Active blocks in sorted order: list of (slot, level, lo[3], hi[3])."""
        out = []
        for i in range(self.lib.amr_run_active(self.h)):
            slot, level = ctypes.c_int32(), ctypes.c_int32()
            lo, hi = (ctypes.c_double * 3)(), (ctypes.c_double * 3)()
            self.lib.amr_run_block(self.h, i, ctypes.byref(slot), ctypes.byref(level), lo, hi)
            out.append((slot.value, level.value, list(lo), list(hi)))
        return out

    def close(self):
        if self.h:
            self.lib.amr_run_destroy(self.h)
            self.h = None


class BatchRun:
    """Many drop-in driver runs stepped together on one GPU
    (include/amr_batch.h, driver half)."""

    def __init__(self, runs, autotune=True, host=None):
        self.lib = host or load_host_library()
        self._keep = [_argv(r, b"miniAMR_batch.x") for r in runs]
        n = len(runs)
        argcs = (ctypes.c_int * n)(*[k[0] for k in self._keep])
        argvs = (ctypes.POINTER(ctypes.c_char_p) * n)(
            *[ctypes.cast(k[1], ctypes.POINTER(ctypes.c_char_p)) for k in self._keep])
        self.n = n
        self.h = self.lib.amr_batch_run_create(n, argcs, argvs, 1 if autotune else 0)
        if not self.h:
            raise AmrError("amr_batch_run_create: " + _host_error(self.lib))

    def step(self):
        rc = self.lib.amr_batch_run_step(self.h)
        if rc < 0:
            raise AmrError("amr_batch_run_step: " + _host_error(self.lib))
        return rc

    def stats(self, i):
        st = RunStats()
        self.lib.amr_batch_run_stats_get(self.h, i, ctypes.byref(st))
        return st

    def tune_ms(self):
        return self.lib.amr_batch_run_tune_ms(self.h)

    def device_ms(self):
        return self.lib.amr_batch_run_device_ms(self.h)

    def shape(self, devlib):
        sh = BatchShape()
        devlib.amr_batch_get_shape(self.lib.amr_batch_run_batch(self.h), ctypes.byref(sh))
        return {f: getattr(sh, f) for f, _ in BatchShape._fields_}

    def finish(self):
        return self.lib.amr_batch_run_finish(self.h)

    def close(self):
        if self.h:
            self.lib.amr_batch_run_destroy(self.h)
            self.h = None


def _ptr(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


class Device:
    """One amr_dev context.  Cells are exchanged as float64 arrays of shape
    (num_vars, nx+2, ny+2, nz+2), the reference's array[var][i][j][k]."""

    def __init__(self, nx, ny, nz, num_vars, max_blocks, stencil=7, precision=PREC_F64,
                 device=0, lib=None):
        self.lib = lib if lib is not None else load_library()
        self.cfg = Config(nx, ny, nz, num_vars, max_blocks, stencil, precision, device)
        self.shape = (num_vars, nx + 2, ny + 2, nz + 2)
        h = ctypes.c_void_p()
        rc = self.lib.amr_dev_create(ctypes.byref(self.cfg), ctypes.byref(h))
        self.h = h
        if rc != AMR_OK:
            msg = self.lib.amr_dev_last_error(h) if h.value else b"no context"
            self.close()
            raise AmrError("amr_dev_create failed (%d): %s" % (rc, msg.decode()))

    def _check(self, rc, what):
        if rc != AMR_OK:
            raise AmrError("%s failed (%d): %s" % (
                what, rc, self.lib.amr_dev_last_error(self.h).decode()))

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.amr_dev_destroy(self.h)
        self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, slot, cells):
        a = np.ascontiguousarray(cells, dtype=np.float64).reshape(self.shape)
        self._check(self.lib.amr_dev_upload_block(self.h, slot, _ptr(a, ctypes.c_double)),
                    "upload")

    def download(self, slot):
        a = np.empty(self.shape, dtype=np.float64)
        self._check(self.lib.amr_dev_download_block(self.h, slot, _ptr(a, ctypes.c_double)),
                    "download")
        return a

    def set_mesh(self, sorted_slots, jobs_by_dir):
        s = np.ascontiguousarray(sorted_slots, dtype=np.int32)
        parts = [np.asarray(j, dtype=FACE_JOB) for j in jobs_by_dir]
        jobs = np.ascontiguousarray(np.concatenate(parts) if parts else np.zeros(0, FACE_JOB))
        cnt = np.array([len(p) for p in parts], dtype=np.int32)
        self._check(self.lib.amr_dev_set_mesh(self.h, _ptr(s, ctypes.c_int32), len(s),
                                              jobs.ctypes.data_as(ctypes.c_void_p),
                                              _ptr(cnt, ctypes.c_int32)), "set_mesh")

    def comm(self, start, num, stage=0, permute=False):
        self._check(self.lib.amr_dev_comm(self.h, start, num, stage, int(permute)), "comm")

    def stencil(self, var):
        self._check(self.lib.amr_dev_stencil(self.h, var), "stencil")

    def stencil_stage(self, var, calc_stage):
        """stencil_driver(var, calc_stage) (ref/stencil.c:43)."""
        self._check(self.lib.amr_dev_stencil_stage(self.h, var, calc_stage), "stencil_stage")

    def set_materials(self, a1, a0):
        """--stencil 0 coefficients (ref/init.c:418-423): a1, a0[num_vars/4]."""
        a = np.ascontiguousarray(a0, dtype=np.float64)
        self._check(self.lib.amr_dev_set_materials(self.h, float(a1), _ptr(a, ctypes.c_double),
                                                   len(a)), "set_materials")

    def check_counts(self):
        """stencil_check's clamps so far: (divides, scales)."""
        c = (ctypes.c_int64 * 2)()
        self._check(self.lib.amr_dev_read_check_counts(self.h, c), "read_check_counts")
        return int(c[0]), int(c[1])

    def check_sum(self, var):
        out = ctypes.c_double()
        self._check(self.lib.amr_dev_check_sum(self.h, var, ctypes.byref(out)), "check_sum")
        return out.value

    def split(self, jobs):
        j = np.ascontiguousarray(jobs, dtype=SPLIT_JOB)
        self._check(self.lib.amr_dev_split(self.h, j.ctypes.data_as(ctypes.c_void_p), len(j)),
                    "split")

    def split_noalloc(self, jobs):
        """amr_dev_split staged in the library's static scratch array."""
        j = np.ascontiguousarray(jobs, dtype=SPLIT_JOB)
        self._check(self.lib.amr_dev_split_noalloc(self.h, j.ctypes.data_as(ctypes.c_void_p),
                                                   len(j)), "split_noalloc")

    def face_msg_len(self, code, dir, face_case, unpack=False):
        """Values per variable of one message (ref/init.c:77-117)."""
        n = self.lib.amr_dev_face_msg_len(self.h, code, dir, face_case, int(unpack))
        self._check(min(n, 0), "face_msg_len")
        return n

    def _codec(self, fn, what, code, dir, start, num, msgs, buf):
        """buf: a torch tensor on the device (its storage is the message
        buffer) or an integer device address with a length: (addr, n)."""
        m = np.ascontiguousarray(msgs, dtype=FACE_MSG)
        if isinstance(buf, tuple):
            addr, n = buf
        else:
            addr, n = buf.data_ptr(), buf.numel()
        self._check(fn(self.h, code, dir, start, num, m.ctypes.data_as(ctypes.c_void_p), len(m),
                       ctypes.c_void_p(addr), n), what)

    def pack_faces(self, code, dir, start, num, msgs, buf):
        self._codec(self.lib.amr_dev_pack_faces, "pack_faces", code, dir, start, num, msgs, buf)

    def unpack_faces(self, code, dir, start, num, msgs, buf):
        self._codec(self.lib.amr_dev_unpack_faces, "unpack_faces", code, dir, start, num, msgs,
                    buf)

    def consolidate(self, jobs):
        j = np.ascontiguousarray(jobs, dtype=MERGE_JOB)
        self._check(self.lib.amr_dev_consolidate(self.h, j.ctypes.data_as(ctypes.c_void_p),
                                                 len(j)), "consolidate")

    def reserve_sums(self, rows):
        self._check(self.lib.amr_dev_reserve_sums(self.h, rows), "reserve_sums")

    def read_sums(self, row0, nrows):
        out = np.empty((nrows, self.cfg.num_vars), dtype=np.float64)
        self._check(self.lib.amr_dev_read_sums(self.h, row0, nrows, _ptr(out, ctypes.c_double)),
                    "read_sums")
        return out

    def stage(self, start, num, stage=0, permute=False, sum_row=-1):
        self._check(self.lib.amr_dev_stage(self.h, start, num, stage, int(permute), sum_row),
                    "stage")

    def stage_fused(self, start, num, stage=0, permute=False, sum_row=-1):
        self._check(self.lib.amr_dev_stage_fused(self.h, start, num, stage, int(permute),
                                                 sum_row), "stage_fused")

    def sync(self):
        self._check(self.lib.amr_dev_sync(self.h), "sync")

    def set_timing(self, on=True):
        self._check(self.lib.amr_dev_set_timing(self.h, int(on)), "set_timing")

    def read_timing(self):
        ms = np.zeros(4, dtype=np.float64)
        self._check(self.lib.amr_dev_read_timing(self.h, _ptr(ms, ctypes.c_double)),
                    "read_timing")
        return ms


def run_driver(args, env=None, cwd=None, timeout=None, exe=EXE_PATH):
    """Run the drop-in driver with a reference-style argument string."""
    if not os.path.exists(exe):
        raise AmrError("%s is not built: run `make -C miniamr_amd`" % exe)
    e = dict(os.environ)
    if env:
        e.update(env)
    argv = [exe] + (args.split() if isinstance(args, str) else list(args))
    return subprocess.run(argv, env=e, cwd=cwd, stdout=subprocess.PIPE,
                          stderr=subprocess.STDOUT, timeout=timeout)
