#!/usr/bin/env python3
"""Benchmark: miniAMR 27-point, 4x4x4 initial blocks of 16^3 cells, 40
variables, 4 refinement levels, two moving objects (the bandwidth-scale
workload, ~3-5.5k active blocks) on one MI355X.

A *step* is one timestep of the drop-in driver: 20 stages, each an
exchange of all 40 variables plus the 27-point stencil over every active
block (checksums every 5 stages), then object motion and, every 4th
timestep, refinement.  The workload has 10 timesteps; steps run through
back-to-back fresh runs of it (step t is timestep t % 10 + 1 of run
t // 10), each run's setup (allocation, initial values, initial
refinement) outside the timed region.  The default --warmup 5 --steps 20
times timesteps 6-10 of the first run, all of the second and 1-5 of the
third: every timestep of the workload exactly twice.  Cells start
resident in HBM; nothing crosses PCIe inside the timed region except the
per-timestep checksum read-back (40 doubles per checksum stage).

value = cell-variable updates per second over the timed timesteps
        (updates = sum over stages of active blocks x 16^3 x 40).
roofline: the stencil kernel.  Algorithmic bytes per launch = one read and
        one write of every updated cell (2 x 8 B in FP64);
        achieved = those bytes / the stencil's device time measured with
        HIP events on the library's stream over the timed region.
cpu_baseline: the reference's OpenMP build (oracle/_ref/miniAMR_omp.x) on
        this host, one MPI rank, on the whole 10-timestep workload, rate
        from its Summary line, next to the GPU's rate over the same
        timesteps.

--workload adjoint times the differentiation form's reverse pass (the
adjoint kernels over the whole tape of this workload, FP32 storage) next
to a forward replay of the same tape; the default run reports a summary
under "differentiation_form".

--workload batch64 instead times the batching form: the 64 independent
runs of miniamr_amd/configs/batch64.txt (8^3/12^3/16^3 blocks, 8/20/40
variables, 2-4 refinement levels, FP32 storage) stepped together by
amr_batch_run (persistent kernels, one graph per timestep, autotuned
shapes), all 10 timesteps timed, next to the same 64 runs done one after
another with the single-run driver.  The default run reports a summary of
it under "batch_form".
"""
import argparse
import ctypes
import json
import os
import re
import subprocess
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

import miniamr_amd as M  # noqa: E402

CFG1 = ("--init_x 4 --init_y 4 --init_z 4 --nx 16 --ny 16 --nz 16 --num_vars 40 "
        "--stencil 27 --num_tsteps 10 --stages_per_ts 20 --checksum_freq 5 --num_refine 4 "
        "--refine_freq 4 --max_blocks 8000 --error_tol -10 --num_objects 2 "
        "--object 2 0 0.15 0.15 0.15 0.02 0.02 0.02 0.1 0.1 0.1 0 0 0 "
        "--object 0 0 0.5 0.5 0.5 0.02 0.02 0.02 0.05 0.05 0.05 0 0 0")
HBM_PEAK_GBS = 8000.0
# device timing level of the timed runs (amr_dev_set_timing); None keeps the
# driver's (1: exchange, stencil, checksum and split/consolidate spans)
TIMING_LEVEL = None
PREC_NAME = {0: "f64", 1: "f32", 2: "bf16"}
PREC_BYTES = {0: 8, 1: 4, 2: 2}


class Stats(ctypes.Structure):
    _fields_ = [("timesteps_done", ctypes.c_int64), ("stages_done", ctypes.c_int64),
                ("block_stages", ctypes.c_int64), ("cell_var_updates", ctypes.c_double),
                ("active_blocks", ctypes.c_int64), ("max_slots", ctypes.c_int64),
                ("device_ms", ctypes.c_double * 4), ("stage_calls", ctypes.c_int64),
                ("nx", ctypes.c_int), ("ny", ctypes.c_int),
                ("nz", ctypes.c_int), ("num_vars", ctypes.c_int), ("stencil", ctypes.c_int),
                ("precision", ctypes.c_int)]


def load_host():
    path = os.path.join(M.PKG, "lib", "libminiamr_host.so")
    if not os.path.exists(path):
        raise M.AmrError("%s is not built: run `make -C miniamr_amd`" % path)
    lib = M._cdll(path)
    lib.amr_run_create.restype = ctypes.c_void_p
    lib.amr_run_create.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p)]
    for name in ("amr_run_step", "amr_run_finish"):
        getattr(lib, name).restype = ctypes.c_int
        getattr(lib, name).argtypes = [ctypes.c_void_p]
    lib.amr_run_stats_get.restype = ctypes.c_int
    lib.amr_run_stats_get.argtypes = [ctypes.c_void_p, ctypes.POINTER(Stats)]
    lib.amr_run_report.argtypes = [ctypes.c_void_p]
    lib.amr_run_destroy.argtypes = [ctypes.c_void_p]
    lib.amr_run_last_error.restype = ctypes.c_char_p
    lib.amr_run_last_error.argtypes = []
    lib.amr_run_stages.restype = ctypes.c_int
    lib.amr_run_stages.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.amr_run_device.restype = ctypes.c_void_p
    lib.amr_run_device.argtypes = [ctypes.c_void_p]
    lib.amr_dev_set_timing.restype = ctypes.c_int
    lib.amr_dev_set_timing.argtypes = [ctypes.c_void_p, ctypes.c_int]
    return lib


def stats(lib, run):
    s = Stats()
    lib.amr_run_stats_get(run, ctypes.byref(s))
    return s


def cpu_baseline(per_ts, timeout=300):
    """The reference OpenMP build on the whole 10-timestep workload, with
    the GPU's rate over the same timesteps (from the timed steps) beside it."""
    exe = os.path.join(REPO, "oracle", "_ref", "miniAMR_omp.x")
    if not os.path.exists(exe):
        return None
    threads = min(16, os.cpu_count() or 1)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close")
    try:
        p = subprocess.run([exe] + CFG1.split(), env=env, stdout=subprocess.PIPE,
                           stderr=subprocess.STDOUT, timeout=timeout, cwd="/tmp")
    except (subprocess.TimeoutExpired, OSError) as e:
        print("cpu baseline skipped: %s" % e, file=sys.stderr)
        return None
    out = p.stdout.decode(errors="replace")
    m = re.search(r"Summary: ranks \d+(?: threads \d+)? ts (\d+) time (\S+) calc (\S+) "
                  r"comm \S+ red \S+ refine \S+ blocks/ts (\S+)", out)
    if not m:
        print("cpu baseline: no Summary line (rc %d)" % p.returncode, file=sys.stderr)
        return None
    ts, t, calc, bpts = int(m.group(1)), float(m.group(2)), float(m.group(3)), float(m.group(4))
    updates = bpts * ts * 20 * 16 ** 3 * 40
    # the GPU over the same timesteps: every timed step of timestep 1..ts
    u = sum(per_ts[k][0] for k in range(1, ts + 1) if k in per_ts)
    s = sum(per_ts[k][1] for k in range(1, ts + 1) if k in per_ts)
    covered = [k for k in range(1, ts + 1) if k in per_ts]
    return {"value": round(updates / t / 1e9, 4), "unit": "GCell-var updates/s",
            "cores": threads, "kind": "reference",
            "sample": "reference openmp/ build, 1 MPI rank, %d OpenMP threads, the whole "
                      "workload (%d timesteps, %.0f blocks/stage, %.4e updates): %.2f s "
                      "(its timer_all: the timesteps plus the initial refinement), %.2f s "
                      "in the stencil => %.3f GCell-var/s stencil-only"
                      % (threads, ts, bpts, updates, t, calc, updates / calc / 1e9),
            "gpu_same_timesteps": {"timesteps": covered,
                                   "value": round(u / s / 1e9, 4) if s > 0 else None,
                                   "speedup": round(u / s / (updates / t), 1) if s > 0 else None}}


def copy_ceiling(bytes_=4 << 30, reps=3):
    """The streaming copy and read bandwidth of this GPU in GB/s
    (amr_dev_copy_bandwidth: one 16-byte nontemporal access per lane, one
    workgroup per 4 KiB tile), or None where it cannot run."""
    dev = M.load_library()
    c, r = ctypes.c_double(), ctypes.c_double()
    if dev.amr_dev_copy_bandwidth(bytes_, reps, ctypes.byref(c), ctypes.byref(r)):
        return None
    return {"copy_gbs": round(c.value, 1), "read_gbs": round(r.value, 1),
            "buffer_bytes": bytes_}


def run_workload(lib, precision, schedule, warmup, steps, sums_path=None):
    """Timesteps warmup+1 .. warmup+steps of the workload, timed.

    The workload is a 10-timestep run.  Bench steps are its timesteps taken
    in order from back-to-back fresh runs: step t (0-based over warmup and
    timed steps) is timestep t % 10 + 1 of run t // 10.  Creating a run
    (parse, allocate, initial values, initial refinement) and finishing it
    happen outside the timed region; every timed step is one whole timestep
    of a real run, and amr_run_step returns after the device is idle.
    The first run's checksums go to sums_path (all 1,640 of the workload)."""
    os.environ["MINIAMR_PRECISION"] = precision
    os.environ["MINIAMR_SCHEDULE"] = schedule
    argv = [b"miniAMR.x"] + [x.encode() for x in CFG1.split()] + [b"--report_perf", b"0"]
    arr = (ctypes.c_char_p * len(argv))(*argv)
    total = warmup + steps
    acc = {"updates": 0.0, "block_stages": 0, "stages": 0, "calls": 0,
           "dev": [0.0] * 4, "elapsed": 0.0}
    per_ts = {}                     # timestep -> [updates, seconds] over timed steps
    setup_s, t, first, last = 0.0, 0, True, None
    while t < total:
        if sums_path and first:
            os.environ["AMR_DUMP_SUMS"] = sums_path
        else:
            os.environ.pop("AMR_DUMP_SUMS", None)
        t_setup = time.time()
        run = lib.amr_run_create(len(argv), arr)
        if not run:
            raise M.AmrError("amr_run_create: " + (lib.amr_run_last_error() or b"").decode())
        if TIMING_LEVEL is not None and lib.amr_dev_set_timing(lib.amr_run_device(run),
                                                               TIMING_LEVEL):
            raise M.AmrError("amr_dev_set_timing failed")
        setup_s += time.time() - t_setup
        ts, rc = 0, 0
        while rc == 0:
            ts += 1
            timed = warmup <= t < total
            s0 = stats(lib, run) if timed else None
            t0 = time.perf_counter()
            rc = lib.amr_run_step(run)
            dt = time.perf_counter() - t0
            if rc < 0:
                raise M.AmrError("amr_run_step: " + (lib.amr_run_last_error() or b"").decode())
            if rc == 2:
                raise SystemExit("conservation check stopped the run")
            if timed:
                s1 = stats(lib, run)
                du = s1.cell_var_updates - s0.cell_var_updates
                acc["updates"] += du
                acc["block_stages"] += s1.block_stages - s0.block_stages
                acc["stages"] += s1.stages_done - s0.stages_done
                acc["calls"] += s1.stage_calls - s0.stage_calls
                for i in range(4):
                    acc["dev"][i] += s1.device_ms[i] - s0.device_ms[i]
                acc["elapsed"] += dt
                p = per_ts.setdefault(ts, [0.0, 0.0])
                p[0] += du
                p[1] += dt
                last = s1
            t += 1
            if t >= total and not (sums_path and first):
                break
        # a run whose checksums are logged finishes all its timesteps
        lib.amr_run_finish(run)
        lib.amr_run_destroy(run)
        first = False
    os.environ.pop("AMR_DUMP_SUMS", None)
    return {"acc": acc, "per_ts": per_ts, "last": last, "setup_s": setup_s,
            "runs": (total + 9) // 10}


def lockstep_cell_errors(lib, prec):
    """The FP64 run and a narrow-storage run of the workload side by side,
    stage chunk by stage chunk (amr_run_stages), their cells compared on the
    device (amr_dev_compare_cells) right after every checksum stage: per
    checksum interval the largest |cell - FP64 cell| over all interior
    cells of all active blocks and variables.  Untimed."""
    dev = M.load_library()
    os.environ["MINIAMR_SCHEDULE"] = "stage"
    os.environ.pop("AMR_DUMP_SUMS", None)
    runs = []
    for p in ("f64", prec):
        argv = [b"miniAMR.x"] + [x.encode() for x in CFG1.split()] + \
            [b"--report_perf", b"0", b"--precision", p.encode()]
        arr = (ctypes.c_char_p * len(argv))(*argv)
        r = lib.amr_run_create(len(argv), arr)
        if not r:
            raise M.AmrError("amr_run_create: " + (lib.amr_run_last_error() or b"").decode())
        runs.append(r)
    st = stats(lib, runs[0])
    nv, spt, cf = st.num_vars, 20, 5
    ends = [s for s in range(0, spt, cf)] + [spt - 1]      # chunks end after these stages
    chunks = [ends[0] + 1] + [b - a for a, b in zip(ends, ends[1:])]
    per_interval, per_var = [], [0.0] * nv
    out = (ctypes.c_double * nv)()
    rc = 0
    while rc == 0:
        for c, end in zip(chunks, ends):
            rcs = [lib.amr_run_stages(r, c) for r in runs]
            if min(rcs) < 0:
                raise M.AmrError("amr_run_stages: " + (lib.amr_run_last_error() or b"").decode())
            if end % cf == 0:                                   # a checksum stage
                if dev.amr_dev_compare_cells(lib.amr_run_device(runs[0]),
                                             lib.amr_run_device(runs[1]), 0, nv, out):
                    raise M.AmrError("amr_dev_compare_cells failed")
                per_interval.append(max(out))
                per_var = [max(a, b) for a, b in zip(per_var, out)]
            rc = max(rcs)
    for r in runs:
        lib.amr_run_finish(r)
        lib.amr_run_destroy(r)
    return {"max_cell_abs_err": float("%.3e" % max(per_interval)),
            "max_cell_abs_err_per_interval": [float("%.2e" % e) for e in per_interval],
            "max_cell_abs_err_per_variable": [float("%.2e" % e) for e in per_var],
            "cells_compared": "every interior cell of every active block and variable, after "
                              "each of the %d checksum stages" % len(per_interval)}


BATCH_TABLE = os.path.join(M.PKG, "configs", "batch64.txt")


def batch_runs():
    with open(BATCH_TABLE) as f:
        return [l.strip() for l in f if l.strip() and not l.startswith("#")]


def quiet(args):
    return args + " --report_perf 0"


def run_batch_workload(runs, autotune=True):
    """All timesteps of the runs as one batch; returns totals and timings."""
    t = time.time()
    B = M.BatchRun([quiet(r) for r in runs], autotune=autotune)
    setup = time.time() - t
    t0 = time.perf_counter()
    steps = 1
    while B.step() == 0:
        steps += 1
    t1 = time.perf_counter()
    st = [B.stats(i) for i in range(B.n)]
    out = {"updates": sum(x.cell_var_updates for x in st), "elapsed": t1 - t0,
           "setup_s": setup, "steps": steps, "shape": B.shape(M.load_library()),
           "tune_ms": B.tune_ms(), "device_s": B.device_ms() * 1e-3,
           "block_stages": [x.block_stages for x in st],
           "stages": [x.stages_done for x in st], "ebytes": PREC_BYTES[st[0].precision],
           "precision": st[0].precision}
    B.finish()
    B.close()
    return out


def run_sequential(runs):
    """The same runs one after another with the single-run driver."""
    host = M.load_host_library()
    updates = elapsed = 0.0
    for r in runs:
        R = M.Run(quiet(r), host=host)
        t0 = time.perf_counter()
        while R.step() == 0:
            pass
        elapsed += time.perf_counter() - t0
        updates += R.stats().cell_var_updates
        R.finish()
        R.close()
    return {"updates": updates, "elapsed": elapsed}


def batch_summary(b, seq):
    v = b["updates"] / b["elapsed"] / 1e9
    sv = seq["updates"] / seq["elapsed"] / 1e9 if seq else None
    avg = [bs / max(1, n) for bs, n in zip(b["block_stages"], b["stages"])]
    return {
        "value": round(v, 4), "unit": "GCell-var updates/s", "dtype": PREC_NAME[b["precision"]],
        "runs": len(avg), "timesteps": b["steps"], "ms_per_step": round(b["elapsed"] / b["steps"] * 1e3, 3),
        "cell_var_updates": b["updates"],
        "active_blocks_per_run": {"min": round(min(avg), 1), "max": round(max(avg), 1),
                                  "total": round(sum(avg), 1)},
        "graph_device_frac_of_hbm": round(b["updates"] * 2 * b["ebytes"] / b["device_s"] / 1e9
                                          / HBM_PEAK_GBS, 4) if b["device_s"] > 0 else None,
        "shape": dict(b["shape"], autotune_stage_ms=round(b["tune_ms"], 4)),
        "setup_s": round(b["setup_s"], 2),
        "sequential_value": round(sv, 4) if sv else None,
        "speedup_vs_sequential": round(v / sv, 3) if sv else None}


ADJ_SPHERE = ((0.5, 0.5, 0.5), 0.25)


def run_adjoint_workload(precision="f32"):
    """The reverse pass of the bandwidth-scale workload: one forward run
    with the tape on, an objective seeded on the blocks that meet a fixed
    sphere (N(0,1) weights), then the adjoint kernels over the whole tape
    (all 200 stages and the refinements, last first), timed, next to a
    forward replay of the same tape."""
    host, lib = M.load_host_library(), M.load_library()
    R = M.Run(quiet(CFG1) + " --tape --precision " + precision, host=host)
    while R.step() == 0:
        pass
    st = R.stats()
    dev = ctypes.c_void_p(R.device())
    lib.amr_dev_tape(dev, 0)
    lib.amr_dev_set_timing(dev, 0)
    n = ctypes.c_int()
    lib.amr_dev_tape_length(dev, ctypes.byref(n))
    first = 0
    op = (ctypes.c_int32 * 8)()
    while first < n.value and lib.amr_dev_tape_op(dev, first, op) == 0 and op[0] == 1:
        first += 1
    adj = ctypes.c_void_p()
    if lib.amr_adj_create(dev, ctypes.byref(adj)):
        raise M.AmrError(lib.amr_adj_last_error(adj).decode())
    import numpy as np
    rng = np.random.default_rng(7)
    (cx, cy, cz), r = ADJ_SPHERE
    seeded = 0
    for slot, _, lo, hi in R.blocks():
        near = [min(max(c, lo[d]), hi[d]) for d, c in enumerate((cx, cy, cz))]
        if sum((near[d] - (cx, cy, cz)[d]) ** 2 for d in range(3)) > r * r:
            continue
        w = rng.standard_normal((st.num_vars, st.nx + 2, st.ny + 2, st.nz + 2))
        lib.amr_adj_set_block(adj, slot, w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        seeded += 1
    lib.amr_adj_sync(adj)
    t0 = time.perf_counter()
    rc = lib.amr_adj_run(adj, first, n.value)
    lib.amr_adj_sync(adj)
    t_adj = time.perf_counter() - t0
    if rc:
        raise M.AmrError(lib.amr_adj_last_error(adj).decode())
    t0 = time.perf_counter()
    rc = lib.amr_dev_replay(dev, first, n.value)
    lib.amr_dev_sync(dev)
    t_fwd = time.perf_counter() - t0
    lib.amr_adj_destroy(adj)
    R.finish()
    R.close()
    return {"updates": st.cell_var_updates, "adjoint_s": t_adj, "replay_s": t_fwd,
            "ops": n.value - first, "seeded_blocks": seeded, "precision": precision}


def adjoint_summary(a):
    return {"value": round(a["updates"] / a["adjoint_s"] / 1e9, 4),
            "unit": "GCell-var updates/s (reverse pass)", "dtype": a["precision"],
            "forward_replay_value": round(a["updates"] / a["replay_s"] / 1e9, 4),
            "adjoint_over_forward_time": round(a["adjoint_s"] / a["replay_s"], 3),
            "adjoint_s": round(a["adjoint_s"], 3), "tape_ops": a["ops"],
            "seeded_blocks": a["seeded_blocks"], "cell_var_updates": a["updates"]}


def main_adjoint(a):
    r = run_adjoint_workload(a.precision if a.precision != "f64" else "f32")
    summ = adjoint_summary(r)
    line = {
        "metric": "GCell-var updates/s", "value": summ["value"], "unit": "GCell-var updates/s",
        "n_gpus": 1, "steps": 10, "warmup": 0,
        "ms_per_step": round(r["adjoint_s"] / 10 * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": summ["dtype"],
        "data": "synthetic: the bandwidth-scale workload's tape, N(0,1) objective weights",
        "config": {"workload": "differentiation form: reverse pass (adjoint kernels) over the "
                               "whole tape of the 4x4x4 x 16^3, 40-variable, 4-level workload",
                   "tape_ops": summ["tape_ops"], "seeded_blocks": summ["seeded_blocks"],
                   "cell_var_updates": r["updates"]},
        "roofline": {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(r["updates"] * 2 * PREC_BYTES[1] / r["adjoint_s"] / 1e9 /
                                   HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "whole reverse pass (k_adj_stencil, k_adj_exchange, split/merge "
                               "adjoints), wall time"},
        "forward_replay": {"value": summ["forward_replay_value"],
                           "adjoint_over_forward_time": summ["adjoint_over_forward_time"]},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def main_batch(a):
    runs = batch_runs()
    b = run_batch_workload(runs)
    seq = None if a.no_sequential else run_sequential(runs)
    summ = batch_summary(b, seq)
    value = summ["value"]
    line = {
        "metric": "GCell-var updates/s",
        "value": value,
        "unit": "GCell-var updates/s",
        "n_gpus": 1,
        "steps": b["steps"],
        "warmup": 0,
        "ms_per_step": summ["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": summ["dtype"],
        "data": "synthetic: uniform [0,1) cell values, seeds 1..64",
        "config": {"workload": "batching form: 64 independent miniAMR 27-pt runs "
                               "(miniamr_amd/configs/batch64.txt), all 10 timesteps",
                   "runs": summ["runs"], "active_blocks_per_run": summ["active_blocks_per_run"],
                   "cell_var_updates": b["updates"], "shape": summ["shape"]},
        "roofline": {"bound": "hbm",
                     "achieved": round(b["updates"] * 2 * b["ebytes"] / b["device_s"] / 1e9, 1)
                     if b["device_s"] > 0 else None,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": summ["graph_device_frac_of_hbm"], "traffic": None,
                     "kernel": "timestep graphs (kb_exchange x3 + kb_stencil + kb_reduce per "
                               "stage), device time of the graph launches"},
        "sequential": {"value": summ["sequential_value"],
                       "speedup": summ["speedup_vs_sequential"]},
        "setup_s": summ["setup_s"],
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def read_sums(path):
    with open(path) as f:
        return [float(x) for x in f.read().split()]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--precision", default=os.environ.get("MINIAMR_PRECISION", "f64"),
                    choices=["f64", "f32", "bf16"])
    ap.add_argument("--schedule", default=os.environ.get("MINIAMR_SCHEDULE", "stage"),
                    choices=["stage", "var"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-precision-report", action="store_true")
    ap.add_argument("--workload", default="cfg1", choices=["cfg1", "batch64", "adjoint"])
    ap.add_argument("--no-adjoint-report", action="store_true")
    ap.add_argument("--no-batch-report", action="store_true")
    ap.add_argument("--no-sequential", action="store_true")
    ap.add_argument("--timing-level", type=int, default=None, choices=[0, 1, 2, 3])
    a = ap.parse_args()
    global TIMING_LEVEL
    TIMING_LEVEL = a.timing_level
    if a.gpus != 1:
        print("bench.py: the benchmark runs one rank on one GPU; --gpus %d is not supported "
              "(blocks would shard across GPUs as the reference's ranks do; not built)" % a.gpus,
              file=sys.stderr)
        sys.exit(2)
    if a.workload == "batch64":
        main_batch(a)
        return
    if a.workload == "adjoint":
        main_adjoint(a)
        return
    if a.steps < 1 or a.warmup < 0:
        print("bench.py: need --steps >= 1 and --warmup >= 0", file=sys.stderr)
        sys.exit(2)

    lib = load_host()
    import tempfile
    tmp = tempfile.mkdtemp(prefix="miniamr_bench_")
    main_sums = os.path.join(tmp, "sums_%s.txt" % a.precision)
    r = run_workload(lib, a.precision, a.schedule, a.warmup, a.steps, main_sums)
    acc, s1, elapsed = r["acc"], r["last"], r["acc"]["elapsed"]
    updates = acc["updates"]
    value = updates / elapsed / 1e9
    ebytes = PREC_BYTES[s1.precision]
    stencil_s = acc["dev"][1] * 1e-3
    launches = acc["calls"]
    achieved = updates * 2 * ebytes / stencil_s / 1e9 if stencil_s > 0 else None
    traffic = None
    pmc = os.path.join(REPO, "profiles", "stencil27_cfg1_%s_pmc.json" % a.precision)
    if os.path.exists(pmc):
        with open(pmc) as f:
            per_bv = json.load(f).get("hbm_bytes_per_block_variable")
        if per_bv and launches:
            block_vars = acc["block_stages"] * s1.num_vars / launches
            traffic = round(per_bv * block_vars)
    avg_blocks = acc["block_stages"] / max(1, acc["stages"])
    ceiling = copy_ceiling()
    timed_ts = sorted(r["per_ts"])
    same = [r["per_ts"][k] for k in timed_ts]
    gpu_sample = {"timesteps": timed_ts,
                  "value": round(sum(u for u, _ in same) / sum(s for _, s in same) / 1e9, 4)}

    precision_form = None
    if not a.no_precision_report and a.precision == "f64":
        ref = read_sums(main_sums)
        precision_form = {}
        for prec in ("f32", "bf16"):
            path = os.path.join(tmp, "sums_%s.txt" % prec)
            q = run_workload(lib, prec, a.schedule, a.warmup, a.steps, path)
            got = read_sums(path)
            n = min(len(ref), len(got))
            errs = [abs(g - f) / abs(f) for g, f in zip(got[:n], ref[:n])]
            v = q["acc"]["updates"] / q["acc"]["elapsed"] / 1e9
            nv = s1.num_vars
            ev = [errs[i:i + nv] for i in range(0, n - n % nv, nv)]   # event 0: initial sums
            precision_form[prec] = {
                "value": round(v, 4), "ratio_vs_f64": round(v / value, 3),
                "checksums_compared": n,
                "max_checksum_rel_err_vs_f64": float("%.3e" % max(errs)) if errs else None,
                "final_event_max_rel_err": float("%.3e" % max(errs[-nv:])) if errs else None,
                # per checksum interval (every 5th stage, all variables) and
                # per variable (over all intervals)
                "checksum_rel_err_per_interval": [float("%.2e" % max(e)) for e in ev[1:]],
                "checksum_rel_err_per_variable": [
                    float("%.2e" % max(e[v] for e in ev[1:])) for v in range(nv)]
                if len(ev) > 1 else None}
            precision_form[prec].update(lockstep_cell_errors(lib, prec))

    line = {
        "metric": "GCell-var updates/s",
        "value": round(value, 4),
        "unit": "GCell-var updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(elapsed / a.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": PREC_NAME[s1.precision],
        "data": "synthetic: uniform [0,1) cell values from the reference's generator and seed",
        "config": {"workload": "miniAMR 27-pt, init 4x4x4 blocks of 16^3, 40 vars, "
                               "num_refine 4, 2 moving objects, 10 timesteps x 20 stages",
                   "steps": "a step is one timestep; steps run through back-to-back fresh "
                            "runs of the 10-timestep workload (step t = timestep t %% 10 + 1 "
                            "of run t // 10), run setup untimed: warmup %d + timed %d = "
                            "%d runs" % (a.warmup, a.steps, r["runs"]),
                   "schedule": a.schedule,
                   "ghosts": ("same-level faces read in place"
                              if os.environ.get("MINIAMR_FUSED", "1") != "0" else "plain"),
                   "avg_active_blocks": round(avg_blocks, 1),
                   "cell_var_updates": updates},
        "roofline": {"bound": "hbm",
                     "achieved": round(achieved, 1) if achieved else None,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
                     "traffic": traffic,
                     # PMC-measured bytes per launch over the launch time: the
                     # fraction of the HBM peak the kernel actually moves
                     "traffic_frac": (round(traffic / (stencil_s / launches) / 1e9 /
                                            HBM_PEAK_GBS, 4)
                                      if traffic and launches and stencil_s > 0 else None),
                     "kernel": "k_stencil_v16 (27-point, 16x16 faces; every stage reads "
                               "same-level faces in place but the last before a refinement "
                               "or the run's end), %d launches, %.3f ms avg" % (
                                   launches, stencil_s / max(1, launches) * 1e3),
                     # the practical ceiling on this GPU: a streaming copy
                     # and read (amr_dev_copy_bandwidth, 4 GiB, best of 3),
                     # and the stencil's PMC-measured bytes against it
                     "copy_ceiling": ceiling,
                     "traffic_frac_of_copy": (round(traffic / (stencil_s / launches) / 1e9 /
                                                    ceiling["copy_gbs"], 4)
                                              if traffic and launches and stencil_s > 0
                                              and ceiling else None),
                     "step_level_frac": round(updates * 2 * ebytes / elapsed / 1e9 /
                                              HBM_PEAK_GBS, 4)},
        "device_ms": {"exchange": round(acc["dev"][0], 2),
                      "stencil": round(acc["dev"][1], 2),
                      "checksum": round(acc["dev"][2], 2),
                      "split_merge": round(acc["dev"][3], 2)},
        "setup_s": round(r["setup_s"], 2),
        "precision_form": precision_form,
        "batch_form": None if a.no_batch_report else batch_summary(
            run_batch_workload(batch_runs()),
            None if a.no_sequential else run_sequential(batch_runs())),
        "differentiation_form": None if a.no_adjoint_report else adjoint_summary(
            run_adjoint_workload("f32")),
        "cpu_baseline": None if a.no_cpu_baseline else cpu_baseline(r["per_ts"]),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
