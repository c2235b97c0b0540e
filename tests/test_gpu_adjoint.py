"""Parity of the differentiation form (GPU).

The adjoint kernels (include/amr_adjoint.h) are checked against FP64 torch
autograd of the torch restatement (oracle/adjoint_torch.py, itself pinned
against the C oracle on the CPU) and against central finite differences
of the GPU forward:
  * a synthetic recorded sequence with every exchange rule, both direction
    orders, separate exchanges and stencils, a split and a consolidation
    with slot reuse (7- and 27-point): the recorded tape equals the
    sequence, and the gradient with respect to every uploaded element
    (ghosts included) matches autograd to 1e-12 of its largest entry;
  * the differentiation workload (2x2x2 blocks of 8^3, 4 variables,
    7-point, num_refine 2 around a static sphere, 40 stages, FP64):
    objective J = sum over the cells whose centres lie in the sphere of
    w * u_final, w ~ N(0,1) seeded; the GPU forward value and the final
    cells equal the restatement's, the adjoint gradient equals autograd's
    (1e-12) and finite differences with h = 1e-6 at 200 seeded cells
    (|FD - g| <= 1e-7 + 1e-6 |g|), a loss that also uses a checkpoint
    (stage 16) gets the right gradient through the segment chain, and the
    FP32 adjoint agrees with FP64 to 1e-4 of the largest entry.
"""
import ctypes
import os
import sys

import numpy as np
import pytest
import torch

import amrcheck as A
import miniamr_amd as M
import tapegen
from conftest import require

sys.path.insert(0, os.path.join(A.REPO, "oracle"))
import adjoint_torch as AT  # noqa: E402

pytestmark = pytest.mark.gpu

CFG4 = ("--init_x 2 --init_y 2 --init_z 2 --nx 8 --ny 8 --nz 8 --num_vars 4 --stencil 7 "
        "--num_refine 2 --num_tsteps 2 --stages_per_ts 20 --num_objects 1 "
        "--object 2 0 0.3 0.3 0.3 0 0 0 0.2 0.2 0.2 0 0 0")
SPHERE = ((0.3, 0.3, 0.3), 0.2)


def _lib():
    require(M.LIB_PATH, "run __graft_entry__.build()")
    return M.load_library(M.LIB_PATH)


# code 1: the plans' jobs move whole ghost planes (--code 1/2); their
# prolongations share coarse ring rows, so the adjoint takes its atomic path
@pytest.mark.parametrize("ghosts", [False, True], ids=["code0", "code1"])
@pytest.mark.parametrize("stencil", [7, 27])
def test_synthetic_tape_gradient_matches_autograd(stencil, ghosts):
    print("case: synthetic tape stencil=%d ghosts=%s" % (stencil, ghosts), flush=True)
    lib = _lib()
    t = tapegen.build(50 + stencil, stencil=stencil, ghosts=ghosts)
    dev = M.Device(lib=lib, **t["cfg"])
    assert lib.amr_dev_tape(dev.h, 1) == 0
    tapegen.apply(dev, t)
    lib.amr_dev_tape(dev.h, 0)
    ops, plans, jobs = AT.read_tape(lib, dev.h, M.FACE_JOB, M.SPLIT_JOB, M.MERGE_JOB)
    assert [o[:5] for o in ops] == [o[:5] for o in t["ops"]]
    assert [o[7] for o in ops] == [o[7] for o in t["ops"]]
    for p in plans:
        assert np.array_equal(plans[p][0], t["plans"][p][0])
        assert plans[p][1].tobytes() == t["plans"][p][1].astype(M.FACE_JOB).tobytes()
    c = t["cfg"]
    R = AT.Restatement(ops, plans, jobs, (c["nx"], c["ny"], c["nz"]), stencil, c["num_vars"],
                       c["max_blocks"])
    u0 = torch.from_numpy(t["u0"]).requires_grad_(True)
    u, _ = R.run(u0, t["uploads"])
    for s in range(c["max_blocks"]):   # the GPU forward equals the restatement
        assert A.ulp_diff(u[:, s].detach().numpy(), dev.download(s)).max() == 0
    w = np.random.default_rng(7).standard_normal((c["max_blocks"],) + t["u0"].shape[1:])
    J = (u * torch.from_numpy(w).transpose(0, 1)).sum()
    (g_ref,) = torch.autograd.grad(J, u0)
    adj = ctypes.c_void_p()
    assert lib.amr_adj_create(dev.h, ctypes.byref(adj)) == 0
    try:
        for s in range(c["max_blocks"]):
            wb = np.ascontiguousarray(w[s])
            assert lib.amr_adj_set_block(adj, s, wb.ctypes.data_as(
                ctypes.POINTER(ctypes.c_double))) == 0
        assert lib.amr_adj_run(adj, len(t["uploads"]), len(ops)) == 0, \
            lib.amr_adj_last_error(adj)
        g = np.empty_like(t["u0"])
        for b, s in enumerate(t["uploads"]):
            assert lib.amr_adj_get_block(adj, s, g[b].ctypes.data_as(
                ctypes.POINTER(ctypes.c_double))) == 0
    finally:
        lib.amr_adj_destroy(adj)
    err = np.abs(g - g_ref.numpy()).max() / np.abs(g_ref.numpy()).max()
    print("max rel gradient error %.3e" % err, flush=True)
    assert err <= 1e-12


@pytest.fixture(scope="module")
def taped():
    require(M.HOST_LIB_PATH, "run __graft_entry__.build()")
    from miniamr_amd import autodiff as AD
    tr = AD.TapedRun(CFG4, k=8)
    yield tr
    tr.close()


def weights(tr, seed=2026):
    rng = np.random.default_rng(seed)
    (cx, cy, cz), r = SPHERE
    w = np.zeros((len(tr.blocks),) + tr.block_shape)
    for b, blk in enumerate(tr.blocks):
        x, y, z = tr.cell_centres(blk)
        inside = (x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2 <= r * r
        w[b] = rng.standard_normal(tr.block_shape) * inside[None]
    return torch.from_numpy(w)


def initial(tr, seed=1):
    rng = np.random.default_rng(seed)
    return torch.from_numpy(rng.random((len(tr.upload_slots),) + tr.block_shape))


def restatement(tr, lib):
    ops, plans, jobs = AT.read_tape(lib, tr.dev, M.FACE_JOB, M.SPLIT_JOB, M.MERGE_JOB)
    used = [s for s, _, _, _ in tr.blocks] + list(tr.upload_slots)
    for j in jobs.values():
        other = j["parent"] if "parent" in j.dtype.names else j["dst"]
        used += [int(x) for x in j["child"]] + [int(x) for x in other]
    for p in plans.values():
        used += [int(s) for s in p[0]]
        used += [int(x) for x in p[1]["dst"]] + [int(x) for x in p[1]["src"]]
    slots = 1 + max(used)
    return AT.Restatement(ops, plans, jobs, (tr.nx, tr.ny, tr.nz), 7, tr.num_vars, slots), ops


def final_cells(tr, u):
    return torch.stack([u[:, s, 1:-1, 1:-1, 1:-1] for s, _, _, _ in tr.blocks])


def test_workload_mesh(taped):
    levels = [lv for _, lv, _, _ in taped.blocks]
    assert [levels.count(l) for l in range(3)] == [4, 24, 64]
    assert taped.n_stages() == 40 and len(taped.segments) == 5


def test_forward_and_gradient_match_autograd(taped):
    print("case: differentiation workload, autograd", flush=True)
    lib = _lib()
    w = weights(taped)
    u0 = initial(taped).requires_grad_(True)
    cells, _ = taped.forward(u0)
    J = (cells * w).sum()
    J.backward()
    g = u0.grad.numpy().copy()
    R, ops = restatement(taped, lib)
    u0t = initial(taped).requires_grad_(True)
    u, _ = R.run(u0t, taped.upload_slots)
    ct = final_cells(taped, u)
    assert A.ulp_diff(cells.detach().numpy(), ct.detach().numpy()).max() == 0
    Jt = (ct * w).sum()
    assert abs(J.item() - Jt.item()) <= 1e-12 * abs(Jt.item())
    (gt,) = torch.autograd.grad(Jt, u0t)
    err = np.abs(g - gt.numpy()).max() / np.abs(gt.numpy()).max()
    print("J %.17g, max rel gradient error %.3e, nonzero %d" % (J.item(), err,
                                                              int((gt != 0).sum())), flush=True)
    assert err <= 1e-12 and (gt != 0).sum() > 1000


def test_gradient_matches_finite_differences(taped):
    print("case: differentiation workload, finite differences", flush=True)
    w = weights(taped)
    u0 = initial(taped).requires_grad_(True)
    cells, _ = taped.forward(u0)
    (cells * w).sum().backward()
    g = u0.grad.numpy()
    rng = np.random.default_rng(200)
    flat = rng.choice(u0.numel(), 200, replace=False)
    # half the probes where the gradient is largest, so the check is not
    # dominated by cells the objective does not reach
    flat[:100] = np.argsort(-np.abs(g).ravel())[:100]
    h = 1e-6
    base = initial(taped).numpy()
    worst = 0.0
    with torch.no_grad():
        for f in flat:
            idx = np.unravel_index(int(f), base.shape)
            up, dn = base.copy(), base.copy()
            up[idx] += h
            dn[idx] -= h
            jp = (taped.forward(torch.from_numpy(up))[0] * w).sum().item()
            jm = (taped.forward(torch.from_numpy(dn))[0] * w).sum().item()
            fd = (jp - jm) / (2 * h)
            worst = max(worst, abs(fd - g[idx]) / (1e-7 + 1e-6 * abs(g[idx])))
    print("worst FD ratio %.3f" % worst, flush=True)
    assert worst <= 1.0


def test_checkpoint_loss(taped):
    """A loss on the state after 16 stages (a segment boundary) as well as
    on the final cells: the gradient flows back through both paths."""
    print("case: differentiation workload, checkpoint loss", flush=True)
    lib = _lib()
    w, w2 = weights(taped), weights(taped, seed=77)
    u0 = initial(taped).requires_grad_(True)
    cells, cps = taped.forward(u0)
    from miniamr_amd.autodiff import Cells
    mid = Cells.apply(cps[1], taped)        # after stages 0-15 (mesh unchanged since ts 0)
    ((cells * w).sum() + (mid * w2).sum()).backward()
    R, ops = restatement(taped, lib)
    cut = taped.segments[1][1]
    u0t = initial(taped).requires_grad_(True)
    u, kept = R.run(u0t, taped.upload_slots, keep=(cut,))
    Jt = (final_cells(taped, u) * w).sum() + (final_cells(taped, kept[cut]) * w2).sum()
    (gt,) = torch.autograd.grad(Jt, u0t)
    err = np.abs(u0.grad.numpy() - gt.numpy()).max() / np.abs(gt.numpy()).max()
    print("max rel gradient error %.3e" % err, flush=True)
    assert err <= 1e-12


def test_fp32_adjoint_tracks_fp64(taped):
    print("case: differentiation workload, FP32", flush=True)
    from miniamr_amd import autodiff as AD
    tr32 = AD.TapedRun(CFG4 + " --precision f32", k=8)
    try:
        w = weights(taped)
        grads = []
        for tr in (taped, tr32):
            u0 = initial(taped).requires_grad_(True)
            cells, _ = tr.forward(u0)
            (cells * w).sum().backward()
            grads.append(u0.grad.numpy())
        err = np.abs(grads[0] - grads[1]).max() / np.abs(grads[0]).max()
        print("FP32 vs FP64 max rel gradient difference %.3e" % err, flush=True)
        assert err <= 1e-4
    finally:
        tr32.close()


def test_tape_refuses_stencil0():
    """The adjoint kernels cover the 7/27-point stencils: taping a
    --stencil 0 context is refused."""
    lib = _lib()
    dev = M.Device(lib=lib, nx=4, ny=4, nz=4, num_vars=8, max_blocks=4, stencil=0)
    assert lib.amr_dev_tape(dev.h, 1) == -1   # AMR_ERR_ARG
