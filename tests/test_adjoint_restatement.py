"""The torch restatement of a recorded run (oracle/adjoint_torch.py), the
checker of the adjoint kernels, pinned on the CPU against the C oracle:
the same synthetic sequence of uploads, exchanges with every rule and
both direction orders, stencils, a split and a consolidation with slot
reuse must give bit-identical cells (ghosts included).  Its gradients are
torch's own; a finite-difference probe confirms the map is differentiated
where it is defined."""
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


def restated(t):
    c = t["cfg"]
    return AT.Restatement(t["ops"], t["plans"], t["jobs"], (c["nx"], c["ny"], c["nz"]),
                          c["stencil"], c["num_vars"], c["max_blocks"])


@pytest.mark.parametrize("ghosts", [False, True], ids=["code0", "code1"])
@pytest.mark.parametrize("stencil", [7, 27])
def test_restatement_matches_c_oracle(stencil, ghosts):
    require(A.ORACLE_SO, "run __graft_entry__.build()")
    t = tapegen.build(5 + stencil, stencil=stencil, ghosts=ghosts)
    dev = M.Device(lib=M.load_library(A.ORACLE_SO), **t["cfg"])
    tapegen.apply(dev, t)
    u, _ = restated(t).run(torch.from_numpy(t["u0"]), t["uploads"])
    for s in range(t["cfg"]["max_blocks"]):
        got = u[:, s].numpy()
        want = dev.download(s)
        assert A.ulp_diff(got, want).max() == 0, "slot %d" % s


def test_restatement_gradient_is_the_linear_map():
    t = tapegen.build(11, n=(4, 4, 4), V=2, stencil=27)
    R = restated(t)
    u0 = torch.from_numpy(t["u0"]).requires_grad_(True)
    w = torch.from_numpy(np.random.default_rng(3).standard_normal(
        (t["cfg"]["num_vars"], t["cfg"]["max_blocks"]) + t["u0"].shape[2:]))
    u, _ = R.run(u0, t["uploads"])
    J = (u * w).sum()
    (g,) = torch.autograd.grad(J, u0)
    rng = np.random.default_rng(4)
    for _ in range(5):
        idx = tuple(int(rng.integers(0, s)) for s in t["u0"].shape)
        e = torch.zeros_like(u0)
        e[idx] = 1.0
        with torch.no_grad():
            jp = (R.run(u0 + e, t["uploads"])[0] * w).sum()
            jm = (R.run(u0 - e, t["uploads"])[0] * w).sum()
        assert abs((jp - jm).item() / 2 - g[idx].item()) <= 1e-10 * max(1.0, abs(g[idx].item()))
