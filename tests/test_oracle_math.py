"""The oracle's arithmetic against what the mathematics fixes (CPU).

These checks pin oracle/amr_cpu.c independently of the golden vectors,
through the same C ABI the product implements (liboracle_dev.so):
closed forms of each exchange kind, stencil invariants, the split/merge
round trip and the slot-reuse semantics of a batch.  Each would fail for
a plausible slip (wrong plane, swapped quarter axes, transposed in-plane
indices, dropped term, /4 on the wrong side).
"""
import numpy as np
import pytest

import amrcheck as A
import meshgen
from conftest import require
from miniamr_amd import Device, load_library, FACE_JOB, FACE_SAME, FACE_PROLONG, \
    FACE_RESTRICT, FACE_BC, FACE_GHOSTS


@pytest.fixture(scope="module")
def olib():
    require(A.ORACLE_SO, "run __graft_entry__.build() or `make -C oracle oracle`")
    return load_library(A.ORACLE_SO)


def dev(olib, nx=4, ny=6, nz=8, nv=2, mb=16, st=7):
    return Device(nx, ny, nz, nv, mb, stencil=st, lib=olib)


def plane(a, d, p):
    """Plane p of axis d of a (nx+2, ny+2, nz+2) array."""
    return a[(slice(None),) * d + (p,)]


@pytest.mark.parametrize("st", [7, 27])
def test_same_level_copy_closed_form(olib, st):
    rng = np.random.default_rng(1)
    d = dev(olib, st=st)
    cells = [rng.random(d.shape) for _ in range(3)]
    for s, c in enumerate(cells):
        d.upload(s, c)
    # slot 2 receives on every face from slot 0 (low faces) / slot 1 (high faces)
    jobs = [np.array([(FACE_SAME, 2 * k, 2, 0, 0, 0), (FACE_SAME, 2 * k + 1, 2, 1, 0, 0)],
                     dtype=FACE_JOB) for k in range(3)]
    d.set_mesh([0, 1, 2], jobs)
    d.comm(0, 2)
    out = d.download(2)
    n = (4, 6, 8)
    for v in range(2):
        for ax in range(3):
            lo = plane(out[v], ax, 0)
            hi = plane(out[v], ax, n[ax] + 1)
            # interior part of the ghost plane: W ghost = neighbour's last plane
            inner = tuple(slice(1, n[b] + 1) for b in range(3) if b != ax)
            np.testing.assert_array_equal(lo[inner], plane(cells[0][v], ax, n[ax])[inner])
            np.testing.assert_array_equal(hi[inner], plane(cells[1][v], ax, 1)[inner])


def test_prolong_restrict_closed_form(olib):
    rng = np.random.default_rng(2)
    nx, ny, nz = 4, 6, 8
    d = dev(olib, nx, ny, nz)
    coarse, fine = rng.random(d.shape), rng.random(d.shape)
    d.upload(0, coarse)
    d.upload(1, fine)
    # fine block 1 is west of coarse block 0, covering quarter (iq=1, jq=0)
    # of the coarse block's west face: fine's E ghost <- coarse/4 and
    # coarse's W ghost quarter <- sums of fine cells.
    jobs = [np.array([(FACE_PROLONG, 1, 1, 0, 1, 0), (FACE_RESTRICT, 0, 0, 1, 1, 0)],
                     dtype=FACE_JOB), np.zeros(0, FACE_JOB), np.zeros(0, FACE_JOB)]
    d.set_mesh([0, 1], jobs)
    d.comm(0, 2)
    c_out, f_out = d.download(0), d.download(1)
    hy, hz = ny // 2, nz // 2
    for v in range(2):
        # ref/comm.c:1618-1625 with j1 = jq*hy = 0, k1 = iq*hz = hz
        for j in range(1, hy + 1):
            for k in range(1, hz + 1):
                q = coarse[v, 1, j, k + hz] / 4.0
                for jj in (2 * j - 1, 2 * j):
                    for kk in (2 * k - 1, 2 * k):
                        assert f_out[v, nx + 1, jj, kk] == q
                s = ((fine[v, nx, 2 * j - 1, 2 * k - 1] + fine[v, nx, 2 * j - 1, 2 * k])
                     + fine[v, nx, 2 * j, 2 * k - 1]) + fine[v, nx, 2 * j, 2 * k]
                assert c_out[v, 0, j, k + hz] == s
        # the other three quarters and the edges are untouched
        untouched = np.ones((ny + 2, nz + 2), bool)
        untouched[1:hy + 1, hz + 1:nz + 1] = False
        np.testing.assert_array_equal(c_out[v, 0][untouched], coarse[v, 0][untouched])


def test_whole_plane_coarse_fine_closed_forms(olib):
    """--code 1/2 coarse/fine jobs over whole ghost planes: a coarse block
    (slot 0) with four fine blocks (slots 1-4) on its west face.  Constant
    fields: every fine ghost cell, ring included, is coarse / 4; the coarse
    ghost plane is covered exactly once by the four quarters, with 4 fine
    cells per interior cell, 2 per ring edge cell and 1 per corner.  A field
    linear in y pins which fine rows pair up."""
    nx, ny, nz = 4, 6, 8
    d = dev(olib, nx, ny, nz, nv=2, st=7)
    quarters = [(iq, jq) for jq in range(2) for iq in range(2)]
    pro = [(FACE_PROLONG | FACE_GHOSTS, 1, 1 + q, 0, iq, jq) for q, (iq, jq) in enumerate(quarters)]
    res = [(FACE_RESTRICT | FACE_GHOSTS, 0, 0, 1 + q, iq, jq) for q, (iq, jq) in enumerate(quarters)]
    jobs = [np.array(pro + res, dtype=FACE_JOB), np.zeros(0, FACE_JOB), np.zeros(0, FACE_JOB)]
    d.set_mesh([0, 1, 2, 3, 4], jobs)
    d.upload(0, np.full(d.shape, 8.0))
    for s in range(1, 5):
        d.upload(s, np.full(d.shape, 1.0))
    d.comm(0, 2)
    for s in range(1, 5):
        np.testing.assert_array_equal(d.download(s)[:, nx + 1], 2.0)
    g = d.download(0)[:, 0]
    want = np.full((ny + 2, nz + 2), 4.0)
    want[[0, -1], :] = 2.0
    want[:, [0, -1]] = 2.0
    want[0, 0] = want[0, -1] = want[-1, 0] = want[-1, -1] = 1.0
    for v in range(2):
        np.testing.assert_array_equal(g[v], want)
    # fine cells = y index j: an interior coarse row j' of a quarter sums
    # fine rows 2j'-1 and 2j' twice (8j' - 2); the ring rows hold 2*0 and
    # 2*(ny+1); ring columns 2j'-1 + 2j'
    f = np.broadcast_to(np.arange(ny + 2, dtype=float)[None, None, :, None], d.shape).copy()
    for s in range(1, 5):
        d.upload(s, f)
    d.comm(0, 2)
    g = d.download(0)[0, 0]
    h = ny // 2
    for j in range(1, ny + 1):
        jp = j - (h if j > h else 0)
        np.testing.assert_array_equal(g[j, 1:-1], 8.0 * jp - 2.0)
        assert g[j, 0] == g[j, -1] == 4.0 * jp - 1.0
    np.testing.assert_array_equal(g[0, 1:-1], 0.0)
    np.testing.assert_array_equal(g[-1, 1:-1], 2.0 * (ny + 1))


@pytest.mark.parametrize("st", [7, 27])
def test_bc_reflects(olib, st):
    rng = np.random.default_rng(3)
    nx, ny, nz = 4, 6, 8
    d = dev(olib, nx, ny, nz, st=st)
    c = rng.random(d.shape)
    d.upload(0, c)
    jobs = [np.array([(FACE_BC, 2 * k, 0, 0, 0, 0), (FACE_BC, 2 * k + 1, 0, 0, 0, 0)],
                     dtype=FACE_JOB) for k in range(3)]
    d.set_mesh([0], jobs)
    d.comm(0, 2)
    o = d.download(0)
    # 7-point: face interiors only; checked on the x faces (first phase)
    np.testing.assert_array_equal(o[:, 0, 1:-1, 1:-1], c[:, 1, 1:-1, 1:-1])
    np.testing.assert_array_equal(o[:, nx + 1, 1:-1, 1:-1], c[:, nx, 1:-1, 1:-1])
    if st == 27:
        # full extent after all three phases: every ghost equals the nearest
        # interior cell (reflection composes to clamping)
        idx = np.ix_(np.clip(np.arange(nx + 2), 1, nx), np.clip(np.arange(ny + 2), 1, ny),
                     np.clip(np.arange(nz + 2), 1, nz))
        for v in range(2):
            np.testing.assert_array_equal(o[v], c[v][idx])


@pytest.mark.parametrize("st", [7, 27])
def test_stencil_constant_and_brute_force(olib, st):
    rng = np.random.default_rng(4)
    nx, ny, nz = 4, 6, 8
    d = dev(olib, nx, ny, nz, st=st)
    const = np.full(d.shape, 0.5)
    d.upload(0, const)
    c = rng.random(d.shape)
    d.upload(1, c)
    d.set_mesh([0, 1], [np.zeros(0, FACE_JOB)] * 3)
    d.stencil(0)
    d.stencil(1)
    np.testing.assert_array_equal(d.download(0), const)
    out = d.download(1)
    # brute force in float64 with the reference's association
    for v in (0,):
        for i in range(1, nx + 1):
            for j in range(1, ny + 1):
                for k in range(1, nz + 1):
                    if st == 7:
                        a = c[v]
                        w = (((((a[i - 1, j, k] + a[i, j - 1, k]) + a[i, j, k - 1]) + a[i, j, k])
                              + a[i, j, k + 1]) + a[i, j + 1, k]) + a[i + 1, j, k]
                        w = w / 7.0
                    else:
                        s = []
                        for p in (i - 1, i, i + 1):
                            acc = c[v, p, j - 1, k - 1]
                            for (jj, kk) in [(j - 1, k), (j - 1, k + 1), (j, k - 1), (j, k),
                                             (j, k + 1), (j + 1, k - 1), (j + 1, k),
                                             (j + 1, k + 1)]:
                                acc = acc + c[v, p, jj, kk]
                            s.append(acc)
                        w = ((s[0] + s[1]) + s[2]) / 27.0
                    assert out[v, i, j, k] == w
    # ghosts untouched
    mask = np.ones(d.shape[1:], bool)
    mask[1:-1, 1:-1, 1:-1] = False
    np.testing.assert_array_equal(out[:, mask], c[:, mask])


def test_checksum_is_interior_sum(olib):
    rng = np.random.default_rng(5)
    d = dev(olib)
    c = [rng.random(d.shape) for _ in range(3)]
    for s in range(3):
        d.upload(s, c[s])
    d.set_mesh([2, 0, 1], [np.zeros(0, FACE_JOB)] * 3)
    for v in range(2):
        ref = sum(float(np.sum(c[s][v, 1:-1, 1:-1, 1:-1])) for s in (2, 0, 1))
        assert abs(d.check_sum(v) - ref) <= 1e-13 * ref


def test_split_consolidate_round_trip_and_reuse(olib):
    rng = np.random.default_rng(6)
    nparents = 3
    mb = meshgen.split_slots_needed(nparents)
    d = dev(olib, mb=mb)
    parents = [rng.random(d.shape) for _ in range(nparents)]
    for p in range(nparents):
        d.upload(p, parents[p])
    jobs = meshgen.split_jobs_with_reuse(rng, nparents)
    assert any(int(t["child"]) < nparents for t in jobs)   # a parent slot is reused
    d.split(jobs)
    nx, ny, nz = 4, 6, 8
    # sequential semantics: the last job writing a slot decides its interior
    expect = {}
    for t in jobs:
        p, ch, o = int(t["parent"]), int(t["child"]), int(t["octant"])
        i1, j1, k1 = (o % 2) * nx // 2, ((o // 2) % 2) * ny // 2, (o // 4) * nz // 2
        ii = (np.arange(1, nx + 1) + 1) // 2 + i1
        jj = (np.arange(1, ny + 1) + 1) // 2 + j1
        kk = (np.arange(1, nz + 1) + 1) // 2 + k1
        expect[ch] = parents[p][:, ii][:, :, jj][:, :, :, kk] / 8.0
    for ch, e in expect.items():
        np.testing.assert_array_equal(d.download(ch)[:, 1:-1, 1:-1, 1:-1], e)
    # merge the last parent's children back: equals the parent up to rounding
    last = [t for t in jobs if int(t["parent"]) == nparents - 1]
    dst = mb - 1
    merge = np.array([(dst, int(t["child"]), int(t["octant"]), 0) for t in last],
                     dtype=meshgen.MERGE_JOB)
    d.consolidate(merge)
    back = d.download(dst)
    p = parents[nparents - 1]
    np.testing.assert_allclose(back[:, 1:-1, 1:-1, 1:-1], p[:, 1:-1, 1:-1, 1:-1],
                               rtol=1e-15, atol=0)
