"""amr_dev_split_noalloc (GPU): the allocation-free split gives the cells of
amr_dev_split, called through the C ABI, on two calls in a row -- the
second with fewer jobs -- in FP64 (bit-exact) and FP32.  The first call
needs several chunks of the static snapshot area, and later jobs write
children into slots that earlier jobs read as parents (the slot reuse of
ref/block.c:116-127, 304-307)."""
import numpy as np
import pytest

import miniamr_amd as M

pytestmark = pytest.mark.gpu

N, V, SLOTS = 16, 40, 260


def jobs_for(parents, children):
    j = np.zeros(len(children), dtype=M.SPLIT_JOB)
    j["parent"] = np.repeat(parents, 8)[:len(children)]
    j["octant"] = np.tile(np.arange(8), len(parents))[:len(children)]
    j["child"] = children
    return j


@pytest.mark.parametrize("prec", [M.PREC_F64, M.PREC_F32], ids=["f64", "f32"])
def test_split_noalloc_matches_split(prec):
    print("case: split_noalloc prec=%d n=%d vars=%d" % (prec, N, V), flush=True)
    rng = np.random.default_rng(2024 + prec)
    a = M.Device(N, N, N, V, SLOTS, stencil=27, precision=prec)
    b = M.Device(N, N, N, V, SLOTS, stencil=27, precision=prec)
    try:
        for s in range(SLOTS):
            cells = rng.random((V, N + 2, N + 2, N + 2))
            a.upload(s, cells)
            b.upload(s, cells)
        # call 1: 150 jobs (one octant snapshot of 40 FP64 variables is
        # 160 KiB, so the 15 MiB area holds 96 jobs: two chunks); parents
        # 0..18; children 100..239, then slots 0..7 (parents read in the
        # first chunk) and 17, 16 (parents read earlier in the same chunk)
        first = jobs_for(np.arange(19),
                         np.concatenate([np.arange(100, 240), np.arange(8), [17, 16]]))
        # call 2: 9 jobs
        second = jobs_for(np.array([30, 31]), np.array([240, 241, 242, 243, 244, 245, 246,
                                                        247, 20]))
        for jobs in (first, second):
            parent0 = a.download(int(jobs["parent"][0]))
            a.split(jobs)
            b.split_noalloc(jobs)
            touched = np.unique(np.concatenate([jobs["child"], jobs["parent"]]))
            for s in touched:
                x, y = a.download(int(s)), b.download(int(s))
                assert np.array_equal(x, y), "slot %d differs after %d jobs" % (s, len(jobs))
            # the children hold parent / 8 (closed form, FP64 only)
            if prec == M.PREC_F64:
                c = a.download(int(jobs["child"][0]))   # octant 0 of parent0
                h = N // 2
                for di in (0, 1):
                    assert np.array_equal(c[:, 1 + di:N + 1:2, 1:N + 1:2, 2:N + 2:2],
                                          parent0[:, 1:h + 1, 1:h + 1, 1:h + 1] / 8.0)
    finally:
        a.close()
        b.close()


def test_split_noalloc_rejects_what_does_not_fit():
    print("case: split_noalloc limits", flush=True)
    d = M.Device(32, 32, 32, 64, 4, stencil=7)   # 64 x 16^3 x 8 B = 2 MiB per job: fits
    try:
        with pytest.raises(M.AmrError, match="bad job"):
            d.split_noalloc(jobs_for(np.array([0]), np.array([9])))
    finally:
        d.close()
    d = M.Device(64, 64, 64, 8, 2, stencil=7)    # 8 x 32^3 x 8 B = 2 MiB per job: fits
    e = M.Device(64, 64, 64, 64, 2, stencil=7)   # 64 x 32^3 x 8 B = 16 MiB: too large
    try:
        d.split_noalloc(jobs_for(np.array([0]), np.array([1])))
        with pytest.raises(M.AmrError, match="static scratch"):
            e.split_noalloc(jobs_for(np.array([0]), np.array([1])))
    finally:
        d.close()
        e.close()
