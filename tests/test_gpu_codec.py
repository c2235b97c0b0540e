"""The halo codec on the GPU (amr_dev_pack_faces / amr_dev_unpack_faces)
against the reference's own pack_face / unpack_face (tests/golden/codec.npz,
written by oracle/ref_codec.c from the reference's comm.o): codes 0-2,
every direction, all 20 face cases, a cubic and a non-cubic block, FP64.

Per direction one pack call carries all 20 messages of block 0 into one
device buffer, and one unpack call writes 20 messages into 20 copies of
block 1 (one per case, since the cases of a side share a ghost plane).
Packed messages must be bit-identical to the reference's and of the
length of ref/init.c:77-117; after unpacking, the ghost plane must be
bit-identical and every other cell unchanged (whole-array hash)."""
import ctypes
import os
import sys

import numpy as np
import pytest
import torch

import codec_golden as G
import miniamr_amd as M

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
import halo_codec as H  # noqa: E402

pytestmark = pytest.mark.gpu


def msgs_array(slots, cases, offsets):
    m = np.zeros(len(slots), dtype=M.FACE_MSG)
    m["slot"], m["face_case"], m["offset"] = slots, cases, offsets
    return m


@pytest.mark.parametrize("key", G.KEYS)
def test_codec_matches_reference_formats(key):
    print("case: codec %s" % key, flush=True)
    r = G.load(key)
    nx, ny, nz = r["n"]
    code, start, num = r["code"], r["start"], r["num_comm"]
    d = M.Device(nx, ny, nz, r["num_vars"], 21, stencil=27)
    try:
        d.upload(0, r["blocks"][0])
        for dir in range(3):
            cases = [c for c in r["cases"] if c["dir"] == dir]
            # pack: all 20 cases of block 0 in one call
            lens = [d.face_msg_len(code, dir, c["case"], unpack=False) for c in cases]
            for c, L in zip(cases, lens):
                assert L == H.msg_len(code, dir, r["n"], c["case"]), (dir, c["case"])
            offs = np.concatenate([[0], np.cumsum([num * L for L in lens])])
            buf = torch.full((int(offs[-1]),), float("nan"), dtype=torch.float64, device="cuda")
            torch.cuda.synchronize()
            d.pack_faces(code, dir, start, num, msgs_array([0] * 20, [c["case"] for c in cases],
                                                           offs[:-1]), buf)
            d.sync()
            got = buf.cpu().numpy()
            for t, c in enumerate(cases):
                assert np.array_equal(got[offs[t]:offs[t + 1]], c["packed"]), (key, dir, c["case"])
            # unpack: case t into slot 1 + t
            ulens = [d.face_msg_len(code, dir, c["case"], unpack=True) for c in cases]
            uoffs = np.concatenate([[0], np.cumsum([num * L for L in ulens])])
            host = np.concatenate([c["unpack_in"][:num * L] for c, L in zip(cases, ulens)])
            ubuf = torch.from_numpy(host).to("cuda")
            torch.cuda.synchronize()
            for t in range(20):
                d.upload(1 + t, r["blocks"][1])
            d.unpack_faces(code, dir, start, num,
                           msgs_array(list(range(1, 21)), [c["case"] for c in cases],
                                      uoffs[:-1]), ubuf)
            d.sync()
            for t, c in enumerate(cases):
                cells = d.download(1 + t)
                assert np.array_equal(G.ghost_plane(cells, dir, c["case"], start, num),
                                      c["plane"]), (key, dir, c["case"])
                assert G.fnv1a(cells) == c["hash"], (key, dir, c["case"])
    finally:
        d.close()


def test_codec_rejects_messages_past_the_buffer():
    print("case: codec bounds", flush=True)
    d = M.Device(4, 4, 4, 2, 2, stencil=27)
    try:
        L = d.face_msg_len(1, 0, 0, unpack=False)
        buf = torch.zeros(L, dtype=torch.float64, device="cuda")
        with pytest.raises(M.AmrError, match="past the buffer"):
            d.pack_faces(1, 0, 0, 2, msgs_array([0], [0], [0]), buf)   # 2 variables
        with pytest.raises(M.AmrError, match="bad message"):
            d.pack_faces(1, 0, 0, 1, msgs_array([5], [0], [0]), buf)   # slot 5 of 2
        d.pack_faces(1, 0, 0, 1, msgs_array([0], [0], [0]), buf)
        d.sync()
    finally:
        d.close()


def test_codec_round_trip_on_a_refined_mesh_face():
    """Closed form on the device: a coarse block's quarter sent with code 0
    (case 6: lower halves of both in-plane axes, value / 4) and unpacked by
    the fine side (case 2) fills each 2x2 group of the fine ghost plane
    with the coarse value / 4."""
    print("case: codec round trip", flush=True)
    rng = np.random.default_rng(5)
    d = M.Device(8, 8, 8, 3, 3, stencil=7)
    try:
        coarse = rng.random((3, 10, 10, 10))
        d.upload(0, coarse)
        d.upload(1, np.zeros((3, 10, 10, 10)))
        L = d.face_msg_len(0, 1, 6, unpack=False)
        buf = torch.zeros(3 * L, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        d.pack_faces(0, 1, 0, 3, msgs_array([0], [6], [0]), buf)
        d.unpack_faces(0, 1, 0, 3, msgs_array([1], [2], [0]), buf)
        d.sync()
        fine = d.download(1)
        expect = coarse[:, 1:5, 1, 1:5] / 4.0          # y face, quarter (lower x, lower z)
        assert np.array_equal(fine[:, 1:9:2, 0, 1:9:2], expect)
        assert np.array_equal(fine[:, 2:10:2, 0, 2:10:2], expect)
    finally:
        d.close()
