"""The halo codec's CPU restatement (oracle/halo_codec.py) against the
reference's own pack_face / unpack_face (tests/golden/codec.npz, from
oracle/ref_codec.c): every direction, every face case 0-19, codes 0-2, a
cubic and a non-cubic block.  Packed messages bit-identical and of the
length of ref/init.c:77-117; after unpacking, the ghost plane bit-identical
and every other cell unchanged (whole-array hash)."""
import os
import sys

import numpy as np
import pytest

import codec_golden as G

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
import halo_codec as H  # noqa: E402


@pytest.mark.parametrize("key", G.KEYS)
def test_oracle_matches_reference_formats(key):
    r = G.load(key)
    code, start, num = r["code"], r["start"], r["num_comm"]
    for c in r["cases"]:
        got = H.pack_face(r["blocks"][0], code, c["dir"], c["case"], start, num)
        assert len(got) == num * H.msg_len(code, c["dir"], r["n"], c["case"]), (c["dir"], c["case"])
        assert np.array_equal(got, c["packed"]), (key, c["dir"], c["case"])
        cells = r["blocks"][1].copy()
        m = num * H.msg_len(code, c["dir"], r["n"], c["case"], recv=True)
        H.unpack_face(cells, c["unpack_in"][:m], code, c["dir"], c["case"], start, num)
        assert np.array_equal(G.ghost_plane(cells, c["dir"], c["case"], start, num), c["plane"])
        assert G.fnv1a(cells) == c["hash"], (key, c["dir"], c["case"])


def test_oracle_unpack_reads_exactly_its_message_length():
    """Each receive-side case consumes exactly msg_len values per variable
    (a shorter buffer fails, the exact one suffices)."""
    r = G.load("6x4x8_c1")
    for c in r["cases"]:
        m = r["num_comm"] * H.msg_len(1, c["dir"], r["n"], c["case"], recv=True)
        H.unpack_face(r["blocks"][1].copy(), c["unpack_in"][:m], 1, c["dir"], c["case"], 1, 2)
        with pytest.raises(StopIteration):
            H.unpack_face(r["blocks"][1].copy(), c["unpack_in"][:m - 1], 1, c["dir"], c["case"],
                          1, 2)


def test_prolong_then_restrict_round_trip():
    """Closed form: a coarse quarter sent (case 6-9, value / 4) and unpacked
    by the fine side (case 2-5) replicates each coarse value / 4 into its
    2x2 fine cells; packing that fine plane back (case 2-5, code 0) sums
    the four copies, giving the coarse cells exactly."""
    rng = np.random.default_rng(3)
    coarse = rng.random((1, 6, 6, 6))
    fine = np.zeros((1, 6, 6, 6))
    for dir in range(3):
        for q in range(4):
            msg = H.pack_face(coarse, 0, dir, 6 + q, 0, 1)
            H.unpack_face(fine, msg, 0, dir, 2 + q, 0, 1)
            back = H.pack_face(np.roll(fine, 1, axis=1 + dir), 0, dir, 2 + q, 0, 1)
            assert np.array_equal(back, msg * 4.0)


@pytest.mark.parametrize("key", G.KEYS)
def test_oracle_library_codec_matches_reference(key):
    """The C oracle library's amr_dev_pack_faces / amr_dev_unpack_faces
    (oracle/amr_cpu.c, host buffers) against the same golden vectors."""
    import amrcheck as A
    import miniamr_amd as M
    from conftest import require
    require(A.ORACLE_SO, "run __graft_entry__.build()")
    r = G.load(key)
    nx, ny, nz = r["n"]
    code, start, num = r["code"], r["start"], r["num_comm"]
    d = M.Device(nx, ny, nz, r["num_vars"], 2, stencil=27, lib=M.load_library(A.ORACLE_SO))
    try:
        d.upload(0, r["blocks"][0])
        for c in r["cases"]:
            L = d.face_msg_len(code, c["dir"], c["case"])
            assert L == H.msg_len(code, c["dir"], r["n"], c["case"])
            out = np.full(num * L, np.nan)
            msg = np.zeros(1, M.FACE_MSG)
            msg["face_case"] = c["case"]
            d.pack_faces(code, c["dir"], start, num, msg, (out.ctypes.data, out.size))
            assert np.array_equal(out, c["packed"]), (key, c["dir"], c["case"])
            d.upload(1, r["blocks"][1])
            U = d.face_msg_len(code, c["dir"], c["case"], unpack=True)
            inp = np.ascontiguousarray(c["unpack_in"][:num * U])
            msg["slot"] = 1
            d.unpack_faces(code, c["dir"], start, num, msg, (inp.ctypes.data, inp.size))
            assert G.fnv1a(d.download(1)) == c["hash"], (key, c["dir"], c["case"])
    finally:
        d.close()
