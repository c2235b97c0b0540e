"""Reader of tests/golden/codec.npz: the reference's pack_face /
unpack_face outputs written by oracle/ref_codec.c, and the inputs it drew
(the same LCG, regenerated here)."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "codec.npz")
MAXBUF = 4096
MASK = (1 << 64) - 1


def _draws(count, state):
    out = np.empty(count, dtype=np.float64)
    x = state[0]
    for t in range(count):
        x = (x * 6364136223846793005 + 1442695040888963407) & MASK
        out[t] = (x >> 11) * (1.0 / 9007199254740992.0)
    state[0] = x
    return out


def fnv1a(cells):
    h = 1469598103934665603
    for byte in np.ascontiguousarray(cells, dtype="<f8").tobytes():
        h = ((h ^ byte) * 1099511628211) & MASK
    return h


def load(key):
    """One (shape, code) record: header fields, the two blocks' initial
    cells, and per (dir, case) the unpack input, the packed message, the
    unpacked ghost plane and the whole-array hash."""
    raw = np.load(GOLDEN)[key].tobytes()
    hdr = np.frombuffer(raw, "<i4", 8)
    assert hdr[0] == 0x414d4331
    nx, ny, nz, code, nv, start, num = (int(v) for v in hdr[1:])
    g = (nv, nx + 2, ny + 2, nz + 2)
    state = [12345]
    blocks = [_draws(int(np.prod(g)), state).reshape(g) for _ in range(2)]
    off = 32
    cases = []
    for dir in range(3):
        plane = (g[2] * g[3]) if dir == 0 else (g[1] * g[3]) if dir == 1 else (g[1] * g[2])
        for c in range(20):
            L = int(np.frombuffer(raw, "<i4", 1, off)[0])
            off += 4
            packed = np.frombuffer(raw, "<f8", L, off).copy()
            off += 8 * L
            pl = np.frombuffer(raw, "<f8", num * plane, off).copy()
            off += 8 * num * plane
            h = int(np.frombuffer(raw, "<u8", 1, off)[0])
            off += 8
            cases.append({"dir": dir, "case": c, "unpack_in": _draws(MAXBUF, state),
                          "packed": packed, "plane": pl, "hash": h})
    assert off == len(raw)
    return {"n": (nx, ny, nz), "code": code, "num_vars": nv, "start": start, "num_comm": num,
            "blocks": blocks, "cases": cases}


def ghost_plane(cells, dir, case, start, num):
    """The unpacked side's ghost plane, as the harness writes it."""
    n = cells.shape[1 + dir] - 2
    q = n + 1 if case >= 10 else 0
    sl = [slice(start, start + num), slice(None), slice(None), slice(None)]
    sl[1 + dir] = q
    return cells[tuple(sl)].reshape(-1)


KEYS = ["4x4x4_c0", "4x4x4_c1", "4x4x4_c2", "6x4x8_c0", "6x4x8_c1", "6x4x8_c2"]
