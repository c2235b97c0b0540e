#!/usr/bin/env python3
"""Golden vectors of the reference's halo message formats (test data only).

Runs ``oracle/_ref/ref_codec.x`` -- the reference's own pack_face and
unpack_face (ref/comm.c:245-1461, linked unmodified from its comm.o),
driven by oracle/ref_codec.c -- for every --code on a cubic and a
non-cubic block, and writes ``codec.npz``: per (shape, code) the raw
harness output as a uint8 array ``<nx>x<ny>x<nz>_c<code>`` (format in
oracle/ref_codec.c).  Build the harness with ``make -C oracle ref``.
Usage: ``python tests/golden/make_codec_golden.py``.
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
EXE = os.path.join(REPO, "oracle", "_ref", "ref_codec.x")
SHAPES = [(4, 4, 4), (6, 4, 8)]


def main():
    if not os.path.exists(EXE):
        sys.exit("build the harness first: make -C oracle ref")
    out = {}
    with tempfile.TemporaryDirectory() as td:
        for shape in SHAPES:
            for code in (0, 1, 2):
                path = os.path.join(td, "c.bin")
                subprocess.run([EXE] + [str(x) for x in shape] + [str(code), path], check=True)
                with open(path, "rb") as f:
                    out["%dx%dx%d_c%d" % (shape + (code,))] = np.frombuffer(f.read(), np.uint8)
    np.savez_compressed(os.path.join(HERE, "codec.npz"), **out)
    print("%d cases, %d bytes" % (len(out), sum(v.size for v in out.values())))


if __name__ == "__main__":
    main()
