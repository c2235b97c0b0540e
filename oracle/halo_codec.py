"""CPU restatement of the reference's halo message formats (test
infrastructure only; nothing in the product imports it).

pack_face / unpack_face of ref/comm.c:245-990 and :993-1461, with the
message lengths of ref/init.c:77-117.  A message for face case c carries,
for each variable m = start .. start+num_comm-1 in turn, the values of one
face plane of the block; the reference writes the same loops three times,
once per direction, with the plane's two in-plane axes renamed.  Here a
direction's plane is viewed as a 2-D array P[a][b] whose axes are the
reference's loop order:

    dir 0 (x, W/E faces):  P[a][b] = array[m][i][a][b]   (a = j, b = k)
    dir 1 (y, S/N faces):  P[a][b] = array[m][a][j][b]   (a = i, b = k)
    dir 2 (z, D/U faces):  P[a][b] = array[m][a][b][k]   (a = i, b = j)

pack reads plane 1 (W/S/D, case < 10) or n (E/N/U, case >= 10) of the
sending block; unpack writes ghost plane 0 (from W/S/D, case < 10) or n+1
(case >= 10) of the receiving block.  Case c (after removing 10):
0 whole face, 1 whole face of the 27-point stencil, 2-5 whole face <->
quarter face (the sender is the fine block; quarter c-2), 6-9 quarter face
<-> whole face (the sender is the coarse block; quarter c-6).  Sums are
formed left to right in the reference's term order; division by 4.0 as in
the reference.  The functions take cells as a float64 array
[num_vars][nx+2][ny+2][nz+2] and follow the cited lines; every branch is
pinned by tests/golden/codec.npz (the reference's own pack_face and
unpack_face, driven by oracle/ref_codec.c).
"""
import numpy as np


def msg_len(code, dir, n, case, recv=False):
    """Values per variable of a message (ref/init.c:77-117) for the
    sender's face case, or with recv=True the receiver's (the case
    selection of ref/comm.c:116-125 and :76-85).  n = (nx, ny, nz)."""
    nx, ny, nz = n
    hx, hy, hz = nx // 2, ny // 2, nz // 2
    c = case - 10 if case >= 10 else case
    if code == 0:
        tab = [[ny * nz, ny * nz, hy * hz, hy * hz],
               [nx * nz, (nx + 2) * nz, hx * hz, hx * hz],
               [nx * ny, (nx + 2) * (ny + 2), hx * hy, hx * hy]]
    elif code == 1:
        tab = [[(ny + 2) * (nz + 2)] * 2 + [(hy + 1) * (hz + 1), (hy + 2) * (hz + 2)],
               [(nx + 2) * (nz + 2)] * 2 + [(hx + 1) * (hz + 1), (hx + 2) * (hz + 2)],
               [(nx + 2) * (ny + 2)] * 2 + [(hx + 1) * (hy + 1), (hx + 2) * (hy + 2)]]
    else:
        tab = [[(ny + 2) * (nz + 2)] * 2 + [(hy + 1) * (hz + 1), (ny + 2) * (nz + 2)],
               [(nx + 2) * (nz + 2)] * 2 + [(hx + 1) * (hz + 1), (nx + 2) * (nz + 2)],
               [(nx + 2) * (ny + 2)] * 2 + [(hx + 1) * (hy + 1), (nx + 2) * (ny + 2)]]
    # a sender's 2-5 send the [2] length and 6-9 the [3] length; a
    # receiver's 2-5 receive [3] and 6-9 receive [2]
    quarter = 2 if (c < 6) != recv else 3
    return tab[dir][0 if c == 0 else 1 if c == 1 else quarter]


def _plane(cells, m, dir, idx):
    if dir == 0:
        return cells[m, idx, :, :]
    if dir == 1:
        return cells[m, :, idx, :]
    return cells[m, :, :, idx]


def _extents(cells, dir):
    nx, ny, nz = (s - 2 for s in cells.shape[1:])
    n = (nx, ny, nz)
    return n, (ny, nz) if dir == 0 else (nx, nz) if dir == 1 else (nx, ny)


def pack_face(cells, code, dir, case, start, num_comm):
    """ref/comm.c:245-990: the send buffer of one face."""
    n, (na, nb) = _extents(cells, dir)
    ha, hb = na // 2, nb // 2
    c = case - 10 if case >= 10 else case
    fixed = n[dir] if case >= 10 else 1
    out = []
    for m in range(start, start + num_comm):
        P = _plane(cells, m, dir, fixed)
        if code == 0:
            if c < 2:
                # x: both cases interior (:264-269); y/z case 1 adds the
                # ring along a (y) or along a and b (z) (:312-321, :358-367)
                ra = range(0, na + 2) if (c == 1 and dir > 0) else range(1, na + 1)
                rb = range(0, nb + 2) if (c == 1 and dir == 2) else range(1, nb + 1)
                out += [P[a, b] for a in ra for b in rb]
            elif c <= 5:                                   # :270-278
                for a in range(1, na, 2):
                    for b in range(1, nb, 2):
                        out.append(P[a, b] + P[a, b + 1] + P[a + 1, b] + P[a + 1, b + 1])
            else:                                          # :279-299
                # the halves come from the bits of the case number itself
                # (6, 7: lower b half; 8, 9: upper; even: lower a half)
                ra = range(1, ha + 1) if c % 2 == 0 else range(ha + 1, na + 1)
                rb = range(1, hb + 1) if (c // 2) % 2 == 1 else range(hb + 1, nb + 1)
                out += [P[a, b] / 4.0 for a in ra for b in rb]
        elif c < 2:                                        # codes 1, 2: :413-418
            out += [P[a, b] for a in range(na + 2) for b in range(nb + 2)]
        elif c <= 5:                                       # :419-465, :670-718
            lead_b, trail_b = (c // 2) % 2 == 1, (c // 2) % 2 == 0

            def row(a_pair):
                if lead_b:
                    out.append(sum_cells(P, a_pair, (0,)))
                for b in range(1, nb, 2):
                    out.append(sum_cells(P, a_pair, (b, b + 1)))
                if trail_b:
                    out.append(sum_cells(P, a_pair, (nb + 1,)))
            if c % 2 == 0:
                row((0,))
            for a in range(1, na, 2):
                row((a, a + 1))
            if c % 2 == 1:
                row((na + 1,))
        elif code == 1:                                    # :466-486
            ra = range(0, ha + 2) if c % 2 == 0 else range(ha, na + 2)
            rb = range(0, hb + 2) if (c // 2) % 2 == 1 else range(hb, nb + 2)
            out += [P[a, b] / 4.0 for a in ra for b in rb]
        else:                                              # code 2: :719-763
            a0, a1 = (1, ha) if c % 2 == 0 else (ha + 1, na)
            b0, b1 = (1, hb) if (c // 2) % 2 == 1 else (hb + 1, nb)

            def row(a):
                out.append(P[a, b0 - 1] / 4.0)
                for b in range(b0, b1 + 1):
                    v = P[a, b] / 4.0
                    out.extend((v, v))
                out.append(P[a, b1 + 1] / 4.0)
            row(a0 - 1)
            for a in range(a0, a1 + 1):
                row(a)
                row(a)
            row(a1 + 1)
    return np.array(out, dtype=np.float64)


def sum_cells(P, ra, rb):
    """Left-to-right sum over ra (outer) x rb (inner), the reference's
    term order (e.g. [j][k] + [j][k+1] + [j+1][k] + [j+1][k+1])."""
    s = None
    for a in ra:
        for b in rb:
            s = P[a, b] if s is None else s + P[a, b]
    return s


def unpack_face(cells, buf, code, dir, case, start, num_comm):
    """ref/comm.c:993-1461: write one received face into cells (in place)."""
    n, (na, nb) = _extents(cells, dir)
    ha, hb = na // 2, nb // 2
    c = case - 10 if case >= 10 else case
    fixed = n[dir] + 1 if case >= 10 else 0
    it = iter(buf)
    for m in range(start, start + num_comm):
        P = _plane(cells, m, dir, fixed)
        if code == 0:
            if c < 2:                                      # :1015-1019, :1056-1065
                ra = range(0, na + 2) if (c == 1 and dir > 0) else range(1, na + 1)
                rb = range(0, nb + 2) if (c == 1 and dir == 2) else range(1, nb + 1)
                for a in ra:
                    for b in rb:
                        P[a, b] = next(it)
            elif c <= 5:                                   # :1020-1028
                for a in range(1, na, 2):
                    for b in range(1, nb, 2):
                        P[a:a + 2, b:b + 2] = next(it)
            else:                                          # :1029-1048
                ra = range(1, ha + 1) if c % 2 == 0 else range(ha + 1, na + 1)
                rb = range(1, hb + 1) if (c // 2) % 2 == 1 else range(hb + 1, nb + 1)
                for a in ra:
                    for b in rb:
                        P[a, b] = next(it)
        elif c < 2 or (code == 2 and c <= 5):              # :1162-1166, :1365-1369
            for a in range(na + 2):
                for b in range(nb + 2):
                    P[a, b] = next(it)
        elif c <= 5:                                       # code 1: :1167-1198
            groups_a = [(0,)] + [(a, a + 1) for a in range(1, na, 2)] + [(na + 1,)]
            groups_b = [(0,)] + [(b, b + 1) for b in range(1, nb, 2)] + [(nb + 1,)]
            for ga in groups_a:
                for gb in groups_b:
                    v = next(it)
                    for a in ga:
                        for b in gb:
                            P[a, b] = v
        else:                                              # codes 1, 2: :1199-1218
            ra = range(0, ha + 1) if c % 2 == 0 else range(ha + 1, na + 2)
            rb = range(0, hb + 1) if (c // 2) % 2 == 1 else range(hb + 1, nb + 2)
            for a in ra:
                for b in rb:
                    P[a, b] = next(it)
    return cells
