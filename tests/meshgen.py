"""Synthetic inputs for kernel-level parity tests.

A face-job plan built here obeys the two rules a real exchange direction
obeys (see ref/comm.c:159-203): no cell is written by two jobs, and no job
reads a cell another job of the same direction writes.  Destination slots
and source slots are disjoint, each destination face gets one fill
(SAME/PROLONG/BC) or four quarter restrictions, so any implementation
that applies the jobs correctly -- in any order -- gives one answer.
"""
import numpy as np

from miniamr_amd import FACE_JOB, SPLIT_JOB, MERGE_JOB, FACE_SAME, FACE_PROLONG, \
    FACE_RESTRICT, FACE_BC, FACE_GHOSTS


def random_plan(rng, n_src, n_dst, kinds=(FACE_SAME, FACE_PROLONG, FACE_RESTRICT, FACE_BC),
                ghosts=False):
    """Jobs for 3 directions; sources are slots [0, n_src), destinations
    [n_src, n_src + n_dst).  ghosts: every job moves whole ghost planes
    (the --code 1/2 kinds)."""
    by_dir = []
    for d in range(3):
        jobs = []
        for dst in range(n_src, n_src + n_dst):
            for face in (2 * d, 2 * d + 1):
                kind = int(rng.choice(kinds))
                if kind == FACE_RESTRICT:
                    for iq in range(2):
                        for jq in range(2):
                            jobs.append((kind, face, dst, int(rng.integers(0, n_src)), iq, jq))
                elif kind == FACE_BC:
                    jobs.append((kind, face, dst, dst, 0, 0))
                else:
                    jobs.append((kind, face, dst, int(rng.integers(0, n_src)),
                                 int(rng.integers(0, 2)), int(rng.integers(0, 2))))
        if ghosts:
            jobs = [(j[0] | FACE_GHOSTS,) + tuple(j[1:]) for j in jobs]
        rng.shuffle(jobs)
        by_dir.append(np.array(jobs, dtype=FACE_JOB))
    return by_dir


def random_cells(rng, shape, n):
    """n random slots of cells in [0, 1) with random (non-zero) ghosts."""
    return [rng.random(shape) for _ in range(n)]


def split_jobs_with_reuse(rng, n_parents):
    """Split jobs as the reference's allocator can produce them: the slot of
    parent p (freed after its 8 children are written) is reused by a child
    of a later parent.  Parents live in slots [0, n_parents)."""
    free = list(range(n_parents, n_parents + 8))
    next_new = n_parents + 8
    jobs = []
    for p in range(n_parents):
        kids = [free.pop(0) for _ in range(8)]
        for o, c in enumerate(kids):
            jobs.append((p, c, o, 0))
        # the parent's slot is now the lowest free slot; top up to 8 free
        free = [p] + list(range(next_new, next_new + 7))
        next_new += 7
    return np.array(jobs, dtype=SPLIT_JOB)


def split_slots_needed(n_parents):
    return n_parents + 8 + 7 * n_parents


def merge_jobs_with_reuse(rng, n_merges):
    """Consolidation jobs where each merged block lands in the first child
    slot of the previous merge (freed by it)."""
    jobs = []
    next_free = 8 * n_merges
    first = next_free
    for m in range(n_merges):
        kids = list(range(8 * m, 8 * m + 8))
        dst = first if m == 0 else 8 * (m - 1)
        for o, c in enumerate(kids):
            jobs.append((dst, c, o, 0))
    return np.array(jobs, dtype=MERGE_JOB)
