/*
 * amr_adjoint.h -- reverse mode of the miniAMR hot path (the
 * differentiation form; no counterpart in the reference, which has no
 * derivatives).
 *
 * Tape.  With taping on, a device context records every call that changes
 * cell values, in order: block uploads, mesh plans (amr_dev_set_mesh),
 * split and consolidate job lists, exchanges, stencils and stages.  The
 * tape can be replayed forward from any saved cell state (the runs of a
 * forward pass with other initial values, e.g. for finite differences) and
 * walked backward by the adjoint kernels.
 *
 * Adjoint.  Every recorded operation is linear in the cell values, so its
 * adjoint needs no forward values, only the mesh plan and job lists the
 * tape holds.  An adjoint context carries one value per element of the
 * forward pool (the same record layout and storage precision, so the
 * adjoint of each element -- interior cell, ghost cell or z-plane copy -- is
 * exactly the transpose of the forward program at element level, stale
 * ghosts included).  The adjoint kernels are:
 *   stencil      the transposed 7/27-point average: each cell receives
 *                lambda(o)/7 (or /27) from every output o whose stencil
 *                reads it; interior cells are overwritten by the forward
 *                stencil, ghost cells accumulate;
 *   exchange     per direction, in reverse order: a ghost cell's adjoint
 *                moves to the cells it was filled from -- copy for
 *                same-level and boundary faces, prolongation's adjoint
 *                (sum of the 4 fine ghosts / 4) onto the coarse cell, and
 *                restriction's adjoint (the coarse ghost's adjoint to each
 *                of its 4 fine cells): prolongation and restriction swap;
 *   split        each parent cell receives the sum of its 8 child cells'
 *                adjoints / 8 (a restriction);
 *   consolidate  each child cell receives its merged cell's adjoint.
 * Uploads are not reversed: the adjoint of an uploaded block is read with
 * amr_adj_get_block at that tape position.
 *
 * State buffers (`ptr`) hold amr_dev_state_bytes bytes (the whole pool,
 * elements in the storage precision) and may be device memory of the
 * context's GPU or host memory.
 */
#ifndef AMR_ADJOINT_H
#define AMR_ADJOINT_H

#include <stddef.h>
#include <stdint.h>
#include "amr_dev.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Tape operation kinds (op[0] of amr_dev_tape_op). */
enum {
   AMR_OP_UPLOAD = 1,    /* op[1] slot */
   AMR_OP_SET_MESH = 2,  /* op[1] plan index */
   AMR_OP_SPLIT = 3,     /* op[1] first job, op[2] job count (split job list) */
   AMR_OP_MERGE = 4,     /* op[1] first job, op[2] job count (merge job list) */
   AMR_OP_COMM = 5,      /* op[1] start, op[2] num, op[3] stage, op[4] permute */
   AMR_OP_STENCIL = 6,   /* op[1] var */
   AMR_OP_STAGE = 7      /* op[1] start, op[2] num, op[3] stage, op[4] permute */
};
/* op[7] of every operation: the plan index in force (-1 before the first
 * amr_dev_set_mesh). */

/* Start (on = 1) or stop (on = 0) recording; starting clears the tape. */
int amr_dev_tape(amr_dev *dev, int on);
int amr_dev_tape_length(const amr_dev *dev, int *n_ops);
int amr_dev_tape_op(const amr_dev *dev, int i, int32_t op[8]);
/* Plan `plan`: pass NULL arrays to query *n_active and count[3] first. */
int amr_dev_tape_plan(const amr_dev *dev, int plan, int32_t *n_active, int32_t *sorted,
                      int32_t count[3], amr_face_job *jobs);
/* Job lists of a SPLIT / MERGE operation i (op[2] entries). */
int amr_dev_tape_split_jobs(const amr_dev *dev, int i, amr_split_job *jobs);
int amr_dev_tape_merge_jobs(const amr_dev *dev, int i, amr_merge_job *jobs);

/* Re-run recorded operations [op0, op1) forward on the current cells
 * (uploads are skipped; checksums are not recomputed).  Not recorded. */
int amr_dev_replay(amr_dev *dev, int op0, int op1);

/* The whole cell pool, for checkpoints. */
int amr_dev_state_bytes(const amr_dev *dev, size_t *bytes);
int amr_dev_save_state(amr_dev *dev, void *ptr);
int amr_dev_load_state(amr_dev *dev, const void *ptr);
/* Set every element of one slot (interior, ghosts and copies) to zero. */
int amr_dev_clear_block(amr_dev *dev, int slot);

typedef struct amr_adj amr_adj;

/* Adjoint context of `dev` (same layout and precision), all zero. */
int amr_adj_create(amr_dev *dev, amr_adj **out);
int amr_adj_destroy(amr_adj *adj);
const char *amr_adj_last_error(const amr_adj *adj);
int amr_adj_zero(amr_adj *adj);
int amr_adj_load(amr_adj *adj, const void *ptr);   /* amr_dev_state_bytes */
int amr_adj_save(amr_adj *adj, void *ptr);
/* Adjoint of one slot's cells as [num_vars][nx+2][ny+2][nz+2] doubles: a
 * cell's value is the sum over its elements (its home and, on the z = 1 /
 * z = nz planes, its copy), i.e. the derivative with respect to a value
 * written to that cell (as an upload does).  set writes the home and
 * clears the copy. */
int amr_adj_get_block(amr_adj *adj, int slot, double *cells);
int amr_adj_set_block(amr_adj *adj, int slot, const double *cells);
/* Reverse pass over recorded operations [op0, op1), last first. */
int amr_adj_run(amr_adj *adj, int op0, int op1);
int amr_adj_sync(amr_adj *adj);

#ifdef __cplusplus
}
#endif

#endif
