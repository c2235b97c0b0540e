/*
 * amr_run.h -- driver-level C API of the miniAMR drop-in (one rank, one GPU).
 *
 * The reference's program is main() -> driver() (ref/main.c:38-362,
 * ref/driver.c:37-165).  This API exposes the same run, step by step, to
 * callers that want to time or drive it (the benchmark, Python via ctypes):
 *
 *   amr_run_create   parse the reference's command line, validate it
 *                    (ref/main.c:657-723), allocate, build the initial mesh
 *                    and values and run the initial refinement
 *                    (ref/driver.c:45-56);
 *   amr_run_step     one timestep: all stages (exchange, stencil, checksum
 *                    and conservation test), then object motion and, every
 *                    refine_freq timesteps, refinement (ref/driver.c:67-160);
 *   amr_run_finish   end-of-run bookkeeping; amr_run_report prints the
 *                    Summary line (ref/profile.c:1298);
 *   amr_run_destroy  free everything.
 *
 * Errors: the reference ends the process on an invalid command line or a
 * fatal condition (printf + exit(-1), ref/main.c:657-723, ref/block.c:72-75,
 * ref/comm.c:199-200).  The drop-in executable keeps that; these library
 * calls print the same message and return instead: amr_run_create returns
 * NULL, amr_run_step / amr_run_finish return -1, and amr_run_last_error
 * gives the message (per thread).  After a failed step only
 * amr_run_destroy may be called.
 */
#ifndef AMR_RUN_H
#define AMR_RUN_H

#include <stdint.h>
#include "amr_dev.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct amr_run amr_run;

typedef struct {
   int64_t timesteps_done;      /* timesteps completed so far */
   int64_t stages_done;         /* stages completed so far */
   int64_t block_stages;        /* sum over stages of active blocks */
   double cell_var_updates;     /* block_stages * nx*ny*nz * num_vars */
   int64_t active_blocks;       /* now */
   int64_t max_slots;           /* high-water mark of used block slots */
   double device_ms[4];         /* exchange, stencil, checksum, split/merge */
   int64_t stage_calls;         /* amr_dev_stage calls (one stencil launch each) */
   int nx, ny, nz, num_vars, stencil, precision;
} amr_run_stats;

amr_run *amr_run_create(int argc, char **argv);
/* 0: more timesteps follow; 1: last timestep done; 2: stopped by the
 * conservation check (the reference's early return); -1: error. */
int amr_run_step(amr_run *run);
/* The next n stages of the current timestep (stage-major order), finished
 * on the device when the call returns, so the caller can look at the
 * cells between stages -- e.g. amr_dev_compare_cells against a run of the
 * same mesh in another precision after every checksum interval.  The call
 * that reaches the timestep's last stage also finishes the timestep
 * (checksums, conservation test, motion, refinement) and returns what
 * amr_run_step would; the others return 0.  Do not mix with amr_run_step
 * inside a timestep. */
int amr_run_stages(amr_run *run, int n);
/* Message of the last failed call on this thread ("" if none). */
const char *amr_run_last_error(void);
int amr_run_finish(amr_run *run);
int amr_run_stats_get(const amr_run *run, amr_run_stats *out);
void amr_run_report(amr_run *run);
amr_dev *amr_run_device(amr_run *run);
/* Active blocks now, and block `in` (0 <= in < amr_run_active) of the
 * sorted list (ref/block.h:72-77): slot, level and extent [lo, hi] in the
 * unit domain.  Returns 0, or -1 for a bad index. */
int amr_run_active(const amr_run *run);
int amr_run_block(const amr_run *run, int in, int32_t *slot, int32_t *level, double lo[3],
                  double hi[3]);
void amr_run_destroy(amr_run *run);

#ifdef __cplusplus
}
#endif

#endif
