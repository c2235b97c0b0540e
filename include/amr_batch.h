/*
 * amr_batch.h -- many independent miniAMR instances stepped together on one
 * MI355X (the batching form; no counterpart in the reference, which runs
 * one mesh per process).
 *
 * Device level.  A batch owns member contexts: each member is an ordinary
 * amr_dev (include/amr_dev.h) with its own block size, variable count, slot
 * pool and mesh, and every amr_dev_* call works on it as usual (upload,
 * set_mesh, split, consolidate, checksum table, download).  What the batch
 * adds is stepping all members at once: one persistent kernel per stage
 * phase (x, y, z exchange; stencil, one kernel per block-size class
 * present; checksum reduce) walks a flat list of work items drawn from
 * every member, so small meshes share the GPU instead of each launching
 * its own underfilled grids.  A whole timestep of such
 * launches is captured once in a hipGraph and replayed until a member's
 * mesh changes; the launch shapes (persistent grid size, variables per work
 * item) come from amr_batch_autotune, which times candidates on the actual
 * meshes and restores the cell state afterwards.
 *
 * All members must share the storage precision and the stencil (7 or 27)
 * of the batch, and have faces of at most 16 x 16 cells.
 *
 * Driver level (amr_batch_run_*): N instances of the drop-in driver, each
 * with its own reference-style command line (e.g. different --nx,
 * --num_vars, --num_refine, --seed and objects), stepped in lockstep; each
 * instance keeps its own refinement decisions, checksums and conservation
 * test, exactly as if run alone.
 */
#ifndef AMR_BATCH_H
#define AMR_BATCH_H

#include "amr_dev.h"
#include "amr_run.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct amr_batch amr_batch;

int amr_batch_create(int device, int precision, int stencil, amr_batch **out);
int amr_batch_destroy(amr_batch *b);            /* destroys remaining members */
const char *amr_batch_last_error(const amr_batch *b);

/* New member context (cfg->precision / stencil / device must match). */
int amr_batch_add(amr_batch *b, const amr_dev_config *cfg, amr_dev **member);

/* Leave a member out of (on = 0) or back into (on = 1) the batch's stages,
 * e.g. once its run has finished. */
int amr_batch_set_active(amr_batch *b, amr_dev *member, int on);

/* Launch shape of the batch's kernels (amr_batch_autotune chooses one). */
typedef struct {
   int vpw;            /* variables per stencil work item: 1 or 2 */
   int stencil_wgs;    /* persistent workgroups of a stencil kernel; 0: one
                          per work item */
   int exchange_wgs;   /* the same for the exchange kernels */
   int wg_classes;     /* bit c set: stencil items of block size class c are
                          computed by a whole 256-thread workgroup instead of
                          one wavefront (class 0: faces up to 64 cells, e.g.
                          8^3 blocks; 1: up to 192, e.g. 12^3; 2: up to 256,
                          e.g. 16^3) */
   int exchange_vars;  /* variables per exchange work item: 4, 8 or 16 */
   int exchange_threads; /* threads per exchange work item: 64 (one wavefront),
                          128 or 256 (one workgroup) */
} amr_batch_shape;

int amr_batch_set_shape(amr_batch *b, const amr_batch_shape *shape);
int amr_batch_get_shape(const amr_batch *b, amr_batch_shape *shape);

/* One stage of every member, all variables: exchange x, y, z (permuted
 * like amr_dev_comm when permute != 0), stencil, and when sum_row >= 0 the
 * checksums into each member's row sum_row.  Enqueue only. */
int amr_batch_stage(amr_batch *b, int stage, int permute, int sum_row);

/* Stages stage0 .. stage0+nstages-1 of every member as one graph launch;
 * stage s (0-based within the call) writes checksum row s / checksum_freq
 * when s % checksum_freq == 0 (checksum_freq 0: none).  Enqueue only. */
int amr_batch_timestep(amr_batch *b, int stage0, int nstages, int checksum_freq, int permute);

/* Time candidate launch shapes on the current meshes (cells restored
 * afterwards) and keep the fastest; *ms receives the chosen stage time. */
int amr_batch_autotune(amr_batch *b, double *ms);

int amr_batch_sync(amr_batch *b);

/* Device time of the batch's work enqueued since the last call (ms):
 * [0] exchange, [1] stencil, [2] checksum reduce (amr_batch_stage calls),
 * [3] whole graph launches (amr_batch_timestep calls).  Requires
 * amr_batch_set_timing(b, 1). */
int amr_batch_set_timing(amr_batch *b, int on);
int amr_batch_read_timing(amr_batch *b, double ms[4]);

/* Driver level. */
typedef struct amr_batch_run amr_batch_run;
/* argv of instance i is argvs[i][0 .. argcs[i]) (argv[0] is a program name).
 * Parses and validates every command line, builds each run's initial mesh,
 * values and refinement, then autotunes if asked.  An invalid command line
 * or a fatal error prints the reference's message and returns NULL
 * (amr_run_last_error gives it), as amr_run_create does. */
amr_batch_run *amr_batch_run_create(int n, const int *argcs, char ***argvs, int autotune);
/* 0: more timesteps follow; 1: all instances finished; 2: an instance's
 * conservation check stopped it (the others continue); -1: error (only
 * amr_batch_run_destroy may follow). */
int amr_batch_run_step(amr_batch_run *r);
int amr_batch_run_finish(amr_batch_run *r);
int amr_batch_run_instances(const amr_batch_run *r);
int amr_batch_run_stats_get(const amr_batch_run *r, int inst, amr_run_stats *out);
amr_run *amr_batch_run_get(amr_batch_run *r, int inst);  /* e.g. for amr_run_report */
amr_batch *amr_batch_run_batch(amr_batch_run *r);
double amr_batch_run_tune_ms(const amr_batch_run *r);    /* autotuned stage time */
double amr_batch_run_device_ms(const amr_batch_run *r);  /* device time of the timesteps */
void amr_batch_run_destroy(amr_batch_run *r);

#ifdef __cplusplus
}
#endif

#endif
