/*
 * batch_run.c -- the driver-level batch of include/amr_batch.h: N runs of
 * the drop-in driver, each with its own command line, stepped in lockstep
 * on one device.
 *
 * Each run keeps everything the single-run driver keeps (mesh, objects,
 * refinement, slot allocation, checksum table and conservation test) and
 * its device context is a member of one amr_batch.  Per timestep:
 *   1. every unfinished run prints its timestep line (driver "begin");
 *   2. the batch enqueues that timestep's stages for all of them as one
 *      graph launch (amr_batch_timestep);
 *   3. every run reads its checksums, tests conservation, moves its
 *      objects and refines (driver "end"), exactly as when run alone.
 * Runs that finish (or stop on the conservation test) leave the batch.
 * The runs must agree on stages_per_ts, checksum_freq, --permute,
 * --stencil and the storage precision, since they share the stage
 * sequence; everything else (block size, variables, refinement levels,
 * objects, seed, timesteps) may differ.
 *
 * With AMR_DUMP_SUMS / AMR_DUMP_CELLS / AMR_DUMP_BSUMS / AMR_DUMP_SAMPLE set, run i writes
 * its checksum log and final dumps to "<path>.<i>".
 */
#include <setjmp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "amr.h"
#include "amr_batch.h"
#include "amr_run.h"

struct amr_batch_run {
   int n;
   amr_run **runs;
   amr_batch *batch;
   int stopped;          /* some run stopped on the conservation test */
   double tune_ms;       /* stage time chosen by the autotuner, 0 if not run */
   double graph_ms;      /* device time of the timestep graphs so far */
   int failed;           /* a call hit a fatal error */
};

static char *suffixed(const char *base, int i)
{
   if (!base)
      return NULL;
   size_t len = strlen(base) + 16;
   char *p = (char *) amr_xmalloc(len);
   snprintf(p, len, "%s.%d", base, i);
   return p;
}

static void release(amr_batch_run *B)
{
   for (int i = 0; i < B->n; i++)
      if (B->runs && B->runs[i]) {
         amr_free(B->runs[i]);   /* leaves the member context to the batch */
         free(B->runs[i]);
      }
   if (B->batch)
      amr_batch_destroy(B->batch);
   free(B->runs);
   free(B);
}

/* Fatal errors return here as a NULL batch (amr_fatal, mesh.c). */
amr_batch_run *amr_batch_run_create(int n, const int *argcs, char ***argvs, int autotune)
{
   jmp_buf jb;
   amr_batch_run *volatile B = NULL;
   if (setjmp(jb)) {
      amr_catch(NULL);
      if (B)
         release(B);
      return NULL;
   }
   amr_catch(&jb);
   if (n < 1)
      amr_fatal("batch: at least one run is needed\n");
   B = (amr_batch_run *) amr_xmalloc(sizeof(amr_batch_run));
   B->runs = (amr_run **) amr_xmalloc(sizeof(amr_run *) * n);
   B->n = n;
   for (int i = 0; i < n; i++) {
      amr_run *R = (amr_run *) amr_xmalloc(sizeof(amr_run));
      B->runs[i] = R;
      amr_parse_args(&R->P, argcs[i], argvs[i]);
      if (amr_check_input(&R->P))
         amr_fatal_quiet("invalid input");
      if (!R->P.block_change)
         R->P.block_change = R->P.num_refine;
   }
   const amr_params *P0 = &B->runs[0]->P;
   for (int i = 1; i < n; i++) {
      const amr_params *P = &B->runs[i]->P;
      if (P->stages_per_ts != P0->stages_per_ts || P->checksum_freq != P0->checksum_freq ||
          P->permute != P0->permute || P->stencil != P0->stencil ||
          P->precision != P0->precision || P->device != P0->device)
         amr_fatal("batch: run %d differs from run 0 in stages_per_ts, checksum_freq, "
                   "permute, stencil, precision or device\n", i);
   }
   int rc = amr_batch_create(P0->device, P0->precision, P0->stencil, &B->batch);
   if (rc != AMR_OK)
      amr_fatal("ERROR: could not create the batch (%d): %s\n", rc,
                amr_batch_last_error(B->batch));
   const char *sp = getenv("AMR_DUMP_SUMS");
   for (int i = 0; i < n; i++) {
      amr_run *R = B->runs[i];
      R->batch = B->batch;
      amr_allocate(R);
      if (sp) {
         char *p = suffixed(sp, i);
         amr_driver_set_sum_log(R, p);
         free(p);
      }
      amr_driver_begin(R);
   }
   if (autotune) {
      rc = amr_batch_autotune(B->batch, &B->tune_ms);
      if (rc != AMR_OK)
         amr_fatal("ERROR: batch autotune failed (%d): %s\n", rc,
                   amr_batch_last_error(B->batch));
   }
   amr_batch_set_timing(B->batch, 1);
   double ms[4];
   amr_batch_read_timing(B->batch, ms);
   amr_catch(NULL);
   return B;
}

static int batch_step(amr_batch_run *B);

/* -1 after a fatal error; the batch then only accepts destroy. */
int amr_batch_run_step(amr_batch_run *B)
{
   jmp_buf jb;
   if (!B || B->failed)
      return -1;
   if (setjmp(jb)) {
      amr_catch(NULL);
      B->failed = 1;
      return -1;
   }
   amr_catch(&jb);
   int rc = batch_step(B);
   amr_catch(NULL);
   return rc;
}

static int batch_step(amr_batch_run *B)
{
   int live = 0;
   int *go = (int *) amr_xmalloc(sizeof(int) * B->n);
   for (int i = 0; i < B->n; i++) {
      go[i] = !amr_driver_step_begin(B->runs[i]);
      live += go[i];
      amr_batch_set_active(B->batch, B->runs[i]->dev, go[i]);
   }
   if (!live) {
      free(go);
      return B->stopped ? 2 : 1;
   }
   const amr_params *P0 = &B->runs[0]->P;
   /* all live runs are at the same stage: they started together and every
    * timestep advances each by stages_per_ts */
   int stage0 = -1;
   for (int i = 0; i < B->n; i++)
      if (go[i]) {
         if (stage0 < 0)
            stage0 = B->runs[i]->loop.comm_stage;
         else if (B->runs[i]->loop.comm_stage != stage0)
            amr_fatal("batch: runs out of step\n");
      }
   int rc = amr_batch_timestep(B->batch, stage0, P0->stages_per_ts, P0->checksum_freq,
                               P0->permute);
   if (rc != AMR_OK)
      amr_fatal("ERROR: batch timestep failed (%d): %s\n", rc, amr_batch_last_error(B->batch));
   double ms[4];
   amr_batch_read_timing(B->batch, ms);
   B->graph_ms += ms[3];
   for (int i = 0; i < B->n; i++)
      if (go[i]) {
         amr_run *R = B->runs[i];
         amr_driver_step_advance(R);
         R->loop.stage_calls += R->P.stages_per_ts;
         amr_count_comm(R, R->P.stages_per_ts *
                              ((R->P.num_vars + R->P.comm_vars - 1) / R->P.comm_vars));
         if (amr_driver_step_end(R) == 2)
            B->stopped = 1;
      }
   free(go);
   for (int i = 0; i < B->n; i++)
      if (!B->runs[i]->loop.state)
         return 0;
   return B->stopped ? 2 : 1;
}

int amr_batch_run_finish(amr_batch_run *B)
{
   if (!B || B->failed)
      return -1;
   for (int i = 0; i < B->n; i++)
      amr_driver_end(B->runs[i]);
   return B->stopped;
}

int amr_batch_run_instances(const amr_batch_run *B)
{
   return B->n;
}

int amr_batch_run_stats_get(const amr_batch_run *B, int inst, amr_run_stats *out)
{
   if (inst < 0 || inst >= B->n)
      return -1;
   return amr_run_stats_get(B->runs[inst], out);
}

amr_run *amr_batch_run_get(amr_batch_run *B, int inst)
{
   return (inst >= 0 && inst < B->n) ? B->runs[inst] : NULL;
}

double amr_batch_run_tune_ms(const amr_batch_run *B)
{
   return B->tune_ms;
}

double amr_batch_run_device_ms(const amr_batch_run *B)
{
   return B->graph_ms;
}

amr_batch *amr_batch_run_batch(amr_batch_run *B)
{
   return B->batch;
}

void amr_batch_run_destroy(amr_batch_run *B)
{
   jmp_buf jb;
   if (!B)
      return;
   const char *pc = getenv("AMR_DUMP_CELLS");
   const char *pb = getenv("AMR_DUMP_BSUMS");
   const char *ps = getenv("AMR_DUMP_SAMPLE");
   if (!B->failed && !setjmp(jb)) {
      amr_catch(&jb);
      for (int i = 0; i < B->n; i++) {
         char *c = suffixed(pc, i), *b = suffixed(pb, i), *q = suffixed(ps, i);
         amr_dump_final_to(B->runs[i], c, b, q);
         free(c);
         free(b);
         free(q);
      }
   }
   amr_catch(NULL);
   release(B);
}
