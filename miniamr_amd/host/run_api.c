/*
 * run_api.c -- include/amr_run.h: the driver of main.c, one timestep per
 * call, for callers that time or drive the run themselves.
 */
#include <setjmp.h>
#include <stdlib.h>
#include <string.h>

#include "amr.h"
#include "amr_run.h"

/* Library entry points arm a recovery point for amr_fatal (mesh.c): a
 * fatal error inside them returns here as an error instead of ending the
 * caller's process.  A run that failed mid-step keeps failing; only
 * amr_run_destroy may follow. */
amr_run *amr_run_create(int argc, char **argv)
{
   jmp_buf jb;
   amr_run *volatile R = NULL;
   if (setjmp(jb)) {
      amr_catch(NULL);
      if (R) {
         amr_free(R);
         free(R);
      }
      return NULL;
   }
   amr_catch(&jb);
   R = (amr_run *) amr_xmalloc(sizeof(amr_run));
   amr_parse_args(&R->P, argc, argv);
   if (amr_check_input(&R->P))
      amr_fatal_quiet("invalid input");
   if (!R->P.block_change)
      R->P.block_change = R->P.num_refine;
   amr_allocate(R);
   amr_driver_begin(R);
   amr_catch(NULL);
   return R;
}

int amr_run_step(amr_run *R)
{
   jmp_buf jb;
   if (!R || R->failed)
      return -1;
   if (setjmp(jb)) {
      amr_catch(NULL);
      R->failed = 1;
      return -1;
   }
   amr_catch(&jb);
   int rc = amr_driver_step(R);
   amr_catch(NULL);
   return rc;
}

int amr_run_stages(amr_run *R, int n)
{
   jmp_buf jb;
   if (!R || R->failed)
      return -1;
   if (setjmp(jb)) {
      amr_catch(NULL);
      R->failed = 1;
      return -1;
   }
   amr_catch(&jb);
   int rc = amr_driver_stages(R, n);
   amr_catch(NULL);
   return rc;
}

const char *amr_run_last_error(void)
{
   return amr_error_message();
}

int amr_run_finish(amr_run *R)
{
   jmp_buf jb;
   if (!R || R->failed)
      return -1;
   if (setjmp(jb)) {
      amr_catch(NULL);
      R->failed = 1;
      return -1;
   }
   amr_catch(&jb);
   amr_driver_end(R);
   amr_catch(NULL);
   return R->loop.state == 2;
}

int amr_run_stats_get(const amr_run *R, amr_run_stats *out)
{
   const amr_params *P = &R->P;
   memset(out, 0, sizeof(*out));
   out->timesteps_done = R->loop.ts - 1;
   out->stages_done = out->timesteps_done * P->stages_per_ts;
   out->block_stages = R->total_blocks;
   out->cell_var_updates = (double) R->total_blocks * P->nx * P->ny * P->nz * P->num_vars;
   out->active_blocks = amr_n_active(R);
   out->max_slots = R->max_active_block;
   memcpy(out->device_ms, R->loop.dev_ms, sizeof(out->device_ms));
   out->stage_calls = R->loop.stage_calls;
   out->nx = P->nx; out->ny = P->ny; out->nz = P->nz;
   out->num_vars = P->num_vars;
   out->stencil = P->stencil;
   out->precision = P->precision;
   return 0;
}

void amr_run_report(amr_run *R)
{
   amr_report(R);
   fflush(stdout);
}

amr_dev *amr_run_device(amr_run *R)
{
   return R->dev;
}

int amr_run_active(const amr_run *R)
{
   return amr_n_active(R);
}

/* Block `in` of the sorted list: its slot, level and extent in the unit
 * domain (the corners ref/move.c:80-88 computes). */
int amr_run_block(const amr_run *R, int in, int32_t *slot, int32_t *level, double lo[3],
                  double hi[3])
{
   if (in < 0 || in >= amr_n_active(R))
      return -1;
   int s = R->sorted_slot[in];
   const amr_block *b = &R->blocks[s];
   int sz = R->p2[R->P.num_refine - b->level];
   if (slot)
      *slot = s;
   if (level)
      *level = b->level;
   for (int d = 0; d < 3; d++) {
      if (lo)
         lo[d] = ((double) (b->cen[d] - sz)) / ((double) R->mesh_size[d]);
      if (hi)
         hi[d] = ((double) (b->cen[d] + sz)) / ((double) R->mesh_size[d]);
   }
   return 0;
}

void amr_run_destroy(amr_run *R)
{
   jmp_buf jb;
   if (!R)
      return;
   if (!setjmp(jb)) {
      amr_catch(&jb);
      if (!R->failed)
         amr_dump_final(R);
   }
   amr_catch(NULL);
   amr_free(R);
   free(R);
}
