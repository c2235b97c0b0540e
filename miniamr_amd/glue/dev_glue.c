/*
 * dev_glue.c -- the reference miniAMR's own program running its data path
 * on the MI355X through include/amr_dev.h.
 *
 * This is the binding a maintainer adds to the reference tree.  It needs
 * no edit of a reference source: the reference's objects are linked with
 * this file, the device library and
 *
 *   -Wl,--wrap=init,--wrap=refine,--wrap=comm,--wrap=stencil_driver,
 *   -Wl,--wrap=check_sum,--wrap=driver
 *
 * so the calls the reference's driver makes to those functions land here
 * (oracle/Makefile, target glue, builds it as oracle/_ref/miniAMR_ref_glue.x
 * from the sources under /root/reference/ref).  The seam is the reference's
 * own (ref/proto.h):
 *
 *   init()            ref/init.c:38     the reference fills the cells on the
 *                                       host; then every active block is
 *                                       uploaded and the exchange plan set
 *   comm(s, n, st)    ref/comm.c:42     amr_dev_comm
 *   stencil_driver    ref/stencil.c:43  amr_dev_stencil_stage (--stencil 0:
 *                                       amr_dev_set_materials after init)
 *   check_sum(var)    ref/check_sum.c:36  amr_dev_check_sum, then the
 *                                       reference's MPI_Allreduce
 *   refine(ts)        ref/refine.c:38   cells come back to the host arrays,
 *                                       the reference's refine (topology and
 *                                       split_blocks/consolidate_blocks data
 *                                       loops) runs, cells go back and the
 *                                       plan is rebuilt
 *   driver()          ref/driver.c:37   after the reference's timestep loop
 *                                       the cells come back, so the end-of-
 *                                       run code sees the final values
 *
 * Cells live on the device between refinements; the host arrays of
 * ref/block.h hold them only around refine() and at the end.  Until init()
 * has uploaded, check_sum runs the reference's own (init.c calls it for the
 * initial sums).  Errors print the library's message and end the program,
 * as the reference's fatal paths do.
 */
#include <stdio.h>
#include <stdlib.h>
#include <mpi.h>

#include "block.h"
#include "comm.h"
#include "proto.h"
#include "amr_dev.h"

void __real_init(void);
void __real_refine(int ts);
void __real_comm(int start, int num_comm, int stage);
void __real_stencil_driver(int var, int calc_stage);
double __real_check_sum(int var);
void __real_driver(void);

static amr_dev *dev;
static int on_device;
static double *cellbuf;          /* one block: [var][i][j][k], ghosts included */

static void check(int rc, const char *what)
{
   if (rc != AMR_OK) {
      printf("ERROR: device %s failed (%d): %s\n", what, rc, amr_dev_last_error(dev));
      exit(-1);
   }
}

static void create(void)
{
   if (dev)
      return;
   amr_dev_config c = { x_block_size, y_block_size, z_block_size, num_vars, max_num_blocks,
                        stencil, AMR_PREC_F64, 0 };
   check(amr_dev_create(&c, &dev), "create");
   cellbuf = (double *) malloc(sizeof(double) * num_vars * (x_block_size + 2) *
                               (y_block_size + 2) * (z_block_size + 2));
}

static int n_active(void)
{
   return sorted_index[num_refine + 1];
}

/* Between the reference's ragged array[var][i][j][k] and a contiguous one. */
static void copy_block(block *bp, int to_host)
{
   const int gx = x_block_size + 2, gy = y_block_size + 2, gz = z_block_size + 2;
   double *p = cellbuf;
   for (int v = 0; v < num_vars; v++)
      for (int i = 0; i < gx; i++)
         for (int j = 0; j < gy; j++, p += gz)
            for (int k = 0; k < gz; k++) {
               if (to_host)
                  bp->array[v][i][j][k] = p[k];
               else
                  p[k] = bp->array[v][i][j][k];
            }
}

static void upload_all(void)
{
   for (int in = 0; in < n_active(); in++) {
      const int n = sorted_list[in].n;
      copy_block(&blocks[n], 0);
      check(amr_dev_upload_block(dev, n, cellbuf), "upload");
   }
}

static void download_all(void)
{
   for (int in = 0; in < n_active(); in++) {
      const int n = sorted_list[in].n;
      check(amr_dev_download_block(dev, n, cellbuf), "download");
      copy_block(&blocks[n], 1);
   }
}

/* One face job per ghost plane of every active block, direction by
 * direction: what comm()'s on-rank loop fills (ref/comm.c:159-203) --
 * same level, one restriction per finer quarter, the prolongation from the
 * coarser neighbour, or the boundary. */
static void set_mesh(void)
{
   const int na = n_active(), ghosts = code ? AMR_FACE_GHOSTS : 0;
   int32_t *slots = (int32_t *) malloc(sizeof(int32_t) * (na ? na : 1));
   amr_face_job *jobs = (amr_face_job *) malloc(sizeof(amr_face_job) * (24 * na + 1));
   int32_t count[3] = { 0, 0, 0 };
   int nj = 0;
   for (int in = 0; in < na; in++)
      slots[in] = sorted_list[in].n;
   for (int dir = 0; dir < 3; dir++) {
      const int before = nj;
      for (int in = 0; in < na; in++) {
         const int n = sorted_list[in].n;
         const block *bp = &blocks[n];
         for (int f = 2 * dir; f < 2 * dir + 2; f++) {
            amr_face_job j = { 0, f, n, n, 0, 0 };
            if (bp->nei_level[f] == bp->level) {
               j.kind = AMR_FACE_SAME | ghosts;
               j.src = bp->nei[f][0][0];
               jobs[nj++] = j;
            } else if (bp->nei_level[f] == bp->level + 1) {
               for (int iq = 0; iq < 2; iq++)
                  for (int jq = 0; jq < 2; jq++) {
                     j.kind = AMR_FACE_RESTRICT | ghosts;
                     j.src = bp->nei[f][iq][jq];
                     j.iq = iq;
                     j.jq = jq;
                     jobs[nj++] = j;
                  }
            } else if (bp->nei_level[f] == bp->level - 1) {
               const int m = bp->nei[f][0][0], g = f ^ 1;
               for (int iq = 0; iq < 2; iq++)
                  for (int jq = 0; jq < 2; jq++)
                     if (blocks[m].nei[g][iq][jq] == n) {
                        j.kind = AMR_FACE_PROLONG | ghosts;
                        j.src = m;
                        j.iq = iq;
                        j.jq = jq;
                        jobs[nj++] = j;
                     }
            } else if (bp->nei_level[f] == -2) {
               j.kind = AMR_FACE_BC | ghosts;
               jobs[nj++] = j;
            } else {
               printf("ERROR: misconnected block\n");
               exit(-1);
            }
         }
      }
      count[dir] = nj - before;
   }
   check(amr_dev_set_mesh(dev, slots, na, jobs, count), "set_mesh");
   free(slots);
   free(jobs);
}

void __wrap_init(void)
{
   __real_init();
   create();
   if (stencil == 0)
      check(amr_dev_set_materials(dev, a1, a0, mat), "set_materials");
   upload_all();
   set_mesh();
   on_device = 1;
}

void __wrap_refine(int ts)
{
   if (on_device) {
      download_all();
      on_device = 0;
   }
   __real_refine(ts);
   upload_all();
   set_mesh();
   on_device = 1;
}

void __wrap_comm(int start, int num_comm, int stage)
{
   if (!on_device) {
      __real_comm(start, num_comm, stage);
      return;
   }
   check(amr_dev_comm(dev, start, num_comm, stage, permute), "comm");
}

void __wrap_stencil_driver(int var, int calc_stage)
{
   if (!on_device) {
      __real_stencil_driver(var, calc_stage);
      return;
   }
   check(amr_dev_stencil_stage(dev, var, calc_stage), "stencil");
}

double __wrap_check_sum(int var)
{
   double sum, gsum;
   if (!on_device)
      return __real_check_sum(var);
   check(amr_dev_check_sum(dev, var, &sum), "check_sum");
   MPI_Allreduce(&sum, &gsum, 1, MPI_DOUBLE, MPI_SUM, MPI_COMM_WORLD);
   return gsum;
}

void __wrap_driver(void)
{
   __real_driver();
   if (on_device) {
      download_all();
      on_device = 0;
   }
}
