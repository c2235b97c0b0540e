/*
 * driver.c -- the timestep loop of the drop-in (ref/driver.c:37-214).
 *
 * The reference runs, per stage, "exchange a group of variables, then
 * stencil and checksum each variable" and tests conservation right after
 * each checksum.  Variables are independent (the stencil and every
 * exchange rule act on one variable at a time), so the order in which
 * (stage, variable) pairs are computed does not change any value.  This
 * driver therefore enqueues a whole timestep on the device without
 * waiting, in one of two orders:
 *   stage-major (default): stage by stage, all variables per launch, as
 *                          the reference does;
 *   variable-major (MINIAMR_SCHEDULE=var): a group of variables through all
 *                          stages of the timestep, then the next group, so
 *                          a group's cells stay in the device's last-level
 *                          cache across stages.
 * The variable-work stencils of --stencil 0 couple variables (each reads
 * others, some already updated in the stage), so that mode always runs
 * stage-major, variables in order within a stage.
 * The checksums of the timestep land in a device table and are then
 * checked on the host in the reference's order, printing the same lines.
 * The only observable difference is after a failed conservation check:
 * the reference stops mid-stage, while here the rest of that timestep's
 * cell updates have already run (the printed output and the Summary line
 * are the same).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "amr.h"

double amr_checksum(amr_run *R, int var)
{
   double s = 0.0;
   amr_dev_check(R, amr_dev_check_sum(R->dev, var, &s), "check_sum");
   R->total_red++;
   return s;
}

static void log_sum(amr_run *R, double s)
{
   if (R->sum_log) {
      fprintf(R->sum_log, "%.17g\n", s);
      fflush(R->sum_log);
   }
}

/* Checksum log of this run (the AMR_DUMP_SUMS format); NULL closes it. */
void amr_driver_set_sum_log(amr_run *R, const char *path)
{
   if (R->sum_log)
      fclose(R->sum_log);
   R->sum_log = path ? fopen(path, "w") : NULL;
}

/* ref/driver.c:171-214 */
double amr_calc_time_step(amr_run *R)
{
   const amr_params *P = &R->P;
   double delta, tmp, inv_cell_size[3];
   if (P->use_tsteps)
      return P->end_time / (double) P->num_tsteps;
   inv_cell_size[0] = (double) (R->mesh_size[0] * P->nx);
   inv_cell_size[1] = (double) (R->mesh_size[1] * P->ny);
   inv_cell_size[2] = (double) (R->mesh_size[2] * P->nz);
   delta = 0.0;
   for (int o = 0; o < P->num_objects; o++) {
      const amr_object *op = &P->objects[o];
      int done = 0;
      if (op->size[0] < 0.0 || op->size[1] < 0.0 || op->size[2] < 0.0)
         break;
      for (int d = 0; d < 3; d++)
         if (op->cen[d] < 0.0) {
            if (op->cen[d] + op->size[d] < 0.0)
               done = 1;
         } else if (op->cen[d] > 1.0) {
            if (op->cen[d] - op->size[d] > 1.0)
               done = 1;
         }
      if (done)
         break;
      for (int d = 0; d < 3; d++) {
         tmp = (fabs(op->move[d]) + fabs(op->inc[d])) * inv_cell_size[d];
         if (tmp > delta)
            delta = tmp;
      }
   }
   return delta > 0.0 ? 1.0 / delta : 1.0;
}

/* Whether every ghost plane must hold what the reference's would at the
 * end of the current timestep: before a refinement (split, consolidate and
 * the new mesh's plan reuse slots whose ghost cells later stencils read as
 * they are), at the end of the run (dumps, saved state), and whenever the
 * run's length is not known in advance (--time).  Between two timesteps on
 * the same mesh nothing reads a ghost plane that a fused stage leaves
 * unwritten: the next stages read those faces in place again.
 * MINIAMR_PLAIN_LAST=1 ends every timestep with a plain stage. */
static int ghosts_needed_after_timestep(const amr_run *R)
{
   const amr_params *P = &R->P;
   static int every = -1;
   if (every < 0) {
      const char *e = getenv("MINIAMR_PLAIN_LAST");
      every = e && atoi(e) != 0;
   }
   const int ts = R->loop.ts;
   return every || P->use_time || ts >= P->num_tsteps ||
          (P->num_refine && !P->uniform_refine && !(ts % P->refine_freq));
}

/* One stage of variables [start, start+number).  Every stage of a
 * timestep but the last reads same-level faces in place (amr_dev_stage_fused:
 * the same cells and checksums); the last is a plain stage where the ghost
 * planes are looked at after it (ghosts_needed_after_timestep), so that
 * there every ghost plane holds what the reference's would. */
static void enqueue_stage(amr_run *R, int start, int number, int stage, int comm_stage,
                          int row)
{
   const amr_params *P = &R->P;
   if (stage < P->stages_per_ts - 1 || !ghosts_needed_after_timestep(R))
      amr_dev_check(R, amr_dev_stage_fused(R->dev, start, number, comm_stage, P->permute, row),
                    "stage");
   else
      amr_dev_check(R, amr_dev_stage(R->dev, start, number, comm_stage, P->permute, row),
                    "stage");
}

/* --stencil 0 operation counts of one stage's stencils of variables
 * [start, start+number) (ref/stencil.c:100-101 for the 7-point average,
 * :147-958 for the variable-work updates): per variable, by calculation
 * stage % 6 and class, per cell of every active block.  The clamp's
 * data-dependent counts (ref/stencil.c:966-981) come from the device. */
static void count_flops0(amr_run *R, int start, int number, int calc_stage)
{
   const amr_params *P = &R->P;
   const double cells = (double) R->global_active * P->nx * P->ny * P->nz;
   const int mat = R->mat, op = calc_stage % 6;
   for (int var = start; var < start + number; var++) {
      double adds, muls = 0.0, divs = 1.0;
      const int cls = var / mat;
      if (var == 0 || var >= 4 * mat)
         adds = 6.0;
      else if (op == 0) {
         if (var == 1) { adds = mat; muls = mat; divs = 0.0; }
         else if (cls == 0) { adds = 3.0; muls = 2.0; divs = 0.0; }
         else if (cls == 1) { adds = 3.0; muls = 3.0; }
         else if (cls == 2) { adds = 4.0; muls = 3.0; }
         else { adds = 6.0; muls = 6.0; }
      } else if (op <= 3) {
         if (var == 1) { adds = mat + 1; muls = mat; }
         else if (cls == 0) { adds = 4.0; muls = 2.0; }
         else { adds = 12.0; muls = 3.0; }
      } else if (op == 4) {
         adds = 7.0;
         muls = 8.0;
      } else {
         adds = 27.0;
      }
      R->total_fp_adds += adds * cells;
      R->total_fp_muls += muls * cells;
      R->total_fp_divs += divs * cells;
   }
}

/* Stages s0 .. s1-1 of a timestep in stage-major order; checksum event e
 * of the timestep goes to row e. */
static void enqueue_stage_range(amr_run *R, int comm_stage0, int s0, int s1)
{
   const amr_params *P = &R->P;
   int V = P->num_vars;
   for (int stage = s0; stage < s1; stage++) {
      int row = (P->checksum_freq && !(stage % P->checksum_freq))
                   ? stage / P->checksum_freq : -1;
      for (int start = 0; start < V; start += P->comm_vars) {
         int number = (start + P->comm_vars > V) ? V - start : P->comm_vars;
         enqueue_stage(R, start, number, stage, comm_stage0 + stage, row);
         if (P->stencil == 0)
            count_flops0(R, start, number, comm_stage0 + stage);
         R->loop.stage_calls++;
         amr_count_comm(R, 1);
      }
   }
}

/* Enqueue one timestep. */
static void enqueue_timestep(amr_run *R, int comm_stage0)
{
   const amr_params *P = &R->P;
   int V = P->num_vars;
   if (P->schedule == 0) {
      enqueue_stage_range(R, comm_stage0, 0, P->stages_per_ts);
   } else {
      /* Variables per group: as many as keep the group's active slabs
       * within ~128 MiB, so they stay in the 256 MiB last-level cache
       * across the stages of the timestep (MINIAMR_VAR_GROUP overrides). */
      static const int elem[3] = { 8, 4, 2 };
      double slab = (double) (P->nx + 2) * (P->ny + 2) * (P->nz + 2) * elem[P->precision];
      double per_var = slab * (amr_n_active(R) > 0 ? amr_n_active(R) : 1);
      int group = (int) ((128.0 * 1024 * 1024) / per_var);
      if (group < 1)
         group = 1;
      if (group > V)
         group = V;
      const char *g = getenv("MINIAMR_VAR_GROUP");
      if (g && atoi(g) > 0)
         group = atoi(g);
      for (int start = 0; start < V; start += group) {
         int number = (start + group > V) ? V - start : group;
         for (int stage = 0; stage < P->stages_per_ts; stage++) {
            int row = (P->checksum_freq && !(stage % P->checksum_freq))
                         ? stage / P->checksum_freq : -1;
            enqueue_stage(R, start, number, stage, comm_stage0 + stage, row);
            R->loop.stage_calls++;
            amr_count_comm(R, 1);
         }
      }
   }
}

/* Faces exchanged by `calls` comm calls on the current plan. */
void amr_count_comm(amr_run *R, int calls)
{
   for (int d = 0; d < 3; d++) {
      R->face_same[d] += (int64_t) calls * R->plan_faces[d][0];
      R->face_diff[d] += (int64_t) calls * R->plan_faces[d][1];
      R->face_bc[d] += (int64_t) calls * R->plan_faces[d][2];
   }
}

/* Check this timestep's checksums in the reference's order
 * (ref/driver.c:88-103).  Returns 1 if the run must stop. */
static int check_timestep(amr_run *R, int ts, const double *sums, int64_t *blocks_done)
{
   const amr_params *P = &R->P;
   int V = P->num_vars;
   for (int stage = 0; stage < P->stages_per_ts; stage++) {
      *blocks_done += R->global_active;
      if (!(P->checksum_freq && !(stage % P->checksum_freq)))
         continue;
      const double *row = sums + (size_t) (stage / P->checksum_freq) * V;
      for (int var = 0; var < V; var++) {
         double sum = row[var];
         R->total_red++;
         log_sum(R, sum);
         if (P->report_diffusion)
            printf("%d var %d sum %lf old %lf diff %lf %lf tol %lf\n", ts, var, sum,
                   R->grid_sum[var], (sum - R->grid_sum[var]),
                   (fabs(sum - R->grid_sum[var]) / R->grid_sum[var]), R->tol);
         if (P->stencil || var == 0)
            if (fabs(sum - R->grid_sum[var]) / R->grid_sum[var] > R->tol) {
               printf("Time step %d sum %lf (old %lf) variable %d difference too large\n",
                      ts, sum, R->grid_sum[var], var);
               return 1;
            }
         R->grid_sum[var] = sum;
      }
   }
   return 0;
}

/* ref/driver.c:37-63: initial mesh, values and refinement. */
int amr_driver_begin(amr_run *R)
{
   amr_params *P = &R->P;
   amr_loop *L = &R->loop;
   const char *sp = getenv("AMR_DUMP_SUMS");
   if (sp && !R->sum_log && !R->batch)
      amr_driver_set_sum_log(R, sp);

   double t2 = amr_now();
   amr_init(R);
   for (int v = 0; v < P->num_vars; v++)
      log_sum(R, R->grid_sum[v]);
   double t1 = amr_now();
   R->t_init = t1 - t2;
   R->total_red = 0;   /* the initial sums are not counted (ref/driver.c:45-46) */
   amr_dev_memory(R->dev, &R->mem_bytes_init, &R->mem_buffers_init);

   if (P->num_refine || P->uniform_refine)
      amr_refine(R, 0);
   amr_dev_check(R, amr_dev_sync(R->dev), "sync");
   t2 = amr_now();
   R->t_refine_all += R->t_refine_init = t2 - t1;
   R->nb_min = R->nb_max = R->global_active;

   L->n_rows = P->checksum_freq ? (P->stages_per_ts + P->checksum_freq - 1) / P->checksum_freq : 0;
   L->sums = (double *) amr_xmalloc(sizeof(double) * (size_t) (L->n_rows ? L->n_rows : 1) *
                                    P->num_vars);
   amr_dev_check(R, amr_dev_reserve_sums(R->dev, L->n_rows ? L->n_rows : 1), "reserve_sums");
   /* exchange time per direction only when a report prints it */
   amr_dev_set_timing(R->dev, (P->report_perf & 7) ? 2 : 1);
   double ms[4];
   amr_dev_read_timing(R->dev, ms);

   L->delta = P->use_time ? amr_calc_time_step(R) : 1.0;
   L->sim_time = 0.0;
   L->comm_stage = 0;
   L->ts = 1;
   L->state = 0;
   L->t_start = t1;
   return 0;
}

/* One timestep of ref/driver.c:67-160, in phases so that a batch can
 * enqueue the stages of many runs together:
 *   begin    -- returns the run's state (nonzero: nothing more to do),
 *               else prints the timestep line;
 *   enqueue  -- this run's stages on its own device context;
 *   advance  -- after the stages were enqueued (by enqueue or a batch);
 *   end      -- checksums, conservation test, motion and refinement.
 * Returns 0 while more timesteps follow, 1 after the last one, 2 if a
 * conservation check stopped the run. */
int amr_driver_step_begin(amr_run *R)
{
   amr_loop *L = &R->loop;
   if (L->state)
      return L->state;
   if (R->P.report_perf & 8)
      printf("Timestep %d time %lf delta %lf\n", L->ts, L->sim_time, L->delta);
   return 0;
}

void amr_driver_step_enqueue(amr_run *R)
{
   enqueue_timestep(R, R->loop.comm_stage);
}

void amr_driver_step_advance(amr_run *R)
{
   R->loop.comm_stage += R->P.stages_per_ts;
}

/* Read and check timestep ts's checksums (ref/driver.c:88-103).  Returns 1
 * if the run must stop. */
static int read_and_check(amr_run *R)
{
   amr_loop *L = &R->loop;
   if (L->n_rows)
      amr_dev_check(R, amr_dev_read_sums(R->dev, 0, L->n_rows, L->sums), "read_sums");
   else
      amr_dev_check(R, amr_dev_sync(R->dev), "sync");
   double t3 = amr_now();
   int stop = check_timestep(R, L->ts, L->sums, &R->total_blocks);
   R->t_cs_host += amr_now() - t3;
   return stop;
}

int amr_driver_step_end(amr_run *R)
{
   amr_params *P = &R->P;
   amr_loop *L = &R->loop;
   int ts = L->ts;
   int stop = L->checked ? L->checked == 2 : read_and_check(R);
   L->overlapped = L->checked != 0;
   L->checked = 0;
   if (stop) {
      double ms[4];
      amr_dev_read_timing(R->dev, ms);
      for (int i = 0; i < 4; i++)
         L->dev_ms[i] += ms[i];
      L->overlapped = 0;
      L->state = 2;
      return 2;
   }
   double t4 = amr_now();

   if (P->num_refine && !P->uniform_refine) {
      amr_move_objects(R, L->delta);
      if (!(ts % P->refine_freq)) {
         amr_refine(R, ts);
         if (R->global_active < R->nb_min)
            R->nb_min = R->global_active;
         if (R->global_active > R->nb_max)
            R->nb_max = R->global_active;
      }
   }
   L->overlapped = 0;
   amr_dev_check(R, amr_dev_sync(R->dev), "sync");
   R->t_refine_all += amr_now() - t4;
   /* device time of the timestep's spans (and of the refinement's data
    * work, which the next timestep would otherwise count) */
   double ms[4];
   amr_dev_read_timing(R->dev, ms);
   for (int i = 0; i < 4; i++)
      L->dev_ms[i] += ms[i];

   L->sim_time += L->delta;
   int done = 0;
   if (P->use_time) {
      if (P->use_tsteps && ts >= P->num_tsteps)
         done = 1;
      L->delta = amr_calc_time_step(R);
      if (L->sim_time >= P->end_time)
         done = 1;
   } else if (ts >= P->num_tsteps) {
      done = 1;
   }
   L->ts++;
   if (done)
      L->state = 1;
   return L->state;
}

/* The stage after which a timestep that ends in a refinement checks its
 * checksums, or -1.  Refinement depends on the objects and the mesh only,
 * never on cell values, so once the last checksum of the timestep has been
 * checked (ref/driver.c:88-103: the test that may stop the run comes
 * first) its host work -- marking, the split and consolidate loops, the new
 * mesh's plans -- can run while the device still works through the
 * timestep's remaining stages, instead of leaving the device idle after
 * them.  Only the stage-major order, and only when stages follow the last
 * checksum stage. */
static int early_check_stage(const amr_run *R)
{
   const amr_params *P = &R->P;
   if (P->schedule != 0 || !R->loop.n_rows || !P->num_refine || P->uniform_refine ||
       (R->loop.ts % P->refine_freq))
      return -1;
   const int last = (P->stages_per_ts - 1) / P->checksum_freq * P->checksum_freq;
   return last < P->stages_per_ts - 1 ? last : -1;
}

int amr_driver_step(amr_run *R)
{
   int st = amr_driver_step_begin(R);
   if (st)
      return st;
   const int c = early_check_stage(R);
   if (c >= 0) {
      amr_loop *L = &R->loop;
      enqueue_stage_range(R, L->comm_stage, 0, c + 1);
      L->checked = read_and_check(R) ? 2 : 1;
      /* the remaining stages run whatever the check said, as when the
       * whole timestep is checked at its end */
      enqueue_stage_range(R, L->comm_stage, c + 1, R->P.stages_per_ts);
   } else {
      amr_driver_step_enqueue(R);
   }
   amr_driver_step_advance(R);
   return amr_driver_step_end(R);
}

/* The next n stages of the current timestep, stage-major, completed on
 * the device before returning; the call that reaches the timestep's last
 * stage finishes the timestep as amr_driver_step does and returns its
 * code, the others return 0. */
int amr_driver_stages(amr_run *R, int n)
{
   amr_loop *L = &R->loop;
   const int spt = R->P.stages_per_ts;
   if (L->stage_pos == 0) {
      int st = amr_driver_step_begin(R);
      if (st)
         return st;
   }
   int s1 = L->stage_pos + (n > 0 ? n : 1);
   if (s1 > spt)
      s1 = spt;
   enqueue_stage_range(R, L->comm_stage, L->stage_pos, s1);
   L->stage_pos = s1;
   if (s1 < spt) {
      amr_dev_check(R, amr_dev_sync(R->dev), "sync");
      return 0;
   }
   L->stage_pos = 0;
   amr_driver_step_advance(R);
   return amr_driver_step_end(R);
}

/* ref/driver.c:162-164 and the timer bookkeeping for the report. */
void amr_driver_end(amr_run *R)
{
   amr_params *P = &R->P;
   amr_loop *L = &R->loop;
   double ms[4];
   amr_dev_read_timing(R->dev, ms);
   for (int i = 0; i < 4; i++)
      L->dev_ms[i] += ms[i];
   R->t_comm = L->dev_ms[0] * 1e-3;
   R->t_calc = L->dev_ms[1] * 1e-3;
   R->t_cs_dev = L->dev_ms[2] * 1e-3;
   R->t_cs = R->t_cs_dev + R->t_cs_host;
   double dir_ms[3];
   amr_dev_read_exchange_timing(R->dev, dir_ms);
   for (int d = 0; d < 3; d++)
      R->t_comm_dir[d] = dir_ms[d] * 1e-3;
   amr_dev_memory(R->dev, &R->mem_bytes, &R->mem_buffers);
   if (L->state != 2) {
      P->end_time = L->sim_time;
      P->num_tsteps = L->ts - 1;
      R->t_all = amr_now() - L->t_start;
   }
   free(L->sums);
   L->sums = NULL;
   amr_driver_set_sum_log(R, NULL);
}

/* ref/driver.c:37-165 */
int amr_driver(amr_run *R)
{
   int rc;
   amr_driver_begin(R);
   while (!(rc = amr_driver_step(R)))
      ;
   amr_driver_end(R);
   return rc == 2;
}

/* Final-state dumps in the formats of the golden-vector harness. */
int amr_dump_final(amr_run *R)
{
   return amr_dump_final_to(R, getenv("AMR_DUMP_CELLS"), getenv("AMR_DUMP_BSUMS"),
                            getenv("AMR_DUMP_SAMPLE"));
}

/* The cell sample of AMR_DUMP_SAMPLE: AMR_SAMPLE_CELLS cells, each drawn as
 * (sorted-list index, variable, i, j, k) with ghosts included, from a
 * splitmix64 sequence seeded AMR_SAMPLE_SEED -- the same draws as the
 * golden-vector harness of the reference build makes. */
#define AMR_SAMPLE_CELLS 16384
#define AMR_SAMPLE_SEED 0x5eed5eed5eedULL

static uint64_t splitmix64(uint64_t *x)
{
   uint64_t z = (*x += 0x9e3779b97f4a7c15ULL);
   z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
   z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
   return z ^ (z >> 31);
}

typedef struct { int32_t in, v, i, j, k, pad; double val; } sample_rec;

int amr_dump_final_to(amr_run *R, const char *pc, const char *pb, const char *ps)
{
   const amr_params *P = &R->P;
   if (!pc && !pb && !ps)
      return 0;
   int gx = P->nx + 2, gy = P->ny + 2, gz = P->nz + 2;
   size_t slab = (size_t) gx * gy * gz;
   int na = amr_n_active(R);
   FILE *fc = pc ? fopen(pc, "wb") : NULL;
   FILE *fb = pb ? fopen(pb, "wb") : NULL;
   int32_t hdr[8] = { 0x414d5231, P->nx, P->ny, P->nz, P->num_vars, na, P->num_refine, 0 };
   if (fc)
      fwrite(hdr, sizeof(hdr), 1, fc);
   hdr[0] = 0x414d5332;
   if (fb)
      fwrite(hdr, sizeof(hdr), 1, fb);
   sample_rec *smp = NULL;
   if (ps && na > 0) {
      uint64_t x = AMR_SAMPLE_SEED;
      smp = (sample_rec *) amr_xmalloc(sizeof(sample_rec) * AMR_SAMPLE_CELLS);
      for (int s = 0; s < AMR_SAMPLE_CELLS; s++) {
         smp[s].in = (int32_t) (splitmix64(&x) % (uint64_t) na);
         smp[s].v = (int32_t) (splitmix64(&x) % (uint64_t) P->num_vars);
         smp[s].i = (int32_t) (splitmix64(&x) % (uint64_t) gx);
         smp[s].j = (int32_t) (splitmix64(&x) % (uint64_t) gy);
         smp[s].k = (int32_t) (splitmix64(&x) % (uint64_t) gz);
      }
   }
   for (int in = 0; in < na; in++) {
      int slot = R->sorted_slot[in];
      const amr_block *b = &R->blocks[slot];
      int64_t number = b->number;
      int32_t ls[2] = { b->level, slot };
      amr_dev_check(R, amr_dev_download_block(R->dev, slot, R->cellbuf), "download");
      if (fc) {
         fwrite(&number, sizeof(number), 1, fc);
         fwrite(ls, sizeof(ls), 1, fc);
         fwrite(R->cellbuf, sizeof(double), slab * P->num_vars, fc);
      }
      if (fb) {
         fwrite(&number, sizeof(number), 1, fb);
         fwrite(ls, sizeof(ls), 1, fb);
         for (int v = 0; v < P->num_vars; v++) {
            double s[2] = { 0.0, 0.0 };
            const double *a = R->cellbuf + slab * v;
            for (int i = 0; i < gx; i++)
               for (int j = 0; j < gy; j++)
                  for (int k = 0; k < gz; k++) {
                     int inner = i >= 1 && i <= P->nx && j >= 1 && j <= P->ny &&
                                 k >= 1 && k <= P->nz;
                     s[inner ? 0 : 1] += a[((size_t) i * gy + j) * gz + k];
                  }
            fwrite(s, sizeof(s), 1, fb);
         }
      }
      if (smp)
         for (int s = 0; s < AMR_SAMPLE_CELLS; s++)
            if (smp[s].in == in)
               smp[s].val = R->cellbuf[slab * smp[s].v + ((size_t) smp[s].i * gy + smp[s].j) * gz
                                       + smp[s].k];
   }
   if (fc) fclose(fc);
   if (fb) fclose(fb);
   if (ps) {
      FILE *f = fopen(ps, "wb");
      if (f) {
         hdr[0] = 0x414d5333;
         hdr[7] = smp ? AMR_SAMPLE_CELLS : 0;
         fwrite(hdr, sizeof(hdr), 1, f);
         if (smp)
            fwrite(smp, sizeof(sample_rec), AMR_SAMPLE_CELLS, f);
         fclose(f);
      }
   }
   free(smp);
   return 0;
}
