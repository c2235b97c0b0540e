/*
 * report.c -- the end-of-run report of ref/profile.c:37-1300, for one rank:
 *   report_perf & 1   results.yaml           (ref/profile.c:52-470)
 *   report_perf & 2   results.txt            (ref/profile.c:472-877)
 *   report_perf & 4   the same text on stdout, with a few more lines, then
 *                     the Summary line       (ref/profile.c:880-1298)
 *
 * Every line, label and number format is the reference's, so scripts that
 * read its reports read these.  With one rank, ave = min = max and the
 * standard deviation is 0.  The values:
 *   times     the GPU's device time in the slots of the phases it replaces
 *             (calc = stencil kernels, comm = exchange kernels, per
 *             direction; gridsum = reduction kernels + the host's
 *             conservation check), host wall time for the rest (total,
 *             main, init, refinement and its parts);
 *   counters  counted as the reference counts them: faces exchanged per
 *             comm call (same level 2 per pair, coarse/fine 2 per quarter
 *             pair, 1 per boundary face), refinement steps, blocks split
 *             and reformed, floating-point operations of the stencils;
 *   zeros     what only exists between ranks or in phases this driver does
 *             not have: messages, bytes, packing, waits, partners, load
 *             balance, redistribution, plots;
 *   memory    the device context's buffers (amr_dev_memory) where the
 *             reference reports its own mallocs.
 */
#include <math.h>
#include <string.h>

#include "amr.h"

static const char *kVersion = "1.7.0 (modified)";

/* ref/profile.c:913-958 object descriptions, by object type. */
static const char *object_kind(int type)
{
   static const struct { int type; const char *text; } k[] = {
      {0, "surface of a rectangle"}, {1, "volume of a rectangle"},
      {2, "surface of a spheroid"}, {3, "volume of a spheroid"},
      {4, "surface of x+ hemispheroid"}, {5, "volume of x+ hemispheroid"},
      {6, "surface of x- hemispheroid"}, {7, "volume of x- hemispheroid"},
      {8, "surface of y+ hemispheroid"}, {9, "volume of y+ hemispheroid"},
      {10, "surface of y- hemispheroid"}, {11, "volume of y- hemispheroid"},
      {12, "surface of z+ hemispheroid"}, {13, "volume of z+ hemispheroid"},
      {14, "surface of z- hemispheroid"}, {15, "volume of z- hemispheroid"},
      {20, "surface of x axis cylinder"}, {21, "volune of x axis cylinder"},
      {22, "surface of y axis cylinder"}, {23, "volune of y axis cylinder"},
      {24, "surface of z axis cylinder"}, {25, "volune of z axis cylinder"},
   };
   for (size_t i = 0; i < sizeof(k) / sizeof(k[0]); i++)
      if (k[i].type == type)
         return k[i].text;
   return NULL;
}

/* The report's values, named by what they measure. */
typedef struct {
   double total, main_t, init_t, calc, comm, comm_dir[3], cs, cs_red, cs_calc;
   double refine, refine_init, ref_compare, ref_mark, ref_split, ref_coarsen, ref_misc;
   double mem, mem_buffers, mem_init, mem_buffers_init;
   double fp_adds, fp_muls, fp_divs;
   double face_same[3], face_diff[3], face_bc[3];
   double split, reformed, parents;
   double tsum_comm, tsum_calc, tsum_cs, tsum_refine;
   double total_dots;        /* load-balance dots: none on one rank */
   int nrs, nlbs, nrrs, nps, num_over, tot_over, max_groups, max_dots;
} amr_rep;

static void collect(const amr_run *R, amr_rep *r)
{
   const amr_params *P = &R->P;
   memset(r, 0, sizeof(*r));
   double cells = (double) P->nx * P->ny * P->nz;
   double updates = (double) R->total_blocks * cells * P->num_vars;
   r->total = R->t_all;
   r->main_t = R->t_main;
   r->init_t = R->t_init;
   r->calc = R->t_calc;
   r->comm = R->t_comm;
   for (int d = 0; d < 3; d++) {
      r->comm_dir[d] = R->t_comm_dir[d];
      r->face_same[d] = (double) R->face_same[d];
      r->face_diff[d] = (double) R->face_diff[d];
      r->face_bc[d] = (double) R->face_bc[d];
   }
   r->cs_red = R->t_cs_dev;
   r->cs_calc = R->t_cs_host;
   r->cs = r->cs_red + r->cs_calc;
   r->refine = R->t_refine_all;
   r->refine_init = R->t_refine_init;
   r->ref_compare = R->t_ref_compare;
   r->ref_mark = R->t_ref_mark;
   r->ref_split = R->t_ref_split;
   r->ref_coarsen = R->t_ref_coarsen;
   r->ref_misc = R->t_refine_all - (R->t_ref_compare + R->t_ref_mark + R->t_ref_split +
                                    R->t_ref_coarsen);
   if (r->ref_misc < 0.0)
      r->ref_misc = 0.0;
   r->mem = (double) R->mem_bytes;
   r->mem_buffers = (double) R->mem_buffers;
   r->mem_init = (double) R->mem_bytes_init;
   r->mem_buffers_init = (double) R->mem_buffers_init;
   /* ref/stencil.c:100-101,142-143: per stencil call, one divide and 6 or
    * 26 adds per interior cell of every active block */
   r->fp_divs = updates;
   r->fp_adds = (P->stencil == 7 ? 6.0 : 26.0) * updates;
   r->fp_muls = 0.0;
   if (P->stencil == 0) {
      /* counted per stage by the driver, plus the clamps: each divide costs
       * a divide and two adds, each scale a multiply and an add
       * (ref/stencil.c:966-981) */
      int64_t c[2];
      if (amr_dev_read_check_counts(R->dev, c) != AMR_OK)
         amr_fatal("ERROR: device read_check_counts failed: %s\n", amr_dev_last_error(R->dev));
      r->fp_divs = R->total_fp_divs + (double) c[0];
      r->fp_adds = R->total_fp_adds + 2.0 * (double) c[0] + (double) c[1];
      r->fp_muls = R->total_fp_muls + (double) c[1];
   }
   r->split = (double) R->num_refined;
   r->reformed = (double) R->num_reformed;
   r->parents = (double) R->num_parents;
   /* per-timestep sums (ref/driver.c:143-158): on one rank min = max = total */
   r->tsum_comm = R->t_comm;
   r->tsum_calc = R->t_calc;
   r->tsum_cs = r->cs;
   r->tsum_refine = R->t_refine_all - R->t_refine_init;
   r->nrs = R->nrs;
   r->max_groups = 1;
}

/* ---------------------------------------------------------------- yaml */

static void y3(FILE *f, int indent, const char *name, double v)
{
   static const char *sfx[3] = { "ave", "min", "max" };
   for (int i = 0; i < 3; i++)
      fprintf(f, "%*s%s_%s: %lf\n", indent, "", name, sfx[i], v);
}

/* The eight timers and nine counters of a communication block. */
static const char *kCommTimers[8] = { "post_recv", "pack_faces", "send_mess", "exch_same",
                                      "exch_diff", "apply_bc", "wait_time", "unpack_faces" };
static const char *kCommCounts[9] = { "mess_recv", "byte_recv", "face_recv",
                                      "mess_send", "byte_send", "face_send",
                                      "face_exch_same", "face_exch_diff", "face_bc_apply" };

static void yaml_comm_counts(FILE *f, int indent, const char *prefix, double same,
                             double diff, double bc)
{
   char name[64];
   for (int i = 0; i < 9; i++) {
      double v = i == 6 ? same : i == 7 ? diff : i == 8 ? bc : 0.0;
      snprintf(name, sizeof(name), "%s_%s", prefix, kCommCounts[i]);
      y3(f, indent, name, v);
   }
}

static void write_yaml(const amr_run *R, const amr_rep *r, FILE *f)
{
   const amr_params *P = &R->P;
   static const char axis[3] = { 'x', 'y', 'z' };
   char name[64];
   fprintf(f, "code: miniAMR\nversion: %s\nranks: %d\n", kVersion, 1);
   const struct { const char *k; int v; } head[] = {
      {"npx", P->npx}, {"npy", P->npy}, {"npz", P->npz},
      {"init_block_x", P->init_block_x}, {"init_block_y", P->init_block_y},
      {"init_block_z", P->init_block_z}, {"x_block_size", P->nx}, {"y_block_size", P->ny},
      {"z_block_size", P->nz}, {"reorder", P->reorder}, {"permute", P->permute},
      {"max_blocks_allowed", P->max_num_blocks}, {"code", P->code},
      {"num_refine", P->num_refine}, {"block_change", P->block_change},
      {"refine_ghost", P->refine_ghost}, {"uniform_refine", P->uniform_refine},
      {"num_objects", P->num_objects},
   };
   for (size_t i = 0; i < sizeof(head) / sizeof(head[0]); i++)
      fprintf(f, "%s: %d\n", head[i].k, head[i].v);
   for (int i = 0; i < P->num_objects; i++) {
      const amr_object *op = &P->objects[i];
      fprintf(f, "obj%dtype: %d\nobj%dbounce: %d\n", i, op->type, i, op->bounce);
      const char *vec[4] = { "center", "move", "size", "inc" };
      const double *val[4] = { op->cen, op->move, op->size, op->inc };
      for (int q = 0; q < 4; q++)
         for (int d = 0; d < 3; d++)
            fprintf(f, "obj%d%s_%c: %lf\n", i, vec[q], axis[d], val[q][d]);
   }
   const struct { const char *k; int v; } run[] = {
      {"num_tsteps", P->num_tsteps}, {"stages_per_timestep", P->stages_per_ts},
      {"nonblocking", P->nonblocking}, {"checksum_freq", P->checksum_freq},
      {"refine_freq", P->refine_freq}, {"lb_opt", P->lb_opt}, {"inbalance", P->inbalance},
      {"lb_method", P->lb_method}, {"plot_freq", P->plot_freq}, {"num_vars", P->num_vars},
      {"stencil", P->stencil}, {"comm_vars", P->comm_vars}, {"error_tol", P->error_tol},
      {"send_faces", P->send_faces},
   };
   for (size_t i = 0; i < sizeof(run) / sizeof(run[0]); i++)
      fprintf(f, "%s: %d\n", run[i].k, run[i].v);

   y3(f, 0, "total_time", r->total);
   y3(f, 0, "memory_used", r->mem);
   y3(f, 0, "compute_time", r->calc);
   double gflops = (r->fp_adds + r->fp_muls + r->fp_divs) / (r->calc * 1024.0 * 1024.0 * 1024.0);
   fprintf(f, "total_gflops: %lf\nave_gflops: %lf\n", gflops, gflops);

   y3(f, 0, "total_comm", r->comm);
   for (int i = 0; i < 8; i++) {
      snprintf(name, sizeof(name), "total_%s", kCommTimers[i]);
      y3(f, 3, name, 0.0);
   }
   for (int i = 0; i < 2; i++) {
      const char *n = i ? "comm_partners_unique" : "comm_partners_total";
      fprintf(f, "   %s: %lf\n   %s_min: %lf\n   %s_max: %lf\n", n, 0.0, n, 0.0, n, 0.0);
   }
   double same = 0, diff = 0, bc = 0;
   for (int d = 0; d < 3; d++) {
      same += r->face_same[d];
      diff += r->face_diff[d];
      bc += r->face_bc[d];
   }
   yaml_comm_counts(f, 3, "total", same, diff, bc);
   for (int d = 0; d < 3; d++) {
      snprintf(name, sizeof(name), "%c_comm", axis[d]);
      y3(f, 3, name, r->comm_dir[d]);
      for (int i = 0; i < 8; i++) {
         snprintf(name, sizeof(name), "%c_%s", axis[d], kCommTimers[i]);
         y3(f, 6, name, 0.0);
      }
      fprintf(f, "      %c_comm_partners: %lf\n      %c_comm_partners_min: %d\n"
                 "      %c_comm_partners_max: %d\n", axis[d], 0.0, axis[d], 0, axis[d], 0);
      snprintf(name, sizeof(name), "%c", axis[d]);
      yaml_comm_counts(f, 6, name, r->face_same[d], r->face_diff[d], r->face_bc[d]);
   }

   y3(f, 0, "gridsum_time", r->cs);
   y3(f, 3, "gridsum_reduce", r->cs_red);
   y3(f, 3, "gridsum_calc", r->cs_calc);
   y3(f, 0, "refine_time", r->refine);
   double stages = (double) P->num_tsteps * P->stages_per_ts;
   fprintf(f, "   total_blocks_ts_ave: %lf\n   total_blocks_ts_min: %lld\n"
              "   total_blocks_ts_max: %lld\n", (double) R->total_blocks / stages,
           (long long) R->nb_min, (long long) R->nb_max);
   const struct { const char *k; double v; int indent; } ref[] = {
      {"blocks_split", r->split, 3}, {"blocks_reformed", r->reformed, 3},
      {"blocks_moved_tot", 0, 3}, {"blocks_moved_lb", 0, 3}, {"blocks_moved_redist", 0, 3},
      {"blocks_moved_coarsen", 0, 3}, {"time_compare_obj", r->ref_compare, 3},
      {"time_mark_refine", r->ref_mark, 3}, {"time_comm_block1", 0, 3},
      {"time_split_block", r->ref_split, 3}, {"time_comm_block2", 0, 3},
      {"time_sync", 0, 3}, {"time_misc", r->ref_misc, 3},
      {"time_total_coarsen", r->ref_coarsen, 3}, {"time_coarsen", r->ref_coarsen, 6},
      {"time_coarsen_pack", 0, 6}, {"time_coarsen_move", 0, 6},
      {"time_coarsen_unpack", 0, 6}, {"time_total_redist", 0, 3},
      {"time_redist_choose", 0, 6}, {"time_redist_pack", 0, 6}, {"time_redist_move", 0, 6},
      {"time_redist_unpack", 0, 6}, {"time_total_load_bal", 0, 3},
      {"time_load_bal_sort", 0, 6}, {"time_lb_move_dots", 0, 6},
      {"time_lb_move_blocks", 0, 6}, {"time_lb_mb_pack", 0, 9}, {"time_lb_mb_move", 0, 9},
      {"time_lb_mb_unpack", 0, 9}, {"time_lb_mb_misc", 0, 9}, {"plot_time", 0, 0},
   };
   for (size_t i = 0; i < sizeof(ref) / sizeof(ref[0]); i++)
      y3(f, ref[i].indent, ref[i].k, ref[i].v);
}

/* ---------------------------------------------------------------- text */

/* "label: ave stddev min max" with the colon in column 28. */
static void row4(FILE *f, int indent, const char *label, double v)
{
   fprintf(f, "%*s%-*s: %lf %lf %lf %lf\n", indent, "", 28 - indent, label, v, 0.0, v, v);
}

/* "<what>: ave, stddev, min, max (sec): ..." */
static void stat_line(FILE *f, const char *what, double v, const char *tail)
{
   fprintf(f, "%s: ave, stddev, min, max (sec): %lf %lf %lf %lf\n%s", what, v, 0.0, v, v, tail);
}

static void rule(FILE *f, const char *title_line)
{
   fprintf(f, "---------------------------------------------\n%s\n"
              "---------------------------------------------\n\n", title_line);
}

static void write_text(const amr_run *R, const amr_rep *r, FILE *f, int on_stdout)
{
   const amr_params *P = &R->P;
   fprintf(f, "\n ================ Start report ===================\n\n");
   fprintf(f, "          Mantevo miniAMR\n          version %s\n\n", kVersion);
   fprintf(f, "Run on %d ranks arranged in a %d x %d x %d grid\n", 1, P->npx, P->npy, P->npz);
   fprintf(f, "initial blocks per rank %d x %d x %d\n", P->init_block_x, P->init_block_y,
           P->init_block_z);
   fprintf(f, "block size %d x %d x %d\n", P->nx, P->ny, P->nz);
   fprintf(f, P->reorder ? "Initial ranks arranged by RCB across machine\n\n"
                         : "Initial ranks arranged as a grid across machine\n\n");
   if (on_stdout)
      fprintf(f, "change_dir %d group_blocks %d limit_move %d\n", P->change_dir,
              P->group_blocks, P->limit_move);
   if (P->permute)
      fprintf(f, "Order of exchanges permuted\n");
   fprintf(f, "Maximum number of blocks per rank is %d\n", P->max_num_blocks);
   if (P->code)
      fprintf(f, "Code set to code %d\n", P->code);
   fprintf(f, "Number of levels of refinement is %d\n", P->num_refine);
   fprintf(f, "Blocks can change by %d levels per refinement step\n", P->block_change);
   if (P->refine_ghost)
      fprintf(f, "Ghost cells will be used determine is block is refined\n");
   if (P->uniform_refine) {
      fprintf(f, "\nBlocks will be uniformly refined\n");
   } else {
      fprintf(f, "\nBlocks will be refined by %d objects\n\n", P->num_objects);
      for (int i = 0; i < P->num_objects; i++) {
         const amr_object *op = &P->objects[i];
         const char *kind = object_kind(op->type);
         if (kind)
            fprintf(f, "Object %d is the %s\n", i, kind);
         fprintf(f, op->bounce == 1 ? "Object center will bounce off of walls\n"
                  : op->bounce == 2 ? "Object edge will bounce off of walls\n"
                                    : "Object may leave mesh\n");
         fprintf(f, "Center starting at %lf %lf %lf\n", op->orig_cen[0], op->orig_cen[1],
                 op->orig_cen[2]);
         fprintf(f, "Center end at %lf %lf %lf\n", op->cen[0], op->cen[1], op->cen[2]);
         double scale[3];
         for (int d = 0; d < 3; d++)
            scale[d] = (double) R->mesh_size[d] * (d == 0 ? P->nx : d == 1 ? P->ny : P->nz);
         if (P->use_time) {
            double delta = P->end_time / (double) P->num_tsteps;
            fprintf(f, "Velocity of %lf %lf %lf\n", op->orig_move[0], op->orig_move[1],
                    op->orig_move[2]);
            fprintf(f, "   Rate relative to smallest cell size %lf %lf %lf\n",
                    delta * op->orig_move[0] * scale[0], delta * op->orig_move[1] * scale[1],
                    delta * op->orig_move[2] * scale[2]);
         } else {
            fprintf(f, "Moving at %lf %lf %lf per timestep\n", op->orig_move[0],
                    op->orig_move[1], op->orig_move[2]);
            fprintf(f, "   Rate relative to smallest cell size %lf %lf %lf\n",
                    op->orig_move[0] * scale[0], op->orig_move[1] * scale[1],
                    op->orig_move[2] * scale[2]);
         }
         fprintf(f, "Initial size %lf %lf %lf\n", op->orig_size[0], op->orig_size[1],
                 op->orig_size[2]);
         fprintf(f, "Final size %lf %lf %lf\n", op->size[0], op->size[1], op->size[2]);
         fprintf(f, "Size increasing %lf %lf %lf per timestep\n", op->inc[0], op->inc[1],
                 op->inc[2]);
         fprintf(f, "   Rate relative to smallest cell size %lf %lf %lf\n\n",
                 op->inc[0] * scale[0], op->inc[1] * scale[1], op->inc[2] * scale[2]);
      }
   }
   if (P->use_time)
      fprintf(f, "\nTime %lf in %d timesteps\n", P->end_time, P->num_tsteps);
   else
      fprintf(f, "\nNumber of timesteps is %d\n", P->num_tsteps);
   fprintf(f, "Communicaion/computation stages per timestep is %d\n", P->stages_per_ts);
   fprintf(f, P->nonblocking ? "Communication will be performed with nonblocking sends\n"
                             : "Communication will be performed with blocking sends\n");
   fprintf(f, "Will perform checksums every %d stages\n", P->checksum_freq);
   fprintf(f, "Will refine every %d timesteps\n", P->refine_freq);
   static const char *lb[4] = {
      "Load balance by RCB (Recursive Coordinate Bisection)\n",
      "Load balance by Morton Space Filling Curve\n",
      "Load balance by Peano Hilbert Space Filling Curve\n",
      NULL };
   fprintf(f, "%s", P->lb_method >= 0 && P->lb_method < 3 ? lb[P->lb_method]
           : on_stdout ? "Load balance by truncated Hilbert style Space Filling Curve\n"
                       : "Load balance by truncated Hilbert Space Filling Curve\n");
   if (P->lb_opt == 0)
      fprintf(f, "Load balance will not be performed\n");
   else
      fprintf(f, "Load balance when inbalanced by %d%%\n", P->inbalance);
   if (P->lb_opt == 2)
      fprintf(f, "Load balance at each phase of refinement step\n");
   if (P->plot_freq)
      fprintf(f, "Will plot results every %d timesteps\n", P->plot_freq);
   else
      fprintf(f, "Will not plot results\n");
   if (P->stencil)
      fprintf(f, "Calculate on %d variables with %d point stencil\n", P->num_vars, P->stencil);
   else
      fprintf(f, "Calculate on %d variables with variable stencils\n", P->num_vars);
   fprintf(f, "Communicate %d variables at a time\n", P->comm_vars);
   fprintf(f, "Error tolorance for variable sums is 10^(-%d)\n", P->error_tol);
   if (P->send_faces)
      fprintf(f, "Will send data from each face seperately\n");

   fprintf(f, "\nTotal time for test: ave, std, min, max (sec): %lf %lf %lf %lf\n\n",
           r->total, 0.0, r->total, r->total);
   const struct { const char *what; double v; int calls; } mem[6] = {
      {"Number of malloc calls", r->mem_buffers, 1}, {"Amount malloced", r->mem, 0},
      {"Malloc calls in init", r->mem_buffers_init, 1},
      {"Amount malloced in init", r->mem_init, 0},
      {"Malloc calls in timestepping", r->mem_buffers - r->mem_buffers_init, 1},
      {"Amount malloced in timestepping", r->mem - r->mem_init, 0},
   };
   for (int i = 0; i < 6; i++)
      fprintf(f, "\n%s: ave, std, min, max%s: %lf %lf %lf %lf\n%s", mem[i].what,
              mem[i].calls ? " (sec)" : "", mem[i].v, 0.0, mem[i].v, mem[i].v,
              i == 5 ? "\n" : "");
   if (on_stdout) {
      stat_line(f, "Main (parse, allocate) Time", r->main_t, "\n");
      stat_line(f, "Initailize Time", r->init_t, "\n");
   }

   rule(f, "          Computational Performance");
   double fp = r->fp_adds + r->fp_muls + r->fp_divs;
   double gflops = fp / (r->calc * 1024.0 * 1024.0 * 1024.0);
   stat_line(f, "     Time", r->calc, "\n");
   fprintf(f, "     total GFLOPS:             %lf\n", gflops);
   fprintf(f, "     Average GFLOPS per rank:  %lf\n\n", gflops);
   fprintf(f, "     Total floating point ops: %lf\n\n", fp);
   fprintf(f, "        Adds:                  %lf\n", r->fp_adds);
   fprintf(f, "        Muls:                  %lf\n", r->fp_muls);
   fprintf(f, "        Divides:               %lf\n\n", r->fp_divs);
   if (on_stdout)
      fprintf(f, "     Sum of min/max compute times per ts: %lf %lf\n\n", r->tsum_calc,
              r->tsum_calc);

   rule(f, "           Interblock communication");
   stat_line(f, "     Time", r->comm, "\n");
   static const char *block_title[4] = {
      "\nTotal communication:\n\n", "\nX direction communication statistics:\n\n",
      "\nY direction communication statistics:\n\n",
      "\nZ direction communication statistics:\n\n" };
   static const char *timer_label[9] = {
      "Total", "Post IRecv", "Pack faces", "Send messages", "Exchange same level",
      "Exchange diff level", "Apply BC", "Wait time", "Unpack faces" };
   static const char *count_label[9] = {
      "Messages received", "Bytes received", "Faces received", "Messages sent",
      "Bytes sent", "Faces sent", "Faces exchanged same", "Faces exchanged diff",
      "Faces with BC applied" };
   for (int i = 0; i < 4; i++) {
      double t = 0.0, same = 0.0, diff = 0.0, bc = 0.0;
      for (int d = 0; d < 3; d++)
         if (!i || d == i - 1) {
            t += r->comm_dir[d];
            same += r->face_same[d];
            diff += r->face_diff[d];
            bc += r->face_bc[d];
         }
      fprintf(f, "%s", block_title[i]);
      fprintf(f, "                              average    stddev  minimum  maximum\n");
      for (int k = 0; k < 9; k++) {
         row4(f, 5, timer_label[k], k == 0 ? t : 0.0);
         if (k == 8)
            fprintf(f, "\n");
      }
      if (!i) {
         static const char *pl[6] = { "Comm partners total ave", "Comm partners total min",
                                      "Comm partners total max", "Comm partners uniq ave ",
                                      "Comm partners uniq min ", "Comm partners uniq max " };
         for (int k = 0; k < 6; k++)
            row4(f, 5, pl[k], 0.0);
      } else {
         static const char *pl[3] = { "Comm partners average", "Comm partners minimum",
                                      "Comm partners maximum" };
         for (int k = 0; k < 3; k++)
            row4(f, 5, pl[k], 0.0);
      }
      for (int k = 0; k < 9; k++)
         row4(f, 5, count_label[k], k == 6 ? same : k == 7 ? diff : k == 8 ? bc : 0.0);
   }
   if (on_stdout) {
      fprintf(f, "\n     Sum of min/max communicate times per ts: %lf %lf\n", r->tsum_comm,
              r->tsum_comm);
      fprintf(f, "\n     Sum of min/max calc and comm times per ts: %lf %lf\n",
              r->tsum_calc + r->tsum_comm, r->tsum_calc + r->tsum_comm);
   }

   fprintf(f, "\n");
   rule(f, "             Gridsum performance");
   stat_line(f, "     Time", r->cs, "\n");
   stat_line(f, "        red ", r->cs_red, "\n");
   stat_line(f, "        calc", r->cs_calc, "\n");
   fprintf(f, "     total number:             %d\n", (int) R->total_red);
   fprintf(f, "     number per timestep:      %d\n\n", P->num_vars);
   if (on_stdout)
      fprintf(f, "     Sum of min/max gridsum times per ts: %lf %lf\n\n", r->tsum_cs,
              r->tsum_cs);

   rule(f, "               Mesh Refinement");
   stat_line(f, "     Time", r->refine, "\n");
   fprintf(f, "     Number of refinement steps: %d\n\n", r->nrs);
   fprintf(f, "     Number of load balance steps: %d\n\n", r->nlbs);
   fprintf(f, "     Number of redistributing steps: %d\n\n", r->nrrs);
   double stages = (double) P->num_tsteps * P->stages_per_ts;
   fprintf(f, "     Total blocks           : %lld\n", (long long) R->total_blocks);
   fprintf(f, "     Blocks/timestep ave, min, max : %lf %lld %lld\n",
           (double) R->total_blocks / stages, (long long) R->nb_min, (long long) R->nb_max);
   fprintf(f, "     Max blocks on a processor at any time: %d\n", R->local_max_b);
   fprintf(f, "     total blocks split     : %lf\n", r->split);
   fprintf(f, "     total blocks reformed  : %lf\n\n", r->reformed);
   fprintf(f, "     total blocks moved     : %lf\n", 0.0);
   fprintf(f, "     total moved load bal   : %lf\n", 0.0);
   fprintf(f, "     total moved redistribut: %lf\n", 0.0);
   fprintf(f, "     total moved coasening  : %lf\n\n", 0.0);
   fprintf(f, "     parents ave, min, max  : %lf %d %d\n", r->parents, (int) r->parents,
           (int) r->parents);
   fprintf(f, "     max dots ave, min, max : %lf %d %d\n", (double) r->max_dots, r->max_dots,
           r->max_dots);
   fprintf(f, "     ave dots used          : %lf\n\n", r->total_dots / r->nlbs);
   if (on_stdout) {
      fprintf(f, "     number of times more sets of blocks than ranks: %d\n", r->num_over);
      fprintf(f, "     total sets of blocks over num of ranks: %d\n", r->tot_over);
      fprintf(f, "     max number of sets of blocks on a rank: %d\n\n", r->max_groups);
   }
   fprintf(f, "                              average    stddev  minimum  maximum\n");
   fprintf(f, "     Per processor:\n");
   const struct { const char *label; double v; } per[6] = {
      {"total blocks split", r->split}, {"total blocks reformed", r->reformed},
      {"Total blocks moved", 0}, {"Blocks moved load bal", 0},
      {"Blocks moved redistribu", 0}, {"Blocks moved coarsening", 0} };
   for (int i = 0; i < 6; i++)
      row4(f, 5, per[i].label, per[i].v);
   fprintf(f, "     Time:\n");
   if (on_stdout) {
      row4(f, 8, "initial refine", r->refine_init);
      fprintf(f, "\n");
   }
   const struct { const char *label; double v; int indent; int stdout_only; } rt[] = {
      {"compare objects", r->ref_compare, 8, 0}, {"mark refine/coarsen", r->ref_mark, 8, 0},
      {"communicate block 1", 0, 8, 0}, {"split blocks", r->ref_split, 8, 0},
      {"communicate block 2", 0, 8, 0}, {"sync time", 0, 8, 0},
      {"time for groups", 0, 8, 1}, {"misc time", r->ref_misc, 8, 0},
      {"total coarsen blocks", r->ref_coarsen, 8, 0}, {"coarsen blocks", r->ref_coarsen, 11, 0},
      {"pack blocks", 0, 11, 0}, {"move blocks", 0, 11, 0}, {"unpack blocks", 0, 11, 0},
      {"total redistribute", 0, 8, 0}, {"choose blocks", 0, 11, 0}, {"pack blocks", 0, 11, 0},
      {"move blocks", 0, 11, 0}, {"unpack blocks", 0, 11, 0},
      {"total load balance", 0, 8, 0}, {"sort", 0, 11, 0}, {"move dots back", 0, 11, 0},
      {"move blocks total", 0, 11, 0}, {"pack blocks", 0, 14, 0}, {"move blocks", 0, 14, 0},
      {"unpack blocks", 0, 14, 0}, {"misc", 0, 14, 0},
   };
   for (size_t i = 0; i < sizeof(rt) / sizeof(rt[0]); i++)
      if (on_stdout || !rt[i].stdout_only)
         row4(f, rt[i].indent, rt[i].label, rt[i].v);
   fprintf(f, "\n");
   if (on_stdout)
      fprintf(f, "     Sum of min/max refinement times per ts: %lf %lf\n\n", r->tsum_refine,
              r->tsum_refine);
   rule(f, "                   Plot");
   stat_line(f, "     Time", 0.0, "\n");
   fprintf(f, "     Number of plot steps: %d\n", r->nps);
   fprintf(f, "\n ================== End report ===================\n");
}

/* ref/profile.c:37-1300 for one rank. */
void amr_report(amr_run *R)
{
   const amr_params *P = &R->P;
   amr_rep r;
   collect(R, &r);
   if (P->report_perf & 1) {
      FILE *f = fopen("results.yaml", "w");
      if (f) {
         write_yaml(R, &r, f);
         fclose(f);
      }
   }
   if (P->report_perf & 2) {
      FILE *f = fopen("results.txt", "w");
      if (f) {
         write_text(R, &r, f, 0);
         fclose(f);
      }
   }
   if (P->report_perf & 4) {
      write_text(R, &r, stdout, 1);
      double stages = (double) P->num_tsteps * P->stages_per_ts;
      printf("Summary: ranks %d ts %d time %lf calc %lf comm %lf red %lf refine %lf blocks/ts %lf max_blocks %d\n",
             1, P->num_tsteps, r.total, r.calc, r.comm, r.cs, r.refine,
             stages > 0 ? ((double) R->total_blocks) / stages : 0.0, R->local_max_b);
      fflush(NULL);
   }
}
