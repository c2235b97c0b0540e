/*
 * params.c -- command line of the drop-in driver.
 *
 * Same flags, argument counts, defaults and validation as the reference
 * (ref/main.c:57-191 parsing, ref/param.h:27-66 defaults,
 * ref/main.c:657-723 check_input), so existing miniAMR command lines run
 * unchanged.  The error messages are kept word for word because scripts
 * around miniAMR match on them.  Flags that select features outside this
 * build's scope (--hilbert/--trunc_hilbert
 * initial orderings, plotting) are accepted by the parser and rejected by
 * amr_check_input with an explicit message rather than silently ignored.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "amr.h"

static void set_defaults(amr_params *P)
{
   memset(P, 0, sizeof(*P));
   P->max_num_blocks = 500;
   P->num_refine = 5;
   P->nx = P->ny = P->nz = 10;
   P->num_vars = 40;
   P->init_block_x = P->init_block_y = P->init_block_z = 1;
   P->reorder = -1;
   P->npx = P->npy = P->npz = 1;
   P->refine_freq = 5;
   P->error_tol = 8;
   P->num_tsteps = 20;
   P->stages_per_ts = 20;
   P->checksum_freq = 5;
   P->stencil = 7;
   P->report_perf = 12;
   P->lb_opt = 1;
   P->nonblocking = 1;
   P->seed = 1;
}

void amr_print_help(void)
{
   static const char *lines[] = {
      "(Optional) command line input is of the form: ",
      "",
      "--nx, --ny, --nz - block size in x, y, z (even && > 0)",
      "--init_x, --init_y, --init_z - initial blocks in x, y, z (> 0)",
      "--reorder - ordering of blocks if initial number > 1",
      "--npx, --npy, --npz - rank grid (one rank: 1 1 1)",
      "--max_blocks - maximum number of blocks per rank",
      "--num_refine - (>= 0) number of levels of refinement",
      "--block_change - (>= 0) number of levels a block can change in a timestep",
      "--uniform_refine - if 1, then grid is uniformly refined",
      "--refine_freq - frequency (in timesteps) of checking for refinement",
      "--inbalance - percentage inbalance to trigger load balance",
      "--lb_opt - load balancing - 0 = none, 1 = each refine, 2 = each refine phase",
      "--num_vars - number of variables (> 0)",
      "--comm_vars - number of vars to communicate together",
      "--num_tsteps - number of timesteps (> 0)",
      "--time - time to run problem with delta by object speed (> 0.0)",
      "--stages_per_ts - number of comm/calc stages per timestep",
      "--checksum_freq - number of stages between checksums",
      "--stencil - 7 or 27 point",
      "--error_tol - (e^{-error_tol} ; >= 0)",
      "--report_diffusion - report check sums for each variable",
      "--report_perf - bitmask: 1 yaml, 2 text, 4 stdout, 8 block counts",
      "--plot_freq - frequency (timesteps) of plotting (0 for none)",
      "--code - communication variant (0 = minimal ghosts, 1/2 = whole ghost planes)",
      "--permute - alternate directions in communication",
      "--blocking_send, --send_faces - message options (no effect with one rank)",
      "--refine_ghost - use full extent of block (including ghosts) to determine if block is refined",
      "--change_dir, --group_blocks, --break_ties, --limit_move - RCB options (no effect with one rank)",
      "--rcb, --morton - load balance method (identical with one rank)",
      "--num_objects - (>= 0) number of objects to cause refinement",
      "--object - type, bounce, position(3), movement(3), size(3), size rate of change(3)",
      "",
      "Extensions (not in the reference):",
      "--seed - seed of the initial values' generator (default 1, the reference's)",
      "--precision - storage of the cell values: f64 (default), f32 or bf16",
      "--tape - record the device's data operations for the reverse pass (amr_adjoint.h)",
      "",
      "Environment: MINIAMR_PRECISION=f64|f32|bf16 storage precision (default f64),",
      "             MINIAMR_SCHEDULE=stage|var stage-major or variable-major order,",
      "             MINIAMR_DEVICE=<n> HIP device ordinal.",
      NULL };
   for (int i = 0; lines[i]; i++)
      printf("%s\n", lines[i]);
}

/* Prints the message and the help text, then fails as amr_fatal does. */
static void die_usage(const char *msg, const char *arg)
{
   char line[256] = "";
   if (msg)
      snprintf(line, sizeof(line), msg, arg);
   printf("%s", line);
   amr_print_help();
   amr_fatal_quiet(line[0] ? line : "help requested\n");
}

/* Reads the integer operand of argv[*i] (advancing *i) or aborts. */
static int int_arg(int argc, char **argv, int *i)
{
   if (*i + 1 >= argc)
      die_usage("** Error ** missing value for %s\n", argv[*i]);
   return atoi(argv[++*i]);
}

static double dbl_arg(int argc, char **argv, int *i)
{
   if (*i + 1 >= argc)
      die_usage("** Error ** missing value for %s\n", argv[*i]);
   return atof(argv[++*i]);
}

int amr_parse_args(amr_params *P, int argc, char **argv)
{
   /* integer-valued flags and where they go */
   struct { const char *flag; int *dst; int use_tsteps; } iflags[] = {
      {"--max_blocks", &P->max_num_blocks, 0}, {"--num_refine", &P->num_refine, 0},
      {"--block_change", &P->block_change, 0}, {"--uniform_refine", &P->uniform_refine, 0},
      {"--nx", &P->nx, 0}, {"--ny", &P->ny, 0}, {"--nz", &P->nz, 0},
      {"--num_vars", &P->num_vars, 0}, {"--comm_vars", &P->comm_vars, 0},
      {"--init_x", &P->init_block_x, 0}, {"--init_y", &P->init_block_y, 0},
      {"--init_z", &P->init_block_z, 0}, {"--reorder", &P->reorder, 0},
      {"--npx", &P->npx, 0}, {"--npy", &P->npy, 0}, {"--npz", &P->npz, 0},
      {"--inbalance", &P->inbalance, 0}, {"--lb_opt", &P->lb_opt, 0},
      {"--refine_freq", &P->refine_freq, 0}, {"--error_tol", &P->error_tol, 0},
      {"--num_tsteps", &P->num_tsteps, 1}, {"--stages_per_ts", &P->stages_per_ts, 0},
      {"--checksum_freq", &P->checksum_freq, 0}, {"--stencil", &P->stencil, 0},
      {"--report_perf", &P->report_perf, 0}, {"--plot_freq", &P->plot_freq, 0},
      {"--code", &P->code, 0}, {"--limit_move", &P->limit_move, 0},
      {"--seed", &P->seed, 0},
      {NULL, NULL, 0} };
   /* flags without operand: flag, destination, value stored */
   struct { const char *flag; int *dst; int value; } sflags[] = {
      {"--report_diffusion", &P->report_diffusion, 1}, {"--permute", &P->permute, 1},
      {"--blocking_send", &P->nonblocking, 0}, {"--refine_ghost", &P->refine_ghost, 1},
      {"--change_dir", &P->change_dir, 1}, {"--group_blocks", &P->group_blocks, 1},
      {"--break_ties", &P->group_blocks, 2}, {"--send_faces", &P->send_faces, 1},
      {"--rcb", &P->lb_method, 0}, {"--morton", &P->lb_method, 1}, {"--tape", &P->tape, 1},
      {"--hilbert", &P->lb_method, 2}, {"--trunc_hilbert", &P->lb_method, 3},
      {NULL, NULL, 0} };
   int object_num = 0;
   const char *prec_flag = NULL;

   set_defaults(P);
   for (int i = 1; i < argc; i++) {
      const char *a = argv[i];
      int done = 0;
      for (int f = 0; iflags[f].flag && !done; f++)
         if (!strcmp(a, iflags[f].flag)) {
            *iflags[f].dst = int_arg(argc, argv, &i);
            if (iflags[f].use_tsteps)
               P->use_tsteps = 1;
            done = 1;
         }
      for (int f = 0; sflags[f].flag && !done; f++)
         if (!strcmp(a, sflags[f].flag)) {
            *sflags[f].dst = sflags[f].value;
            done = 1;
         }
      if (done)
         continue;
      if (!strcmp(a, "--time")) {
         P->end_time = dbl_arg(argc, argv, &i);
         P->use_time = 1;
      } else if (!strcmp(a, "--num_objects")) {
         P->num_objects = int_arg(argc, argv, &i);
         free(P->objects);
         P->objects = (amr_object *) calloc(P->num_objects > 0 ? P->num_objects : 1,
                                            sizeof(amr_object));
         object_num = 0;
      } else if (!strcmp(a, "--object")) {
         if (object_num >= P->num_objects)
            amr_fatal("object number greater than num_objects\n");
         amr_object *o = &P->objects[object_num++];
         o->type = int_arg(argc, argv, &i);
         o->bounce = int_arg(argc, argv, &i);
         for (int d = 0; d < 3; d++) o->cen[d] = dbl_arg(argc, argv, &i);
         for (int d = 0; d < 3; d++) o->move[d] = dbl_arg(argc, argv, &i);
         for (int d = 0; d < 3; d++) o->size[d] = dbl_arg(argc, argv, &i);
         for (int d = 0; d < 3; d++) o->inc[d] = dbl_arg(argc, argv, &i);
      } else if (!strcmp(a, "--precision")) {
         if (i + 1 >= argc)
            die_usage("** Error ** missing value for %s\n", a);
         prec_flag = argv[++i];
      } else if (!strcmp(a, "--help")) {
         die_usage(NULL, NULL);
      } else {
         die_usage("** Error ** Unknown input parameter %s\n", a);
      }
   }
   if (object_num != P->num_objects)
      amr_fatal("Error - number of objects less than specified");
   if (P->reorder == -1)
      P->reorder = P->lb_method ? 0 : 1;
   for (int o = 0; o < P->num_objects; o++)
      for (int d = 0; d < 3; d++) {
         P->objects[o].orig_cen[d] = P->objects[o].cen[d];
         P->objects[o].orig_move[d] = P->objects[o].move[d];
         P->objects[o].orig_size[d] = P->objects[o].size[d];
      }

   /* build extensions, from the environment so the command line stays the
    * reference's */
   const char *prec = prec_flag ? prec_flag : getenv("MINIAMR_PRECISION");
   P->precision = AMR_PREC_F64;
   if (prec && !strcmp(prec, "f32"))
      P->precision = AMR_PREC_F32;
   else if (prec && !strcmp(prec, "bf16"))
      P->precision = AMR_PREC_BF16;
   else if (prec && strcmp(prec, "f64"))
      amr_fatal("precision must be f64, f32 or bf16\n");
   const char *sch = getenv("MINIAMR_SCHEDULE");
   P->schedule = (sch && !strcmp(sch, "var")) ? 1 : 0;
   if (P->stencil == 0)   /* coupled variables: keep the stage-major order */
      P->schedule = 0;
   const char *mk = getenv("MINIAMR_DEVICE_MARKING");
   P->device_marking = mk ? (atoi(mk) != 0) : -1;
   const char *dev = getenv("MINIAMR_DEVICE");
   P->device = dev ? atoi(dev) : 0;
   return 0;
}

/* Validation of ref/main.c:657-723 for one rank, plus what this driver leaves out. */
int amr_check_input(amr_params *P)
{
   int error = 0;

   if (P->init_block_x < 1 || P->init_block_y < 1 || P->init_block_z < 1) {
      error = amr_note_error("initial blocks on processor must be positive\n");
   }
   if (P->max_num_blocks < P->init_block_x * P->init_block_y * P->init_block_z) {
      error = amr_note_error("max_num_blocks not large enough\n");
   }
   if (P->nx < 1 || P->ny < 1 || P->nz < 1) {
      error = amr_note_error("block size must be positive\n");
   }
   if ((P->nx / 2) * 2 != P->nx) {
      error = amr_note_error("block size in x direction must be even\n");
   }
   if ((P->ny / 2) * 2 != P->ny) {
      error = amr_note_error("block size in y direction must be even\n");
   }
   if ((P->nz / 2) * 2 != P->nz) {
      error = amr_note_error("block size in z direction must be even\n");
   }
   if (P->num_refine < 0) {
      error = amr_note_error("number of refinement levels must be non-negative\n");
   }
   if (P->block_change < 0) {
      error = amr_note_error("number of refinement levels must be non-negative\n");
   }
   if (P->num_vars < 1) {
      error = amr_note_error("number of variables must be positive\n");
   }
   if (P->npx * P->npy * P->npz != 1) {
      error = amr_note_error("number of processors used does not match number allocated\n");
   }
   if (P->stencil != 0 && P->stencil != 7 && P->stencil != 27) {
      error = amr_note_error("illegal value for stencil\n");
   }
   if (P->stencil == 0 && P->num_vars < 8) {
      error = amr_note_error("if stencil is 0, num_vars must be more than 8\n");
   }
   if (P->stencil == 27 && P->num_refine && !P->uniform_refine)
      printf("WARNING: 27 point stencil with non-uniform refinement: answers may diverge\n");
   if (P->comm_vars == 0 || P->comm_vars > P->num_vars)
      P->comm_vars = P->num_vars;
   if (P->code < 0 || P->code > 2) {
      error = amr_note_error("code must be 0, 1, or 2\n");
   }
   if (P->lb_opt < 0 || P->lb_opt > 2) {
      error = amr_note_error("lb_opt must be 0, 1, or 2\n");
   }
   /* what this driver leaves out */
   if (!error && P->lb_method >= 2) {
      error = amr_note_error("miniAMR.x implements the RCB/Morton initial block order only (not --hilbert)\n");
   }
   if (!error && P->plot_freq) {
      error = amr_note_error("miniAMR.x does not write plot files (--plot_freq must be 0)\n");
   }
   return error;
}
