/*
 * amr.h -- host side of the miniAMR drop-in (one rank, one GPU).
 *
 * The reference keeps about a hundred globals (ref/block.h:54-183); here the
 * whole run is one `amr_run` value.  The block/parent/object model keeps the
 * reference's fields that matter on a single rank; fields that only exist
 * for other ranks (off-rank neighbour ids, nei_refine, parent_node,
 * child_node, comm lists) are dropped because every one of them is
 * constant or unused when there is one rank.
 *
 * Cell values never live here: they are owned by the device library
 * (include/amr_dev.h), which this driver calls for every data operation.
 */
#ifndef AMR_HOST_H
#define AMR_HOST_H

#include <setjmp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include "amr_dev.h"

#define AMR_BOUNDARY (-2)   /* nei_level of a face on the domain boundary */

/* One refinement object (ref/block.h:148-158). */
typedef struct {
   int type, bounce;
   double cen[3], move[3], size[3], inc[3];
   double orig_cen[3], orig_move[3], orig_size[3];
} amr_object;

/* Command-line parameters; names and defaults follow ref/param.h:27-66. */
typedef struct {
   int max_num_blocks, num_refine, uniform_refine;
   int nx, ny, nz, num_vars, comm_vars;
   int init_block_x, init_block_y, init_block_z;
   int reorder, npx, npy, npz, inbalance, refine_freq;
   int report_diffusion, error_tol, use_tsteps, num_tsteps, use_time;
   double end_time;
   int stages_per_ts, checksum_freq, stencil, report_perf, plot_freq;
   int num_objects, lb_opt, block_change, code, permute, nonblocking;
   int refine_ghost, change_dir, group_blocks, limit_move, send_faces;
   int lb_method;
   amr_object *objects;
   /* Extensions beyond the reference (environment, not command line). */
   int precision;        /* AMR_PREC_* (MINIAMR_PRECISION=f64|f32|bf16) */
   int schedule;         /* 0: stage-major like the reference,
                            1: variable-major within a refinement interval */
   int device;
   int device_marking;   /* MINIAMR_DEVICE_MARKING: -1 (unset) refinement
                            marking on the host when the refinement overlaps
                            the timestep's last stages (no device round
                            trips beside the queued stages), else on the
                            device when it accepts the mesh, else on the
                            host; 0 host; 1 device or fail */
   int seed;             /* --seed: generator seed of the initial values
                            (1 = the reference's implicit srand(1)) */
   int tape;             /* --tape: record data operations from allocation on */
} amr_params;

/* A block slot (ref/block.h:36-53 minus off-rank fields). */
typedef struct {
   int64_t number;       /* global block number; -1 while the slot is free */
   int level;
   int refine;           /* 1 refine, -1 may coarsen, 0 stay */
   int parent;           /* parent record index, -1 for an initial block */
   int child_number;     /* octant within the parent */
   int nei_level[6];     /* AMR_BOUNDARY on the domain boundary */
   int nei[6][2][2];     /* neighbour slots; [f][0][0] unless finer */
   int cen[3];           /* centre in units of the finest half-cell grid */
} amr_block;

/* A parent record (ref/block.h:56-69 minus off-rank fields). */
typedef struct {
   int64_t number;       /* -1 while the record is free */
   int level;
   int refine;
   int parent;
   int child_number;
   int child[8];         /* slot, or -1 - p when the child is parent p */
   int cen[3];
} amr_parent;

/* Growable job list for the device. */
typedef struct { void *v; int n, cap, elem; } amr_vec;

/* State of the timestep loop between amr_driver_step calls. */
typedef struct {
   int ts, comm_stage, state, n_rows;
   int stage_pos;              /* stages of timestep ts done by amr_driver_stages */
   int checked;                /* timestep ts's checksums already checked: 1 passed, 2 failed */
   int overlapped;             /* the refinement in progress overlaps queued stages */
   double delta, sim_time, t_start;
   double dev_ms[4];
   int64_t stage_calls;
   double *sums;
} amr_loop;

typedef struct amr_run {
   amr_params P;
   amr_loop loop;

   /* mesh */
   amr_block *blocks;          /* max_num_blocks slots */
   amr_parent *parents;        /* max_num_parents records */
   int max_num_parents, num_parents, max_active_parent;
   int num_active, max_active_block, local_max_b;
   int64_t *sorted_number;     /* sorted list (ref/block.h:72-77) */
   int *sorted_slot;
   int *sorted_index;          /* num_refine + 2 entries */
   amr_vec sorted_new;         /* slots given a block since the last rebuild */
   unsigned char *sorted_mark; /* max_num_blocks flags, zero between rebuilds */
   int64_t *num_blocks;        /* blocks per level */
   int64_t *block_start;       /* first block number of each level */
   int64_t *p8;
   int *p2;
   int init_x, init_y, init_z;
   int mesh_size[3];
   int cur_max_level;
   int64_t global_active;
   double tol;
   double *grid_sum;

   /* counters for the report */
   int64_t total_blocks, nb_min, nb_max;
   int64_t num_refined, num_reformed, total_red;
   double total_fp_divs, total_fp_adds, total_fp_muls;   /* --stencil 0 */
   /* --stencil 0 coefficients (ref/init.c:418-423): mat = num_vars / 4 */
   int mat;
   double a1, *a0;

   /* timers (seconds) in the reference's report slots */
   double t_all, t_init, t_refine_all, t_comm, t_calc, t_cs, t_main;
   double t_comm_dir[3];       /* exchange device time per direction */
   double t_cs_dev, t_cs_host; /* checksum: reduction kernels, host check */
   double t_refine_init;       /* the initial refinement */
   double t_ref_compare, t_ref_mark, t_ref_split, t_ref_coarsen;
   int nrs;                    /* refinement steps (ref/refine.c:45) */
   /* faces exchanged per direction, counted per comm call as the
    * reference does (ref/comm.c:169-196), and the current plan's counts */
   int64_t face_same[3], face_diff[3], face_bc[3];
   int32_t plan_faces[3][3];   /* [dir][same, coarse/fine, boundary] */
   /* device memory (amr_dev_memory): after initialisation, at the end */
   size_t mem_bytes, mem_bytes_init;
   int mem_buffers, mem_buffers_init;

   /* The reference draws its initial values from libc rand() with the
    * implicit seed 1 (ref/init.c:494).  The GPU runtime calls srand()/rand()
    * itself while it initialises, so the driver keeps a private state of the
    * same glibc generator, seeded as rand() is by default:
    * initstate_r(1, state, 128) yields rand()'s sequence exactly. */
   struct random_data rng;
   char rng_state[128];

   /* device */
   amr_dev *dev;
   struct amr_batch *batch;    /* when set, dev is created as a member of it */
   FILE *sum_log;              /* checksum log (AMR_DUMP_SUMS), or NULL */
   amr_vec split_jobs, merge_jobs;
   int mesh_dirty;             /* exchange plan must be rebuilt */
   amr_vec face_jobs;
   int *scratch_i;             /* max_num_blocks ints */
   double *cellbuf;            /* one slot's cells */
   int failed;                 /* a library call on this run hit a fatal error */
} amr_run;

/* params.c */
void amr_print_help(void);
int amr_parse_args(amr_params *P, int argc, char **argv);
int amr_check_input(amr_params *P);

/* mesh.c */
int amr_allocate(amr_run *R);
void amr_free(amr_run *R);
int amr_init(amr_run *R);
void amr_sorted_add(amr_run *R, int slot, int64_t number, int level);
void amr_sorted_del(amr_run *R, int64_t number, int level);
void amr_sorted_note(amr_run *R, int slot);
void amr_sorted_rebuild(amr_run *R);
int amr_n_active(const amr_run *R);
double amr_rand01(amr_run *R);
int amr_build_exchange_plan(amr_run *R);
int amr_sync_mesh(amr_run *R);
void amr_fatal(const char *fmt, ...);
void amr_fatal_quiet(const char *msg);
int amr_note_error(const char *msg);
void amr_catch(jmp_buf *jb);
const char *amr_error_message(void);
void *amr_xmalloc(size_t bytes);
void amr_vec_push(amr_vec *v, const void *elem);
double amr_now(void);
int amr_dev_check(amr_run *R, int rc, const char *what);

/* objects.c */
void amr_move_objects(amr_run *R, double delta);
void amr_check_objects(amr_run *R);
int amr_check_block(const amr_run *R, double cor[3][2]);

/* refine.c */
void amr_refine(amr_run *R, int ts);
int amr_refine_level(amr_run *R);
void amr_split_blocks(amr_run *R);
void amr_consolidate_blocks(amr_run *R);

/* driver.c */
int amr_driver(amr_run *R);
int amr_driver_begin(amr_run *R);
int amr_driver_step(amr_run *R);
int amr_driver_stages(amr_run *R, int n);
int amr_driver_step_begin(amr_run *R);
void amr_driver_step_enqueue(amr_run *R);
void amr_driver_step_advance(amr_run *R);
int amr_driver_step_end(amr_run *R);
void amr_driver_end(amr_run *R);
void amr_driver_set_sum_log(amr_run *R, const char *path);
int amr_dump_final_to(amr_run *R, const char *cells_path, const char *bsums_path,
                      const char *sample_path);
double amr_calc_time_step(amr_run *R);
void amr_report(amr_run *R);                 /* report.c */
void amr_count_comm(amr_run *R, int calls);  /* driver.c */
int amr_dump_final(amr_run *R);
double amr_checksum(amr_run *R, int var);

#endif
