/*
 * mesh.c -- block storage, initial mesh, sorted list and exchange plan.
 *
 * Initial mesh: ref/init.c:38-683 for one rank with the RCB/Morton layout
 * (lb_method < 2): blocks numbered x-fastest, slot == number, level 0,
 * reflective faces on the domain boundary, interior values drawn from the
 * glibc generator behind rand() with its default seed (or --seed), in the reference's
 * order (block, var, i, j, k), so the field equals the reference's.
 *
 * Sorted list: ref/block.c:610-657 -- active blocks grouped by level, each
 * level in descending block number.  Its order fixes the checksum
 * accumulation order and the iteration order of refinement passes.
 *
 * Exchange plan: the on-rank part of comm() (ref/comm.c:159-203) visits,
 * per direction, every active block's two faces and performs each
 * same-level, coarse/fine or boundary transfer once.  Within one direction
 * every transfer is a gather from cells that no transfer of that direction
 * writes, so the plan is simply the set of ghost planes each block
 * receives; it is rebuilt only when the mesh changes.
 */
#include <setjmp.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <time.h>

#include "amr.h"
#include "amr_adjoint.h"
#include "amr_batch.h"

/* Fatal errors.  The drop-in executable ends the process the way the
 * reference does (printf, then exit(-1): ref/block.c:72-75,120-124,
 * ref/comm.c:199-200).  The library entry points of amr_run.h and
 * amr_batch.h first arm a recovery point (amr_catch): a fatal error then
 * records its message, prints it as the reference would, and returns to
 * that point, so the entry point fails with an error code instead of
 * ending its caller's process.  State is per thread. */
static __thread jmp_buf *err_jump;
static __thread char err_msg[512];

void amr_catch(jmp_buf *jb)
{
   err_jump = jb;
   if (jb)
      err_msg[0] = 0;
}

const char *amr_error_message(void)
{
   return err_msg;
}

int amr_note_error(const char *msg)
{
   printf("%s", msg);
   if (!err_msg[0])
      snprintf(err_msg, sizeof(err_msg), "%s", msg);
   return 1;
}

void amr_fatal_quiet(const char *msg)
{
   fflush(stdout);
   if (err_jump) {
      if (!err_msg[0])
         snprintf(err_msg, sizeof(err_msg), "%s", msg);
      longjmp(*err_jump, 1);
   }
   exit(-1);
}

void amr_fatal(const char *fmt, ...)
{
   char line[512];
   va_list ap;
   va_start(ap, fmt);
   vsnprintf(line, sizeof(line), fmt, ap);
   va_end(ap);
   printf("%s", line);
   amr_fatal_quiet(line);
}

void *amr_xmalloc(size_t bytes)
{
   void *p = calloc(1, bytes ? bytes : 1);
   if (!p)
      amr_fatal("NULL pointer from malloc call (%zu bytes)\n", bytes);
   return p;
}

void amr_vec_push(amr_vec *v, const void *elem)
{
   if (v->n == v->cap) {
      v->cap = v->cap ? 2 * v->cap : 256;
      v->v = realloc(v->v, (size_t) v->cap * v->elem);
      if (!v->v)
         amr_fatal("out of host memory for job list\n");
   }
   memcpy((char *) v->v + (size_t) v->n * v->elem, elem, v->elem);
   v->n++;
}

double amr_now(void)
{
   struct timespec ts;
   clock_gettime(CLOCK_MONOTONIC, &ts);
   return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int amr_dev_check(amr_run *R, int rc, const char *what)
{
   if (rc != AMR_OK)
      amr_fatal("ERROR: device %s failed (%d): %s\n", what, rc,
                amr_dev_last_error(R->dev));
   return rc;
}

int amr_allocate(amr_run *R)
{
   const amr_params *P = &R->P;
   int nl = P->num_refine + 1;

   R->blocks = (amr_block *) amr_xmalloc(sizeof(amr_block) * P->max_num_blocks);
   for (int n = 0; n < P->max_num_blocks; n++)
      R->blocks[n].number = -1;
   R->max_num_parents = P->max_num_blocks;   /* same guess as ref/main.c:457 */
   R->parents = (amr_parent *) amr_xmalloc(sizeof(amr_parent) * R->max_num_parents);
   for (int n = 0; n < R->max_num_parents; n++)
      R->parents[n].number = -1;
   R->sorted_number = (int64_t *) amr_xmalloc(sizeof(int64_t) * (P->max_num_blocks + 1));
   R->sorted_slot = (int *) amr_xmalloc(sizeof(int) * (P->max_num_blocks + 1));
   R->sorted_index = (int *) amr_xmalloc(sizeof(int) * (nl + 1));
   R->num_blocks = (int64_t *) amr_xmalloc(sizeof(int64_t) * nl);
   R->block_start = (int64_t *) amr_xmalloc(sizeof(int64_t) * nl);
   R->p8 = (int64_t *) amr_xmalloc(sizeof(int64_t) * (nl + 1));
   R->p2 = (int *) amr_xmalloc(sizeof(int) * (nl + 1));
   R->grid_sum = (double *) amr_xmalloc(sizeof(double) * P->num_vars);
   R->scratch_i = (int *) amr_xmalloc(sizeof(int) * P->max_num_blocks);
   R->cellbuf = (double *) amr_xmalloc(sizeof(double) * (size_t) P->num_vars *
                                       (P->nx + 2) * (P->ny + 2) * (P->nz + 2));
   R->sorted_new.elem = sizeof(int);
   R->sorted_mark = (unsigned char *) amr_xmalloc(P->max_num_blocks);
   memset(R->sorted_mark, 0, P->max_num_blocks);
   R->split_jobs.elem = sizeof(amr_split_job);
   R->merge_jobs.elem = sizeof(amr_merge_job);
   R->face_jobs.elem = sizeof(amr_face_job);

   amr_dev_config cfg;
   cfg.nx = P->nx; cfg.ny = P->ny; cfg.nz = P->nz;
   cfg.num_vars = P->num_vars;
   cfg.max_blocks = P->max_num_blocks;
   cfg.stencil = P->stencil;
   cfg.precision = P->precision;
   cfg.device = P->device;
   if (R->batch) {
      int rc = amr_batch_add(R->batch, &cfg, &R->dev);
      if (rc != AMR_OK)
         amr_fatal("ERROR: could not add the device context to the batch (%d): %s\n", rc,
                   amr_batch_last_error(R->batch));
      return 0;
   }
   int rc = amr_dev_create(&cfg, &R->dev);
   if (rc != AMR_OK)
      amr_fatal("ERROR: could not create the device context (%d): %s\n", rc,
                R->dev ? amr_dev_last_error(R->dev) : "no context");
   if (P->tape)
      amr_dev_tape(R->dev, 1);
   return 0;
}

void amr_free(amr_run *R)
{
   if (R->dev && !R->batch)   /* batch members are destroyed with the batch */
      amr_dev_destroy(R->dev);
   free(R->a0);
   free(R->blocks); free(R->parents); free(R->sorted_number); free(R->sorted_slot);
   free(R->sorted_index); free(R->num_blocks); free(R->block_start);
   free(R->p8); free(R->p2); free(R->grid_sum); free(R->scratch_i); free(R->cellbuf);
   free(R->split_jobs.v); free(R->merge_jobs.v); free(R->face_jobs.v);
   free(R->sorted_new.v); free(R->sorted_mark);
   free(R->P.objects);
   memset(R, 0, sizeof(*R));
}

int amr_n_active(const amr_run *R)
{
   return R->sorted_index[R->P.num_refine + 1];
}

/* Insert before the first entry of the level with a smaller number
 * (ref/block.c:610-626). */
void amr_sorted_add(amr_run *R, int slot, int64_t number, int level)
{
   int last = R->P.num_refine + 1;
   int end = R->sorted_index[last];
   int i = R->sorted_index[level];
   while (i < R->sorted_index[level + 1] && !(number > R->sorted_number[i]))
      i++;
   memmove(&R->sorted_number[i + 1], &R->sorted_number[i], sizeof(int64_t) * (end - i));
   memmove(&R->sorted_slot[i + 1], &R->sorted_slot[i], sizeof(int) * (end - i));
   R->sorted_number[i] = number;
   R->sorted_slot[i] = slot;
   for (int l = level + 1; l <= last; l++)
      R->sorted_index[l]++;
}

/* Split and consolidate add and remove hundreds of blocks per pass.  They
 * note each slot given a block (amr_sorted_note) and update the sorted list
 * once at the end (amr_sorted_rebuild) instead of shifting its tail per
 * change.  The list holds the active blocks level by level and, within a
 * level, by number from largest to smallest; numbers are unique, so this
 * is the list that amr_sorted_add / amr_sorted_del calls in any order over
 * the same changes would leave (ref/block.c:610-646).  The rebuild keeps
 * the old entries whose slot still holds the same block, sorts the noted
 * ones (a few hundred) and merges the two, level by level. */
typedef struct {
   int64_t number;
   int slot, level;
} sorted_entry;

static int cmp_sorted(const void *a, const void *b)
{
   const sorted_entry *x = (const sorted_entry *) a, *y = (const sorted_entry *) b;
   if (x->level != y->level)
      return x->level < y->level ? -1 : 1;
   if (x->number != y->number)
      return x->number > y->number ? -1 : 1;
   return 0;
}

void amr_sorted_note(amr_run *R, int slot)
{
   amr_vec_push(&R->sorted_new, &slot);
}

void amr_sorted_rebuild(amr_run *R)
{
   const int last = R->P.num_refine + 1;
   const int *noted = (const int *) R->sorted_new.v;
   const int k = R->sorted_new.n;
   /* the noted slots that still hold a block, each once, in list order */
   sorted_entry *add = (sorted_entry *) amr_xmalloc(sizeof(sorted_entry) * (size_t) (k + 1));
   int na = 0;
   for (int t = 0; t < k; t++) {
      const int s = noted[t];
      if (R->sorted_mark[s])
         continue;
      R->sorted_mark[s] = 1;
      if (R->blocks[s].number >= 0) {
         add[na].number = R->blocks[s].number;
         add[na].slot = s;
         add[na].level = R->blocks[s].level;
         na++;
      }
   }
   qsort(add, (size_t) na, sizeof(sorted_entry), cmp_sorted);
   const int old_n = R->sorted_index[last];
   int64_t *num = (int64_t *) amr_xmalloc(sizeof(int64_t) * (size_t) (old_n + na + 1));
   int *slot = (int *) amr_xmalloc(sizeof(int) * (size_t) (old_n + na + 1));
   int n = 0, a = 0;
   for (int l = 0; l < last; l++) {
      const int i1 = R->sorted_index[l + 1];
      int i = R->sorted_index[l];
      R->sorted_index[l] = n;
      for (;;) {
         /* next surviving old entry of the level: its slot holds the same
          * block and was not given another one in this pass */
         while (i < i1 && (R->sorted_mark[R->sorted_slot[i]] ||
                           R->blocks[R->sorted_slot[i]].number != R->sorted_number[i]))
            i++;
         const int has_new = a < na && add[a].level == l;
         if (i < i1 && (!has_new || R->sorted_number[i] > add[a].number)) {
            num[n] = R->sorted_number[i];
            slot[n++] = R->sorted_slot[i++];
         } else if (has_new) {
            num[n] = add[a].number;
            slot[n++] = add[a++].slot;
         } else {
            break;
         }
      }
   }
   R->sorted_index[last] = n;
   memcpy(R->sorted_number, num, sizeof(int64_t) * (size_t) n);
   memcpy(R->sorted_slot, slot, sizeof(int) * (size_t) n);
   for (int t = 0; t < k; t++)
      R->sorted_mark[noted[t]] = 0;
   R->sorted_new.n = 0;
   free(add);
   free(num);
   free(slot);
}

/* ref/block.c:628-646 */
void amr_sorted_del(amr_run *R, int64_t number, int level)
{
   int last = R->P.num_refine + 1;
   int i = R->sorted_index[level];
   while (i < R->sorted_index[level + 1] && R->sorted_number[i] != number)
      i++;
   if (i >= R->sorted_index[level + 1])
      amr_fatal("ERROR: del_sorted_list on %d - number %lld not found l %d\n",
                0, (long long) number, level);
   int end = R->sorted_index[last];
   memmove(&R->sorted_number[i], &R->sorted_number[i + 1], sizeof(int64_t) * (end - i - 1));
   memmove(&R->sorted_slot[i], &R->sorted_slot[i + 1], sizeof(int) * (end - i - 1));
   for (int l = level + 1; l <= last; l++)
      R->sorted_index[l]--;
}

/* Initial mesh and values (ref/init.c:38-683, one rank, lb_method < 2). */
int amr_init(amr_run *R)
{
   amr_params *P = &R->P;
   int nr = P->num_refine;

   R->tol = pow(10.0, (double) -P->error_tol);
   memset(&R->rng, 0, sizeof(R->rng));
   initstate_r((unsigned int) P->seed, R->rng_state, sizeof(R->rng_state), &R->rng);
   R->p2[0] = 1;
   R->p8[0] = 1;
   for (int i = 0; i <= nr; i++) {
      R->p8[i + 1] = R->p8[i] * 8;
      R->p2[i + 1] = R->p2[i] * 2;
      R->sorted_index[i] = 0;
   }
   R->sorted_index[nr + 1] = 0;

   int n0 = P->init_block_x * P->init_block_y * P->init_block_z;
   R->local_max_b = n0;
   R->block_start[0] = 0;
   R->num_blocks[0] = n0;
   int64_t num = n0;
   for (int i = 1; i <= nr; i++) {
      R->block_start[i] = R->block_start[i - 1] + num;
      num *= 8;
      R->num_blocks[i] = 0;
   }
   R->init_x = P->init_block_x;
   R->init_y = P->init_block_y;
   R->init_z = P->init_block_z;

   /* --stencil 0: the material coefficients, drawn from the generator
    * before the initial values (ref/init.c:418-423) */
   if (P->stencil == 0) {
      R->mat = P->num_vars / 4;
      R->a1 = amr_rand01(R);
      R->a0 = (double *) amr_xmalloc(sizeof(double) * R->mat);
      for (int i = 0; i < R->mat; i++)
         R->a0[i] = amr_rand01(R);
      amr_dev_check(R, amr_dev_set_materials(R->dev, R->a1, R->a0, R->mat), "set_materials");
   }

   R->num_active = R->max_active_block = n0;
   R->global_active = n0;
   R->num_parents = R->max_active_parent = 0;
   int size = R->p2[nr + 1];
   R->mesh_size[0] = R->init_x * size;
   R->mesh_size[1] = R->init_y * size;
   R->mesh_size[2] = R->init_z * size;

   const int gx = P->nx + 2, gy = P->ny + 2, gz = P->nz + 2;
   const size_t slab = (size_t) gx * gy * gz;
   int n = 0;
   for (int k = 0; k < P->init_block_z; k++)
      for (int j = 0; j < P->init_block_y; j++)
         for (int i = 0; i < P->init_block_x; i++, n++) {
            amr_block *b = &R->blocks[n];
            int idx[3] = { i, j, k };
            int cnt[3] = { P->init_block_x, P->init_block_y, P->init_block_z };
            int stride[3] = { 1, P->init_block_x, P->init_block_x * P->init_block_y };
            b->level = 0;
            b->number = n;
            b->refine = 0;
            b->parent = -1;
            b->child_number = 0;
            b->cen[0] = i * size + size / 2;
            b->cen[1] = j * size + size / 2;
            b->cen[2] = k * size + size / 2;
            amr_sorted_add(R, n, n, 0);
            /* zero the whole slot, then the interior from the generator */
            memset(R->cellbuf, 0, sizeof(double) * slab * P->num_vars);
            for (int v = 0; v < P->num_vars; v++) {
               double *a = R->cellbuf + slab * v;
               for (int x = 1; x <= P->nx; x++)
                  for (int y = 1; y <= P->ny; y++)
                     for (int z = 1; z <= P->nz; z++)
                        a[((size_t) x * gy + y) * gz + z] = amr_rand01(R);
            }
            amr_dev_check(R, amr_dev_upload_block(R->dev, n, R->cellbuf), "upload");
            memset(b->nei, 0, sizeof(b->nei));
            for (int d = 0; d < 3; d++) {
               /* low face 2d, high face 2d+1 */
               if (idx[d] == 0) {
                  b->nei_level[2 * d] = AMR_BOUNDARY;
                  b->nei[2 * d][0][0] = 0;
               } else {
                  b->nei_level[2 * d] = 0;
                  b->nei[2 * d][0][0] = n - stride[d];
               }
               if (idx[d] == cnt[d] - 1) {
                  b->nei_level[2 * d + 1] = AMR_BOUNDARY;
                  b->nei[2 * d + 1][0][0] = 0;
               } else {
                  b->nei_level[2 * d + 1] = 0;
                  b->nei[2 * d + 1][0][0] = n + stride[d];
               }
            }
         }
   R->mesh_dirty = 1;
   amr_sync_mesh(R);
   for (int v = 0; v < P->num_vars; v++)
      R->grid_sum[v] = amr_checksum(R, v);
   return 0;
}

/* rand()/RAND_MAX from the private generator (see amr_run.rng). */
double amr_rand01(amr_run *R)
{
   int32_t r;
   random_r(&R->rng, &r);
   return ((double) r) / ((double) RAND_MAX);
}

static int opposite(int f) { return f ^ 1; }

/* Ghost planes each active block receives, per direction
 * (ref/comm.c:159-203 with on_proc_comm, on_proc_comm_diff, apply_bc).
 * --code 1/2 jobs move whole ghost planes (AMR_FACE_GHOSTS). */
int amr_build_exchange_plan(amr_run *R)
{
   amr_vec *v = &R->face_jobs;
   int32_t count[3] = { 0, 0, 0 };
   int na = amr_n_active(R);
   const int whole = R->P.code ? AMR_FACE_GHOSTS : 0;
   v->n = 0;
   for (int dir = 0; dir < 3; dir++) {
      int before = v->n;
      for (int in = 0; in < na; in++) {
         int s = R->sorted_slot[in];
         const amr_block *b = &R->blocks[s];
         for (int f = 2 * dir; f < 2 * dir + 2; f++) {
            amr_face_job j;
            memset(&j, 0, sizeof(j));
            j.face = f;
            j.dst = s;
            if (b->nei_level[f] == b->level) {
               j.kind = AMR_FACE_SAME | whole;
               j.src = b->nei[f][0][0];
               amr_vec_push(v, &j);
            } else if (b->nei_level[f] == b->level + 1) {
               /* coarse side: one restriction per finer neighbour */
               for (int iq = 0; iq < 2; iq++)
                  for (int jq = 0; jq < 2; jq++) {
                     j.kind = AMR_FACE_RESTRICT | whole;
                     j.src = b->nei[f][iq][jq];
                     j.iq = iq;
                     j.jq = jq;
                     amr_vec_push(v, &j);
                  }
            } else if (b->nei_level[f] == b->level - 1) {
               /* fine side: find which quarter of the coarse face we are */
               int m = b->nei[f][0][0], g = opposite(f), found = 0;
               for (int iq = 0; iq < 2 && !found; iq++)
                  for (int jq = 0; jq < 2 && !found; jq++)
                     if (R->blocks[m].nei[g][iq][jq] == s) {
                        j.kind = AMR_FACE_PROLONG | whole;
                        j.src = m;
                        j.iq = iq;
                        j.jq = jq;
                        amr_vec_push(v, &j);
                        found = 1;
                     }
            } else if (b->nei_level[f] == AMR_BOUNDARY) {
               j.kind = AMR_FACE_BC | whole;
               j.src = s;
               amr_vec_push(v, &j);
            } else {
               amr_fatal("ERROR: misconnected block\n");
            }
         }
      }
      count[dir] = v->n - before;
   }
   /* faces per comm call for the report: same level 2 per pair (one job per
    * side), coarse/fine 2 per quarter pair (restriction + prolongation job),
    * 1 per boundary face */
   memset(R->plan_faces, 0, sizeof(R->plan_faces));
   for (int t = 0, dir = 0; dir < 3; dir++)
      for (int e = 0; e < count[dir]; e++, t++) {
         const amr_face_job *jj = (const amr_face_job *) v->v + t;
         int kind = jj->kind & 3;
         int k = kind == AMR_FACE_SAME ? 0 : kind == AMR_FACE_BC ? 2 : 1;
         R->plan_faces[dir][k]++;
      }
   return amr_dev_check(R, amr_dev_set_mesh(R->dev, R->sorted_slot, na,
                                            (const amr_face_job *) v->v, count),
                        "set_mesh");
}

/* Push pending split/consolidate data work and a changed mesh to the device. */
int amr_sync_mesh(amr_run *R)
{
   if (R->mesh_dirty) {
      amr_build_exchange_plan(R);
      R->mesh_dirty = 0;
   }
   return 0;
}
