/*
 * amr_cpu.c -- CPU oracle for the miniAMR hot path (TEST INFRASTRUCTURE).
 *
 * A plain-C implementation of include/amr_dev.h that restates the
 * reference's arithmetic in the reference's exact operation order, FP64
 * only.  It is linked into oracle/_build/miniAMR_oracle.x (with the
 * product's host driver) and built as oracle/_build/liboracle_dev.so for
 * ctypes.  Only tests, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg use it, as the checker; the product never links it.
 *
 * Functions and the reference code they restate:
 *   stencil_7 / stencil_27   ref/stencil.c:82-101 / 102-144
 *   stencil 0 (stencil0_*)   ref/stencil.c:43-73, 147-981 (variable work)
 *   fill_*, --code 1/2       ref/comm.c:1535-1579, 1689-1907
 *   fill_same                ref/comm.c:1473-1534 (on_proc_comm, code 0)
 *   fill_prolong / restrict  ref/comm.c:1597-1688 (on_proc_comm_diff, code 0)
 *   fill_bc                  ref/comm.c:1911-1965 (apply_bc)
 *   check_sum                ref/check_sum.c:36-65
 *   split                    ref/block.c:161-173
 *   consolidate              ref/block.c:411-431
 * Parity: pinned by the golden fixtures of tests/golden (checksums at
 * %.17g and full cell dumps of the reference on the same inputs).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>

#include "amr_dev.h"

struct amr_dev {
   amr_dev_config c;
   int gx, gy, gz;
   size_t slab;                 /* doubles per (var, slot) */
   double *pool;                /* [var][slot][gx][gy][gz] */
   int32_t *sorted;
   int n_active;
   amr_face_job *jobs;
   int32_t count[3];
   double *sums;
   int sum_rows;
   double *work;                /* one slab, stencil scratch */
   int mat;                     /* --stencil 0: num_vars / 4 materials */
   double a1, *a0;
   int64_t checks[2];           /* stencil_check: divides, scales */
   char err[256];
};

#define A(d, v, s, i, j, k) \
   ((d)->pool[((size_t) (v) * (d)->c.max_blocks + (s)) * (d)->slab + \
              ((size_t) (i) * (d)->gy + (j)) * (d)->gz + (k)])

static int fail(amr_dev *d, int rc, const char *msg)
{
   if (d)
      snprintf(d->err, sizeof(d->err), "%s", msg);
   return rc;
}

int amr_dev_create(const amr_dev_config *cfg, amr_dev **out)
{
   *out = NULL;
   if (!cfg || cfg->nx < 2 || cfg->ny < 2 || cfg->nz < 2 || (cfg->nx | cfg->ny | cfg->nz) & 1 ||
       cfg->num_vars < 1 || cfg->max_blocks < 1 ||
       (cfg->stencil != 0 && cfg->stencil != 7 && cfg->stencil != 27) ||
       cfg->precision != AMR_PREC_F64)
      return AMR_ERR_ARG;
   amr_dev *d = (amr_dev *) calloc(1, sizeof(amr_dev));
   d->c = *cfg;
   d->gx = cfg->nx + 2; d->gy = cfg->ny + 2; d->gz = cfg->nz + 2;
   d->slab = (size_t) d->gx * d->gy * d->gz;
   d->pool = (double *) calloc(d->slab * cfg->num_vars * cfg->max_blocks, sizeof(double));
   d->work = (double *) calloc(d->slab, sizeof(double));
   d->sorted = (int32_t *) calloc(cfg->max_blocks, sizeof(int32_t));
   if (!d->pool || !d->work || !d->sorted) {
      free(d->pool); free(d->work); free(d->sorted); free(d);
      return AMR_ERR_NOMEM;
   }
   *out = d;
   return AMR_OK;
}

int amr_dev_destroy(amr_dev *d)
{
   if (!d) return AMR_OK;
   free(d->pool); free(d->work); free(d->sorted); free(d->jobs); free(d->sums);
   free(d->a0);
   free(d);
   return AMR_OK;
}

const char *amr_dev_last_error(const amr_dev *d) { return d ? d->err : "no context"; }

int amr_dev_upload_block(amr_dev *d, int slot, const double *cells)
{
   if (slot < 0 || slot >= d->c.max_blocks)
      return fail(d, AMR_ERR_ARG, "slot out of range");
   for (int v = 0; v < d->c.num_vars; v++)
      memcpy(&A(d, v, slot, 0, 0, 0), cells + d->slab * v, sizeof(double) * d->slab);
   return AMR_OK;
}

int amr_dev_download_block(amr_dev *d, int slot, double *cells)
{
   if (slot < 0 || slot >= d->c.max_blocks)
      return fail(d, AMR_ERR_ARG, "slot out of range");
   for (int v = 0; v < d->c.num_vars; v++)
      memcpy(cells + d->slab * v, &A(d, v, slot, 0, 0, 0), sizeof(double) * d->slab);
   return AMR_OK;
}

int amr_dev_set_mesh(amr_dev *d, const int32_t *sorted, int n_active,
                     const amr_face_job *jobs, const int32_t count[3])
{
   int nj = count[0] + count[1] + count[2];
   memcpy(d->sorted, sorted, sizeof(int32_t) * n_active);
   d->n_active = n_active;
   free(d->jobs);
   d->jobs = (amr_face_job *) malloc(sizeof(amr_face_job) * (nj ? nj : 1));
   memcpy(d->jobs, jobs, sizeof(amr_face_job) * nj);
   memcpy(d->count, count, sizeof(d->count));
   return AMR_OK;
}

/* n of axis a (0 = x, 1 = y, 2 = z) */
static int nsz(const amr_dev *d, int a) { return a == 0 ? d->c.nx : a == 1 ? d->c.ny : d->c.nz; }

/* Cell reference by axis-ordered coordinates: dir-axis index p and the
 * two in-plane indices (a slower, b faster) of a face of direction dir. */
static double *cell(amr_dev *d, int v, int s, int dir, int p, int a, int b)
{
   if (dir == 0) return &A(d, v, s, p, a, b);   /* x face: (j, k) */
   if (dir == 1) return &A(d, v, s, a, p, b);   /* y face: (i, k) */
   return &A(d, v, s, a, b, p);                 /* z face: (i, j) */
}

/* in-plane axes of a face of direction dir: slower, faster */
static void plane_axes(int dir, int *sa, int *sb)
{
   *sa = dir == 0 ? 1 : 0;
   *sb = dir == 2 ? 1 : 2;
}

/* Extended in-plane ranges for the 27-point stencil (ref/comm.c:1496-1533,
 * 1941-1964): y faces extend x, z faces extend x and y, BC extends both.
 * --code 1/2 jobs (AMR_FACE_GHOSTS) cover the whole (n+2)^2 plane at any
 * stencil (ref/comm.c:1535-1579, 1941-1964). */
static void ranges(amr_dev *d, const amr_face_job *j, int *a0, int *a1, int *b0, int *b1)
{
   int dir = j->face / 2, sa, sb;
   plane_axes(dir, &sa, &sb);
   int ext_a = 0, ext_b = 0;
   if (j->kind & AMR_FACE_GHOSTS) {
      ext_a = ext_b = 1;
   } else if (d->c.stencil != 7) {   /* 27, and 0 (ref/comm.c:1496, 1941) */
      if (j->kind == AMR_FACE_BC) ext_a = ext_b = 1;
      else if (dir == 1) ext_a = 1;
      else if (dir == 2) ext_a = ext_b = 1;
   }
   *a0 = ext_a ? 0 : 1; *a1 = nsz(d, sa) + (ext_a ? 1 : 0);
   *b0 = ext_b ? 0 : 1; *b1 = nsz(d, sb) + (ext_b ? 1 : 0);
}

static void run_job(amr_dev *d, const amr_face_job *j, int v)
{
   int dir = j->face / 2, hi = j->face % 2, sa, sb;
   int n = nsz(d, dir);
   int ghost = hi ? n + 1 : 0;     /* plane written */
   int src_pl = hi ? 1 : n;        /* neighbour's plane read (SAME/PROLONG/RESTRICT) */
   plane_axes(dir, &sa, &sb);
   int ha = nsz(d, sa) / 2, hb = nsz(d, sb) / 2;
   int a0, a1, b0, b1;
   int whole = j->kind & AMR_FACE_GHOSTS;
   switch (j->kind & 3) {
   case AMR_FACE_SAME:
      ranges(d, j, &a0, &a1, &b0, &b1);
      for (int a = a0; a <= a1; a++)
         for (int b = b0; b <= b1; b++)
            *cell(d, v, j->dst, dir, ghost, a, b) = *cell(d, v, j->src, dir, src_pl, a, b);
      break;
   case AMR_FACE_BC:
      ranges(d, j, &a0, &a1, &b0, &b1);
      for (int a = a0; a <= a1; a++)
         for (int b = b0; b <= b1; b++)
            *cell(d, v, j->dst, dir, ghost, a, b) = *cell(d, v, j->dst, dir, hi ? n : 1, a, b);
      break;
   case AMR_FACE_PROLONG: {
      /* fine ghost quad <- coarse cell / 4; a offset jq*ha, b offset iq*hb */
      int ao = j->jq * ha, bo = j->iq * hb;
      if (whole) {
         /* the whole fine plane (ref/comm.c:1698-1758 for x): fine ring row
          * 0 takes coarse row ao, fine ring row n+1 coarse row ao + ha + 1,
          * fine rows 2a-1 and 2a coarse row ao + a; columns alike */
         for (int a = 0; a <= 2 * ha + 1; a++) {
            int ca = a == 0 ? ao : a == 2 * ha + 1 ? ao + ha + 1 : ao + (a + 1) / 2;
            for (int b = 0; b <= 2 * hb + 1; b++) {
               int cb = b == 0 ? bo : b == 2 * hb + 1 ? bo + hb + 1 : bo + (b + 1) / 2;
               *cell(d, v, j->dst, dir, ghost, a, b) =
                  *cell(d, v, j->src, dir, src_pl, ca, cb) / 4.0;
            }
         }
         break;
      }
      for (int a = 1; a <= ha; a++)
         for (int b = 1; b <= hb; b++) {
            double q = *cell(d, v, j->src, dir, src_pl, a + ao, b + bo) / 4.0;
            *cell(d, v, j->dst, dir, ghost, 2 * a - 1, 2 * b - 1) = q;
            *cell(d, v, j->dst, dir, ghost, 2 * a - 1, 2 * b) = q;
            *cell(d, v, j->dst, dir, ghost, 2 * a, 2 * b - 1) = q;
            *cell(d, v, j->dst, dir, ghost, 2 * a, 2 * b) = q;
         }
      break;
   }
   case AMR_FACE_RESTRICT: {
      int ao = j->jq * ha, bo = j->iq * hb;
      if (whole) {
         /* the quarter plus the ring row on its side (coarse row 0 for
          * jq 0, n+1 for jq 1) and the ring column on its side
          * (ref/comm.c:1708-1760 for x): ring cells sum the fine ring
          * row's (or column's) two cells, the ring corner copies one */
         int alo = j->jq ? ha + 1 : 0, ahi = j->jq ? 2 * ha + 1 : ha;
         int blo = j->iq ? hb + 1 : 0, bhi = j->iq ? 2 * hb + 1 : hb;
         for (int a = alo; a <= ahi; a++)
            for (int b = blo; b <= bhi; b++) {
               int ra = a == 0 || a == 2 * ha + 1, rb = b == 0 || b == 2 * hb + 1;
               int fa = 2 * (a - ao), fb = 2 * (b - bo);
               double val;
               if (ra && rb)
                  val = *cell(d, v, j->src, dir, src_pl, a, b);
               else if (ra)
                  val = *cell(d, v, j->src, dir, src_pl, a, fb - 1) +
                        *cell(d, v, j->src, dir, src_pl, a, fb);
               else if (rb)
                  val = *cell(d, v, j->src, dir, src_pl, fa - 1, b) +
                        *cell(d, v, j->src, dir, src_pl, fa, b);
               else
                  val = *cell(d, v, j->src, dir, src_pl, fa - 1, fb - 1) +
                        *cell(d, v, j->src, dir, src_pl, fa - 1, fb) +
                        *cell(d, v, j->src, dir, src_pl, fa, fb - 1) +
                        *cell(d, v, j->src, dir, src_pl, fa, fb);
               *cell(d, v, j->dst, dir, ghost, a, b) = val;
            }
         break;
      }
      for (int a = 1; a <= ha; a++)
         for (int b = 1; b <= hb; b++)
            *cell(d, v, j->dst, dir, ghost, a + ao, b + bo) =
               *cell(d, v, j->src, dir, src_pl, 2 * a - 1, 2 * b - 1) +
               *cell(d, v, j->src, dir, src_pl, 2 * a - 1, 2 * b) +
               *cell(d, v, j->src, dir, src_pl, 2 * a, 2 * b - 1) +
               *cell(d, v, j->src, dir, src_pl, 2 * a, 2 * b);
      break;
   }
   }
}

int amr_dev_comm(amr_dev *d, int start, int num, int stage, int permute)
{
   static const int perm[6][3] = { {0, 1, 2}, {1, 2, 0}, {2, 0, 1},
                                   {0, 2, 1}, {1, 0, 2}, {2, 1, 0} };
   if (start < 0 || num < 0 || start + num > d->c.num_vars)
      return fail(d, AMR_ERR_ARG, "variable range");
   int off[3] = { 0, d->count[0], d->count[0] + d->count[1] };
   for (int o = 0; o < 3; o++) {
      int dir = permute ? perm[stage % 6][o] : o;
      for (int v = start; v < start + num; v++)
         for (int t = 0; t < d->count[dir]; t++)
            run_job(d, &d->jobs[off[dir] + t], v);
   }
   return AMR_OK;
}

static void stencil_avg(amr_dev *d, int var, int st)
{
   int nx = d->c.nx, ny = d->c.ny, nz = d->c.nz;
   double *w = d->work;
#define W(i, j, k) w[((size_t) (i) * d->gy + (j)) * d->gz + (k)]
   for (int in = 0; in < d->n_active; in++) {
      int s = d->sorted[in];
      if (st == 7) {
         for (int i = 1; i <= nx; i++)
            for (int j = 1; j <= ny; j++)
               for (int k = 1; k <= nz; k++)
                  W(i, j, k) = (A(d, var, s, i - 1, j, k) + A(d, var, s, i, j - 1, k) +
                                A(d, var, s, i, j, k - 1) + A(d, var, s, i, j, k) +
                                A(d, var, s, i, j, k + 1) + A(d, var, s, i, j + 1, k) +
                                A(d, var, s, i + 1, j, k)) / 7.0;
      } else {
         for (int i = 1; i <= nx; i++)
            for (int j = 1; j <= ny; j++)
               for (int k = 1; k <= nz; k++) {
                  double sp[3];
                  for (int p = 0; p < 3; p++) {
                     int ii = i - 1 + p;
                     double acc = A(d, var, s, ii, j - 1, k - 1);
                     acc = acc + A(d, var, s, ii, j - 1, k);
                     acc = acc + A(d, var, s, ii, j - 1, k + 1);
                     acc = acc + A(d, var, s, ii, j, k - 1);
                     acc = acc + A(d, var, s, ii, j, k);
                     acc = acc + A(d, var, s, ii, j, k + 1);
                     acc = acc + A(d, var, s, ii, j + 1, k - 1);
                     acc = acc + A(d, var, s, ii, j + 1, k);
                     acc = acc + A(d, var, s, ii, j + 1, k + 1);
                     sp[p] = acc;
                  }
                  W(i, j, k) = (sp[0] + sp[1] + sp[2]) / 27.0;
               }
      }
      for (int i = 1; i <= nx; i++)
         for (int j = 1; j <= ny; j++)
            for (int k = 1; k <= nz; k++)
               A(d, var, s, i, j, k) = W(i, j, k);
   }
#undef W
}

/* ---- --stencil 0: the variable-work stencils -------------------------
 * Blocks in sorted order, cells in the reference's i, j, k loop order,
 * updated in place as the reference does (the sweeps of stages 1-3 thus
 * read the new value of the cell before), Jacobi stages 4-5 through the
 * work slab; then stencil_check over every block. */

/* class partners in units of mat: 7-point x, y, z neighbours; 27-point
 * neighbours with 1, 2, 3 non-zero offsets (ref/stencil.c:661-958) */
static const int kMatPartner[4][3] = { {1, 2, 3}, {1, 2, -1}, {1, -2, -1}, {-3, -2, -1} };

/* stage 0, ref/stencil.c:147-225 */
static double stencil0_point(amr_dev *d, int var, int s, int i, int j, int k)
{
#define P(u) A(d, (u), s, i, j, k)
   const int mat = d->mat;
   const double a1 = d->a1, *a0 = d->a0;
   if (var == 1) {
      double r = P(1);
      for (int u = mat; u < 2 * mat; u++)
         r += P(u) * P(0);
      return r;
   }
   double v = P(var);
   if (var < mat)
      return v + v * (P(0) + P(1) - a1 * v);
   if (var < 2 * mat)
      return v * (P(0) + v + a1 * P(var + mat) + (1.0 - a1) * P(var + 2 * mat)) / P(1);
   if (var < 3 * mat)
      return v + P(var - mat) * (a1 * P(0) + a0[var - 2 * mat] * v + (1.0 - a1) * P(var + mat)) /
                 P(1);
   return v + P(var - 2 * mat) * (a1 * P(0) + a0[var - 3 * mat] * v +
                                  (1.0 - a0[var - 3 * mat]) * P(var - mat) +
                                  (1.0 - a1) * P(var - 2 * mat)) / (P(1) * P(1));
}

/* stages 1-3 (ref/stencil.c:233-660): axis (ex, ey, ez); cells before this
 * one along the axis already hold their new value */
static double stencil0_sweep(amr_dev *d, int var, int s, int i, int j, int k, int ex, int ey,
                             int ez)
{
   const int mat = d->mat;
   const double a1 = d->a1, *a0 = d->a0;
   if (var == 1) {
      double r = P(1);
      for (int u = 2; u < mat + 2; u++)
         r += P(u) * P(0);
      return r / (a1 + r);
   }
   double v = P(var);
   if (var < mat)
      return v + v * (P(0) + P(1) - a1 * v) / (a0[var] + P(1));
   const int cls = var / mat, g = cls == 1 ? var : var - (cls - 1) * mat;
   double vm = A(d, var, s, i - ex, j - ey, k - ez), vp = A(d, var, s, i + ex, j + ey, k + ez);
   double tmp1 = fabs(P(g) - A(d, g, s, i - ex, j - ey, k - ez));
   double tmp2 = fabs(P(g) - A(d, g, s, i + ex, j + ey, k + ez));
   double den;
   if (cls == 1)
      den = a1 + a0[var - mat] + vm + v + vp + P(0) + P(1);
   else if (cls == 2)
      den = a1 + a0[var - 2 * mat] + P(var - mat) + P(var + mat) + vm + v + vp;
   else
      den = a1 + a0[var - 3 * mat] + P(var - mat) + P(var - 2 * mat) + vm + v + vp;
   if (tmp1 > tmp2)
      return (tmp1 * vm + (tmp1 - tmp2) * (v + P(1)) + tmp2 * vp) / den;
   return (tmp1 * vm + (tmp2 - tmp1) * (v + P(1)) + tmp2 * vp) / den;
}

/* stage 4 (ref/stencil.c:661-785) */
static double stencil0_7(amr_dev *d, int var, int s, int i, int j, int k)
{
   const int c = var / d->mat, m = d->mat;
   const int qx = var + kMatPartner[c][0] * m, qy = var + kMatPartner[c][1] * m,
             qz = var + kMatPartner[c][2] * m;
   return (A(d, var, s, i - 1, j, k) * A(d, qx, s, i - 1, j, k) +
           A(d, var, s, i, j - 1, k) * A(d, qy, s, i, j - 1, k) +
           A(d, var, s, i, j, k - 1) * A(d, qz, s, i, j, k - 1) +
           P(var) * P(var) +
           A(d, var, s, i, j, k + 1) * A(d, qz, s, i, j, k + 1) +
           A(d, var, s, i, j + 1, k) * A(d, qy, s, i, j + 1, k) +
           A(d, var, s, i + 1, j, k) * A(d, qx, s, i + 1, j, k)) / 7.0 * (d->a1 + P(var));
}

/* stage 5 (ref/stencil.c:787-958): 27 terms in i, j, k order */
static double stencil0_27(amr_dev *d, int var, int s, int i, int j, int k)
{
   const int c = var / d->mat, m = d->mat;
   double sum = 0.0;
   int first = 1;
   for (int di = -1; di <= 1; di++)
      for (int dj = -1; dj <= 1; dj++)
         for (int dk = -1; dk <= 1; dk++) {
            int nzo = (di != 0) + (dj != 0) + (dk != 0);
            int u = nzo ? var + kMatPartner[c][nzo - 1] * m : var;
            double e = A(d, u, s, i + di, j + dj, k + dk);
            sum = first ? e : sum + e;
            first = 0;
         }
   return sum / (d->a1 + 27.0);
}
#undef P

static void stencil0(amr_dev *d, int var, int calc_stage)
{
   int nx = d->c.nx, ny = d->c.ny, nz = d->c.nz, op = calc_stage % 6;
   if (var == 0 || var >= 4 * d->mat) {   /* ref/stencil.c:45-47, 71-72 */
      stencil_avg(d, var, 7);
      return;
   }
   double *w = d->work;
#define W(i, j, k) w[((size_t) (i) * d->gy + (j)) * d->gz + (k)]
   for (int in = 0; in < d->n_active; in++) {
      int s = d->sorted[in];
      for (int i = 1; i <= nx; i++)
         for (int j = 1; j <= ny; j++)
            for (int k = 1; k <= nz; k++) {
               if (op == 0)
                  A(d, var, s, i, j, k) = stencil0_point(d, var, s, i, j, k);
               else if (op <= 3)
                  A(d, var, s, i, j, k) =
                     stencil0_sweep(d, var, s, i, j, k, op == 1, op == 2, op == 3);
               else
                  W(i, j, k) = op == 4 ? stencil0_7(d, var, s, i, j, k)
                                       : stencil0_27(d, var, s, i, j, k);
            }
      if (op >= 4)
         for (int i = 1; i <= nx; i++)
            for (int j = 1; j <= ny; j++)
               for (int k = 1; k <= nz; k++)
                  A(d, var, s, i, j, k) = W(i, j, k);
   }
#undef W
   /* stencil_check, ref/stencil.c:966-981 */
   for (int in = 0; in < d->n_active; in++) {
      int s = d->sorted[in];
      for (int i = 1; i <= nx; i++)
         for (int j = 1; j <= ny; j++)
            for (int k = 1; k <= nz; k++) {
               double *c = &A(d, var, s, i, j, k);
               *c = fabs(*c);
               if (*c >= 1.0) {
                  *c /= (d->a1 + d->a0[0] + *c);
                  d->checks[0]++;
               } else if (*c < 0.1) {
                  *c *= 10.0 - d->a1;
                  d->checks[1]++;
               }
            }
   }
}

int amr_dev_stencil_stage(amr_dev *d, int var, int calc_stage)
{
   if (var < 0 || var >= d->c.num_vars || calc_stage < 0)
      return fail(d, AMR_ERR_ARG, "stencil: variable out of range");
   if (d->c.stencil != 0) {
      stencil_avg(d, var, d->c.stencil);
      return AMR_OK;
   }
   if (!d->a0)
      return fail(d, AMR_ERR_STATE, "stencil 0: amr_dev_set_materials not called");
   stencil0(d, var, calc_stage);
   return AMR_OK;
}

int amr_dev_stencil(amr_dev *d, int var) { return amr_dev_stencil_stage(d, var, 0); }

int amr_dev_set_materials(amr_dev *d, double a1, const double *a0, int n)
{
   if (d->c.stencil != 0 || !a0 || n < 1 || n != d->c.num_vars / 4)
      return fail(d, AMR_ERR_ARG, "set_materials");
   free(d->a0);
   d->a0 = (double *) malloc(sizeof(double) * n);
   memcpy(d->a0, a0, sizeof(double) * n);
   d->mat = n;
   d->a1 = a1;
   return AMR_OK;
}

/* The oracle leaves refinement marking to the host driver's own code
 * (the reference's sequential passes); the device's one-launch marking is
 * checked against the reference's runs directly. */
int amr_dev_mark_refine(amr_dev *d, const amr_mark_params *p, const amr_mark_block *blocks,
                        int n_blocks, const int32_t *level_start, const amr_mark_parent *parents,
                        int n_parents, const amr_mark_object *objects, int8_t *block_refine,
                        int8_t *parent_refine)
{
   (void) p; (void) blocks; (void) n_blocks; (void) level_start; (void) parents;
   (void) n_parents; (void) objects; (void) block_refine; (void) parent_refine;
   return fail(d, AMR_ERR_STATE, "the CPU oracle marks refinement on the host");
}

int amr_dev_read_check_counts(amr_dev *d, int64_t counts[2])
{
   counts[0] = d->checks[0];
   counts[1] = d->checks[1];
   return AMR_OK;
}

int amr_dev_check_sum(amr_dev *d, int var, double *out)
{
   double sum = 0.0;
   for (int in = 0; in < d->n_active; in++) {
      int s = d->sorted[in];
      double bs = 0.0;
      for (int i = 1; i <= d->c.nx; i++)
         for (int j = 1; j <= d->c.ny; j++)
            for (int k = 1; k <= d->c.nz; k++)
               bs += A(d, var, s, i, j, k);
      sum += bs;
   }
   *out = sum;
   return AMR_OK;
}

int amr_dev_split(amr_dev *d, const amr_split_job *jobs, int n)
{
   int hx = d->c.nx / 2, hy = d->c.ny / 2, hz = d->c.nz / 2;
   for (int t = 0; t < n; t++) {
      const amr_split_job *j = &jobs[t];
      int o = j->octant;
      int i1 = (o % 2) * hx, j1 = ((o / 2) % 2) * hy, k1 = (o / 4) * hz;
      for (int v = 0; v < d->c.num_vars; v++)
         for (int i = 1; i <= hx; i++)
            for (int jj = 1; jj <= hy; jj++)
               for (int k = 1; k <= hz; k++) {
                  double q = A(d, v, j->parent, i + i1, jj + j1, k + k1) / 8.0;
                  for (int di = 0; di < 2; di++)
                     for (int dj = 0; dj < 2; dj++)
                        for (int dk = 0; dk < 2; dk++)
                           A(d, v, j->child, 2 * i - 1 + di, 2 * jj - 1 + dj, 2 * k - 1 + dk) = q;
               }
   }
   return AMR_OK;
}

/* The allocation-free variant computes the same cells (include/amr_dev.h). */
int amr_dev_split_noalloc(amr_dev *d, const amr_split_job *jobs, int n)
{
   return amr_dev_split(d, jobs, n);
}

int amr_dev_consolidate(amr_dev *d, const amr_merge_job *jobs, int n)
{
   int hx = d->c.nx / 2, hy = d->c.ny / 2, hz = d->c.nz / 2;
   for (int t = 0; t < n; t++) {
      const amr_merge_job *j = &jobs[t];
      int o = j->octant, c = j->child;
      int i1 = (o % 2) * hx, j1 = ((o / 2) % 2) * hy, k1 = (o / 4) * hz;
      for (int v = 0; v < d->c.num_vars; v++)
         for (int i = 1; i <= hx; i++)
            for (int jj = 1; jj <= hy; jj++)
               for (int k = 1; k <= hz; k++) {
                  int i2 = 2 * i - 1, j2 = 2 * jj - 1, k2 = 2 * k - 1;
                  A(d, v, j->dst, i + i1, jj + j1, k + k1) =
                     A(d, v, c, i2, j2, k2) + A(d, v, c, i2 + 1, j2, k2) +
                     A(d, v, c, i2, j2 + 1, k2) + A(d, v, c, i2 + 1, j2 + 1, k2) +
                     A(d, v, c, i2, j2, k2 + 1) + A(d, v, c, i2 + 1, j2, k2 + 1) +
                     A(d, v, c, i2, j2 + 1, k2 + 1) + A(d, v, c, i2 + 1, j2 + 1, k2 + 1);
               }
   }
   return AMR_OK;
}

int amr_dev_reserve_sums(amr_dev *d, int n_rows)
{
   free(d->sums);
   d->sums = (double *) calloc((size_t) (n_rows > 0 ? n_rows : 1) * d->c.num_vars, sizeof(double));
   d->sum_rows = n_rows;
   return AMR_OK;
}

int amr_dev_read_sums(amr_dev *d, int row0, int n_rows, double *out)
{
   if (row0 < 0 || row0 + n_rows > d->sum_rows)
      return fail(d, AMR_ERR_ARG, "sum rows");
   memcpy(out, d->sums + (size_t) row0 * d->c.num_vars,
          sizeof(double) * (size_t) n_rows * d->c.num_vars);
   return AMR_OK;
}

int amr_dev_stage(amr_dev *d, int start, int num, int stage, int permute, int sum_row)
{
   int rc = amr_dev_comm(d, start, num, stage, permute);
   if (rc) return rc;
   if (sum_row >= d->sum_rows)
      return fail(d, AMR_ERR_ARG, "sum row");
   for (int v = start; v < start + num; v++) {
      amr_dev_stencil_stage(d, v, stage);
      if (sum_row >= 0)
         amr_dev_check_sum(d, v, &d->sums[(size_t) sum_row * d->c.num_vars + v]);
   }
   return AMR_OK;
}

int amr_dev_sync(amr_dev *d) { (void) d; return AMR_OK; }
int amr_dev_set_timing(amr_dev *d, int on) { (void) d; (void) on; return AMR_OK; }
int amr_dev_read_timing(amr_dev *d, double ms[4])
{
   (void) d;
   ms[0] = ms[1] = ms[2] = ms[3] = 0.0;
   return AMR_OK;
}

/* The CPU restatement has one interior per record: the fused stage is the
 * plain stage (same cells and checksums by definition). */
int amr_dev_stage_fused(amr_dev *d, int start, int num, int stage, int permute, int sum_row)
{
   return amr_dev_stage(d, start, num, stage, permute, sum_row);
}

int amr_dev_read_exchange_timing(amr_dev *d, double ms[3])
{
   (void) d;
   ms[0] = ms[1] = ms[2] = 0.0;
   return AMR_OK;
}

int amr_dev_memory(const amr_dev *d, size_t *bytes, int *buffers)
{
   if (!d || !bytes)
      return AMR_ERR_ARG;
   *bytes = 0;
   if (buffers)
      *buffers = 0;
   return AMR_OK;
}

/* ------------------------------------------------------------------------
 * include/amr_batch.h on the CPU: the batch's definition restated as "each
 * member steps alone" (members in order, all variables per stage), which
 * is what the batched kernels must reproduce.  Launch shapes and tuning
 * have no meaning here and are accepted and ignored.
 * ---------------------------------------------------------------------- */
#include "amr_batch.h"

struct amr_batch {
   int device, precision, stencil;
   amr_dev **members;
   char *active;
   int n, cap;
   amr_batch_shape shape;
   char err[256];
};

int amr_batch_create(int device, int precision, int stencil, amr_batch **out)
{
   amr_batch *b = (amr_batch *) calloc(1, sizeof(amr_batch));
   *out = b;
   b->device = device;
   b->precision = precision;
   b->stencil = stencil;
   b->shape.vpw = 1;
   b->shape.exchange_vars = 8;
   b->shape.exchange_threads = 128;
   if (stencil != 7 && stencil != 27) {
      snprintf(b->err, sizeof(b->err), "stencil must be 7 or 27");
      return AMR_ERR_ARG;
   }
   if (precision != AMR_PREC_F64) {
      snprintf(b->err, sizeof(b->err), "the CPU oracle is FP64 only");
      return AMR_ERR_ARG;
   }
   return AMR_OK;
}

int amr_batch_destroy(amr_batch *b)
{
   if (!b)
      return AMR_OK;
   for (int i = 0; i < b->n; i++)
      amr_dev_destroy(b->members[i]);
   free(b->members);
   free(b->active);
   free(b);
   return AMR_OK;
}

const char *amr_batch_last_error(const amr_batch *b) { return b ? b->err : "no batch"; }

int amr_batch_add(amr_batch *b, const amr_dev_config *cfg, amr_dev **member)
{
   if (cfg->precision != b->precision || cfg->stencil != b->stencil) {
      snprintf(b->err, sizeof(b->err), "member differs from the batch");
      return AMR_ERR_ARG;
   }
   int rc = amr_dev_create(cfg, member);
   if (rc) {
      snprintf(b->err, sizeof(b->err), "%s", *member ? amr_dev_last_error(*member) : "create");
      return rc;
   }
   if (b->n == b->cap) {
      b->cap = b->cap ? 2 * b->cap : 16;
      b->members = (amr_dev **) realloc(b->members, sizeof(amr_dev *) * b->cap);
      b->active = (char *) realloc(b->active, b->cap);
   }
   b->members[b->n] = *member;
   b->active[b->n++] = 1;
   return AMR_OK;
}

int amr_batch_set_active(amr_batch *b, amr_dev *member, int on)
{
   for (int i = 0; i < b->n; i++)
      if (b->members[i] == member) {
         b->active[i] = on != 0;
         return AMR_OK;
      }
   return AMR_ERR_ARG;
}

int amr_batch_set_shape(amr_batch *b, const amr_batch_shape *shape)
{
   b->shape = *shape;
   return AMR_OK;
}

int amr_batch_get_shape(const amr_batch *b, amr_batch_shape *shape)
{
   *shape = b->shape;
   return AMR_OK;
}

int amr_batch_stage(amr_batch *b, int stage, int permute, int sum_row)
{
   for (int i = 0; i < b->n; i++)
      if (b->active[i]) {
         amr_dev *d = b->members[i];
         int rc = amr_dev_stage(d, 0, d->c.num_vars, stage, permute, sum_row);
         if (rc)
            return rc;
      }
   return AMR_OK;
}

int amr_batch_timestep(amr_batch *b, int stage0, int nstages, int checksum_freq, int permute)
{
   for (int s = 0; s < nstages; s++) {
      int row = (checksum_freq && !(s % checksum_freq)) ? s / checksum_freq : -1;
      int rc = amr_batch_stage(b, stage0 + s, permute, row);
      if (rc)
         return rc;
   }
   return AMR_OK;
}

int amr_batch_autotune(amr_batch *b, double *ms)
{
   (void) b;
   if (ms)
      *ms = 0.0;
   return AMR_OK;
}

int amr_batch_sync(amr_batch *b) { (void) b; return AMR_OK; }
int amr_batch_set_timing(amr_batch *b, int on) { (void) b; (void) on; return AMR_OK; }
int amr_batch_read_timing(amr_batch *b, double ms[4])
{
   (void) b;
   ms[0] = ms[1] = ms[2] = ms[3] = 0.0;
   return AMR_OK;
}

/* include/amr_adjoint.h: the reverse pass has its own checker (the torch
 * restatement used by the tests); the CPU oracle records nothing. */
#include "amr_adjoint.h"

int amr_dev_tape(amr_dev *d, int on)
{
   if (on)
      return fail(d, AMR_ERR_STATE, "the CPU oracle does not record a tape");
   return AMR_OK;
}

/* ---- halo message codec: pack_face / unpack_face (ref/comm.c:245-1461),
 * lengths ref/init.c:77-117.  Each direction's face plane is addressed as
 * P[a][b] with (a, b) the reference's loop axes: (j, k) for x, (i, k) for
 * y, (i, j) for z; `p` is the plane's index along the direction.  The
 * reference's three per-direction copies of every loop are one loop here
 * over (a, b).  Buffers are host memory in this CPU library. ---- */

static double *plane_cell(amr_dev *d, int m, int slot, int dir, int p, int a, int b)
{
   int i = dir == 0 ? p : a, j = dir == 0 ? a : dir == 1 ? p : b, k = dir == 2 ? p : b;
   return &A(d, m, slot, i, j, k);
}

static void plane_ext(const amr_dev *d, int dir, int *na, int *nb, int *nd)
{
   *na = dir == 0 ? d->c.ny : d->c.nx;
   *nb = dir == 2 ? d->c.ny : d->c.nz;
   *nd = dir == 0 ? d->c.nx : dir == 1 ? d->c.ny : d->c.nz;
}

/* One message; returns the number of values written (per call), or -1. */
static long pack_one(amr_dev *d, int code, int dir, int fc, int slot, int start, int num,
                     double *out)
{
   int na, nb, nd;
   plane_ext(d, dir, &na, &nb, &nd);
   const int c = fc % 10, p = fc >= 10 ? nd : 1, ha = na / 2, hb = nb / 2;
   const int lo_a = c % 2 == 0, lo_b = (c / 2) % 2 == 1;
   long n = 0;
   for (int m = start; m < start + num; m++) {
#define P(a, b) (*plane_cell(d, m, slot, dir, p, (a), (b)))
      if (code == 0 && c < 2) {                 /* ref/comm.c:264-269, 307-321, 353-367 */
         int ra = c == 1 && dir > 0, rb = c == 1 && dir == 2;
         for (int a = ra ? 0 : 1; a <= (ra ? na + 1 : na); a++)
            for (int b = rb ? 0 : 1; b <= (rb ? nb + 1 : nb); b++)
               out[n++] = P(a, b);
      } else if (code == 0 && c <= 5) {         /* :270-278 and twins */
         for (int a = 1; a < na; a += 2)
            for (int b = 1; b < nb; b += 2)
               out[n++] = P(a, b) + P(a, b + 1) + P(a + 1, b) + P(a + 1, b + 1);
      } else if (code == 0) {                   /* :279-299 and twins */
         for (int a = lo_a ? 1 : ha + 1; a <= (lo_a ? ha : na); a++)
            for (int b = lo_b ? 1 : hb + 1; b <= (lo_b ? hb : nb); b++)
               out[n++] = P(a, b) / 4.0;
      } else if (c < 2) {                       /* :413-418, 665-670 */
         for (int a = 0; a <= na + 1; a++)
            for (int b = 0; b <= nb + 1; b++)
               out[n++] = P(a, b);
      } else if (c <= 5) {                      /* :419-465, 670-718 */
         /* rows: the ring row 0 first (even cases), the pairs, the ring
          * row na+1 last (odd cases); in each row the ring column 0 first
          * (cases 2, 3) or nb+1 last (4, 5); pairs sum as [a][b] + [a][b+1]
          * + [a+1][b] + [a+1][b+1] */
         int ga[64][2], gb[64][2], nga = 0, ngb = 0;
         if (lo_a) { ga[nga][0] = ga[nga][1] = 0; nga++; }
         for (int a = 1; a < na; a += 2, nga++) { ga[nga][0] = a; ga[nga][1] = a + 1; }
         if (!lo_a) { ga[nga][0] = ga[nga][1] = na + 1; nga++; }
         if (lo_b) { gb[ngb][0] = gb[ngb][1] = 0; ngb++; }
         for (int b = 1; b < nb; b += 2, ngb++) { gb[ngb][0] = b; gb[ngb][1] = b + 1; }
         if (!lo_b) { gb[ngb][0] = gb[ngb][1] = nb + 1; ngb++; }
         for (int x = 0; x < nga; x++)
            for (int y = 0; y < ngb; y++) {
               double v = 0.0;
               int first = 1;
               for (int a = ga[x][0]; a <= ga[x][1]; a++)
                  for (int b = gb[y][0]; b <= gb[y][1]; b++) {
                     v = first ? P(a, b) : v + P(a, b);
                     first = 0;
                  }
               out[n++] = v;
            }
      } else if (code == 1) {                   /* :466-486 and twins */
         for (int a = lo_a ? 0 : ha; a <= (lo_a ? ha + 1 : na + 1); a++)
            for (int b = lo_b ? 0 : hb; b <= (lo_b ? hb + 1 : nb + 1); b++)
               out[n++] = P(a, b) / 4.0;
      } else {                                  /* code 2: :719-763 and twins */
         int a0 = lo_a ? 1 : ha + 1, a1 = lo_a ? ha : na;
         int b0 = lo_b ? 1 : hb + 1, b1 = lo_b ? hb : nb;
         for (int r = a0 - 1; r <= a1 + 1; r++)
            for (int rep = 0; rep < ((r >= a0 && r <= a1) ? 2 : 1); rep++) {
               out[n++] = P(r, b0 - 1) / 4.0;
               for (int b = b0; b <= b1; b++) {
                  out[n] = out[n + 1] = P(r, b) / 4.0;
                  n += 2;
               }
               out[n++] = P(r, b1 + 1) / 4.0;
            }
      }
   }
   return n;
}

static long unpack_one(amr_dev *d, int code, int dir, int fc, int slot, int start, int num,
                       const double *in)
{
   int na, nb, nd;
   plane_ext(d, dir, &na, &nb, &nd);
   const int c = fc % 10, p = fc >= 10 ? nd + 1 : 0, ha = na / 2, hb = nb / 2;
   const int lo_a = c % 2 == 0, lo_b = (c / 2) % 2 == 1;
   long n = 0;
   for (int m = start; m < start + num; m++) {
      if (code == 0 && c < 2) {                 /* ref/comm.c:1015-1019 and twins */
         int ra = c == 1 && dir > 0, rb = c == 1 && dir == 2;
         for (int a = ra ? 0 : 1; a <= (ra ? na + 1 : na); a++)
            for (int b = rb ? 0 : 1; b <= (rb ? nb + 1 : nb); b++)
               P(a, b) = in[n++];
      } else if (code == 0 && c <= 5) {         /* :1020-1028 and twins */
         for (int a = 1; a < na; a += 2)
            for (int b = 1; b < nb; b += 2, n++)
               P(a, b) = P(a, b + 1) = P(a + 1, b) = P(a + 1, b + 1) = in[n];
      } else if (code == 0) {                   /* :1029-1048 and twins */
         for (int a = lo_a ? 1 : ha + 1; a <= (lo_a ? ha : na); a++)
            for (int b = lo_b ? 1 : hb + 1; b <= (lo_b ? hb : nb); b++)
               P(a, b) = in[n++];
      } else if (c < 2 || (code == 2 && c <= 5)) {   /* :1162-1166, 1365-1369 */
         for (int a = 0; a <= na + 1; a++)
            for (int b = 0; b <= nb + 1; b++)
               P(a, b) = in[n++];
      } else if (c <= 5) {                      /* code 1: :1167-1198 and twins */
         for (int a = 0; a <= na + 1; a += (a == 0 || a == na + 1) ? 1 : 2) {
            int a1 = (a == 0 || a == na + 1) ? a : a + 1;
            for (int b = 0; b <= nb + 1; b += (b == 0 || b == nb + 1) ? 1 : 2, n++) {
               int b1 = (b == 0 || b == nb + 1) ? b : b + 1;
               for (int x = a; x <= a1; x++)
                  for (int y = b; y <= b1; y++)
                     P(x, y) = in[n];
            }
         }
      } else {                                  /* codes 1, 2: :1199-1218 and twins */
         for (int a = lo_a ? 0 : ha + 1; a <= (lo_a ? ha : na + 1); a++)
            for (int b = lo_b ? 0 : hb + 1; b <= (lo_b ? hb : nb + 1); b++)
               P(a, b) = in[n++];
      }
#undef P
   }
   return n;
}

int amr_dev_face_msg_len(amr_dev *d, int code, int dir, int face_case, int unpack)
{
   static double sink[2 * 1024 * 1024];
   if (!d || code < 0 || code > 2 || dir < 0 || dir > 2 || face_case < 0 || face_case > 19)
      return fail(d, AMR_ERR_ARG, "face_msg_len: bad arguments");
   if (d->c.nx > 120 || d->c.ny > 120 || d->c.nz > 120)
      return fail(d, AMR_ERR_ARG, "face_msg_len: blocks above 120 cells per axis");
   if ((size_t) (d->gx + 2) * (d->gy + 2) * 4 > sizeof(sink) / sizeof(double))
      return fail(d, AMR_ERR_ARG, "face_msg_len: block too large for this CPU library");
   /* run the format on slot 0, variable 0 and count (unpack writes the
    * values back, leaving the cells as they were) */
   if (!unpack)
      return (int) pack_one(d, code, dir, face_case, 0, 0, 1, sink);
   double *save = (double *) malloc(sizeof(double) * d->slab);
   memcpy(save, &A(d, 0, 0, 0, 0, 0), sizeof(double) * d->slab);
   long n = unpack_one(d, code, dir, face_case, 0, 0, 1, sink);
   memcpy(&A(d, 0, 0, 0, 0, 0), save, sizeof(double) * d->slab);
   free(save);
   return (int) n;
}

static int codec_args(amr_dev *d, int code, int dir, int start, int num, const amr_face_msg *msgs,
                      int n, const void *buf, size_t buf_elems, int unpack)
{
   if (!d || code < 0 || code > 2 || dir < 0 || dir > 2 || start < 0 || num < 1 ||
       start + num > d->c.num_vars || n < 0 || (n && (!msgs || !buf)))
      return fail(d, AMR_ERR_ARG, "pack/unpack_faces: bad arguments");
   for (int t = 0; t < n; t++) {
      if (msgs[t].slot < 0 || msgs[t].slot >= d->c.max_blocks || msgs[t].face_case < 0 ||
          msgs[t].face_case > 19 || msgs[t].offset < 0)
         return fail(d, AMR_ERR_ARG, "pack/unpack_faces: bad message");
      int len = amr_dev_face_msg_len(d, code, dir, msgs[t].face_case, unpack);
      if ((size_t) msgs[t].offset + (size_t) num * len > buf_elems)
         return fail(d, AMR_ERR_ARG, "pack/unpack_faces: message ends past the buffer");
   }
   return AMR_OK;
}

int amr_dev_pack_faces(amr_dev *d, int code, int dir, int start, int num_comm,
                       const amr_face_msg *msgs, int n, void *buf, size_t buf_elems)
{
   int rc = codec_args(d, code, dir, start, num_comm, msgs, n, buf, buf_elems, 0);
   for (int t = 0; !rc && t < n; t++)
      pack_one(d, code, dir, msgs[t].face_case, msgs[t].slot, start, num_comm,
               (double *) buf + msgs[t].offset);
   return rc;
}

int amr_dev_unpack_faces(amr_dev *d, int code, int dir, int start, int num_comm,
                         const amr_face_msg *msgs, int n, const void *buf, size_t buf_elems)
{
   int rc = codec_args(d, code, dir, start, num_comm, msgs, n, buf, buf_elems, 1);
   for (int t = 0; !rc && t < n; t++)
      unpack_one(d, code, dir, msgs[t].face_case, msgs[t].slot, start, num_comm,
                 (const double *) buf + msgs[t].offset);
   return rc;
}

/* The device's streaming bandwidth has no CPU meaning: no device here. */
int amr_dev_copy_bandwidth(size_t bytes, int reps, double *copy_gbs, double *read_gbs)
{
   (void) bytes;
   (void) reps;
   (void) copy_gbs;
   (void) read_gbs;
   return AMR_ERR_NODEV;
}

/* Largest |a - b| over the active blocks' interior cells, per variable
 * (FP64 contexts only in this CPU library). */
int amr_dev_compare_cells(amr_dev *a, amr_dev *b, int start, int num, double *max_abs)
{
   if (!a || !b || !max_abs || start < 0 || num < 1 || start + num > a->c.num_vars ||
       a->c.nx != b->c.nx || a->c.ny != b->c.ny || a->c.nz != b->c.nz ||
       a->c.num_vars != b->c.num_vars || a->n_active != b->n_active)
      return fail(a, AMR_ERR_ARG, "compare_cells: bad arguments or different contexts");
   for (int in = 0; in < a->n_active; in++)
      if (a->sorted[in] != b->sorted[in])
         return fail(a, AMR_ERR_STATE, "compare_cells: the contexts hold different meshes");
   for (int v = start; v < start + num; v++) {
      double m = 0.0;
      for (int in = 0; in < a->n_active; in++) {
         const int s = a->sorted[in];
         for (int i = 1; i <= a->c.nx; i++)
            for (int j = 1; j <= a->c.ny; j++)
               for (int k = 1; k <= a->c.nz; k++) {
                  double d = fabs(A(a, v, s, i, j, k) - A(b, v, s, i, j, k));
                  if (d > m || d != d)
                     m = d != d ? INFINITY : d;
               }
      }
      max_abs[v - start] = m;
   }
   return AMR_OK;
}
