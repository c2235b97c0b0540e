/*
 * objects.c -- object motion and block/object intersection.
 *
 * Refinement decisions must match the reference bit for bit, so every
 * comparison below evaluates the same IEEE expression as the reference:
 * squared terms are ((corner - centre) / size) squared and summed in axis
 * order, single-axis tests use fabs(corner - centre) < size, and the
 * quirks of the reference are kept (they decide which blocks refine):
 *   - move(): only bounce == 1 reflects; the bounce == 2 branch of
 *     ref/move.c:37-63 is unreachable (it is the else of the inner if), so
 *     bounce 2 moves like bounce 0;
 *   - rectangle surface (type 0): when the block overlaps in x but neither
 *     straddles in x nor overlaps in y, the reference still refines when z
 *     overlaps and straddles (ref/move.c:183-196);
 *   - cylinder surfaces (20/22/24): the "straddles an end" test always uses
 *     the x extent whatever the axis (ref/move.c:517-521).
 */
#include <math.h>
#include "amr.h"

/* ref/move.c:37-63 */
void amr_move_objects(amr_run *R, double delta)
{
   for (int i = 0; i < R->P.num_objects; i++) {
      amr_object *o = &R->P.objects[i];
      for (int d = 0; d < 3; d++) {
         o->cen[d] += delta * o->move[d];
         o->size[d] += delta * o->inc[d];
         if (o->bounce == 1) {
            if (o->cen[d] >= 1.0) {
               o->cen[d] = 2.0 - o->cen[d];
               o->move[d] = -o->move[d];
            } else if (o->cen[d] <= 0.0) {
               o->cen[d] = 0.0 - o->cen[d];
               o->move[d] = -o->move[d];
            }
         }
      }
   }
}

/* ((cor - c) / s)^2 */
static double sq_ratio(double cor, double c, double s)
{
   double t = (cor - c) / s;
   return t * t;
}

/* Where the object's centre lies relative to the block on one axis:
 * inside (*in = 1) or on the low/high side (*v = 0/1).  With `mid`, an
 * inside centre also picks the nearer half (ref/move.c:271-281). */
static void locate(const double cor[2], double c, int mid, int *in, int *v)
{
   *in = 0;
   *v = 0;
   if (c < cor[0])
      *v = 0;
   else if (c > cor[1])
      *v = 1;
   else {
      *in = 1;
      if (mid)
         *v = (c < (cor[0] + cor[1]) / 2.0) ? 0 : 1;
   }
}

/* "Near" test over the axes listed in ax[0..na): the nearest corner of the
 * block lies inside the ellipsoid/ellipse.  One axis: |cor - c| < s; more:
 * sum of squared ratios < 1 (the reference's two forms are kept). */
static int near_inside(const double cor[3][2], const double *c, const double *s,
                       const int *ax, const int *v, int na)
{
   if (na == 0)
      return 1;
   if (na == 1)
      return fabs(cor[ax[0]][v[ax[0]]] - c[ax[0]]) < s[ax[0]];
   double acc = sq_ratio(cor[ax[0]][v[ax[0]]], c[ax[0]], s[ax[0]]);
   for (int i = 1; i < na; i++)
      acc = acc + sq_ratio(cor[ax[i]][v[ax[i]]], c[ax[i]], s[ax[i]]);
   return acc < 1.0;
}

static int spheroid_surface(const amr_object *o, const double cor[3][2])
{
   int in[3], v[3], ax[3], na = 0;
   for (int d = 0; d < 3; d++) {
      locate(cor[d], o->cen[d], 1, &in[d], &v[d]);
      if (!in[d])
         ax[na++] = d;
   }
   double far = sq_ratio(cor[0][1 - v[0]], o->cen[0], o->size[0]) +
                sq_ratio(cor[1][1 - v[1]], o->cen[1], o->size[1]) +
                sq_ratio(cor[2][1 - v[2]], o->cen[2], o->size[2]);
   if (na == 0)
      return far > 1.0;
   return near_inside(cor, o->cen, o->size, ax, v, na) && far > 1.0;
}

static int spheroid_solid(const amr_object *o, const double cor[3][2])
{
   int in[3], v[3], ax[3], na = 0;
   for (int d = 0; d < 3; d++) {
      locate(cor[d], o->cen[d], 0, &in[d], &v[d]);
      if (!in[d])
         ax[na++] = d;
   }
   return near_inside(cor, o->cen, o->size, ax, v, na);
}

/* Half-space cut that turns a spheroid into a hemispheroid
 * (ref/move.c:356-370 and 450-464): returns 1 if the block lies entirely
 * in the excluded half.  k = (type - 2) / 2 for surfaces, (type - 3) / 2
 * for solids: 1 = below x, 2 = above x, 3/4 = y, 5/6 = z. */
static int hemi_excluded(int k, const amr_object *o, const double cor[3][2])
{
   if (k <= 0)
      return 0;
   int d = (k - 1) / 2;
   if (k % 2)
      return cor[d][1] < o->cen[d];
   return cor[d][0] > o->cen[d];
}

static int overlap(const double cor[2], double c, double s)
{
   return cor[1] > (c - s) && cor[0] < (c + s);
}

static int straddle(const double cor[2], double c, double s)
{
   return (cor[0] < (c - s) && cor[1] < (c + s)) ||
          (cor[0] > (c - s) && cor[1] > (c + s));
}

static int rectangle_surface(const amr_object *o, const double cor[3][2])
{
   const double *c = o->cen, *s = o->size;
   if (!overlap(cor[0], c[0], s[0]))
      return 0;
   if (straddle(cor[0], c[0], s[0]))
      return overlap(cor[1], c[1], s[1]) && overlap(cor[2], c[2], s[2]);
   if (overlap(cor[1], c[1], s[1]))
      return straddle(cor[1], c[1], s[1]) && overlap(cor[2], c[2], s[2]);
   return overlap(cor[2], c[2], s[2]) && straddle(cor[2], c[2], s[2]);
}

static int rectangle_solid(const amr_object *o, const double cor[3][2])
{
   return overlap(cor[0], o->cen[0], o->size[0]) &&
          overlap(cor[1], o->cen[1], o->size[1]) &&
          overlap(cor[2], o->cen[2], o->size[2]);
}

/* Cylinders: axis ca, cross-section axes c1, c2 (ref/move.c:477-620). */
static int cylinder(const amr_object *o, const double cor[3][2], int surface)
{
   static const int axes[3][3] = { {0, 1, 2}, {1, 2, 0}, {2, 0, 1} };
   int which = ((o->type - 20) / 2);
   int ca = axes[which][0], c1 = axes[which][1], c2 = axes[which][2];
   const double *c = o->cen, *s = o->size;
   if (!overlap(cor[ca], c[ca], s[ca]))
      return 0;
   int in[3] = { 0, 0, 0 }, v[3] = { 0, 0, 0 }, ax[2], na = 0;
   locate(cor[c1], c[c1], surface, &in[c1], &v[c1]);
   locate(cor[c2], c[c2], surface, &in[c2], &v[c2]);
   if (!in[c1]) ax[na++] = c1;
   if (!in[c2]) ax[na++] = c2;
   int near = near_inside(cor, c, s, ax, v, na);
   if (!surface)
      return near;
   if (straddle(cor[0], c[0], s[0]))
      return near;
   double far = sq_ratio(cor[c1][1 - v[c1]], c[c1], s[c1]) +
                sq_ratio(cor[c2][1 - v[c2]], c[c2], s[c2]);
   return near && far > 1.0;
}

/* ref/move.c:143-623 */
int amr_check_block(const amr_run *R, double cor[3][2])
{
   int intersect = 0;
   for (int i = 0; i < R->P.num_objects; i++) {
      const amr_object *o = &R->P.objects[i];
      if (intersect > 0 || o->size[0] < 0.0 || o->size[1] < 0.0 || o->size[2] < 0.0)
         continue;
      int t = o->type;
      if (t == 0)
         intersect = rectangle_surface(o, cor);
      else if (t == 1)
         intersect = rectangle_solid(o, cor);
      else if (t >= 2 && t <= 14 && !(t % 2))
         intersect = spheroid_surface(o, cor) && !hemi_excluded((t - 2) / 2, o, cor);
      else if (t >= 3 && t <= 15 && (t % 2))
         intersect = spheroid_solid(o, cor) && !hemi_excluded((t - 3) / 2, o, cor);
      else if (t == 20 || t == 22 || t == 24)
         intersect = cylinder(o, cor, 1);
      else if (t == 21 || t == 23 || t == 25)
         intersect = cylinder(o, cor, 0);
      else
         printf("undefined object %d\n", t);
   }
   return intersect;
}

/* Block extent in the unit cube, optionally grown by `ghosts` cells of a
 * block of half-size sz (ref/move.c:79-101). */
static void grow(const amr_run *R, double cor[3][2], int sz, double f)
{
   const int n[3] = { R->P.nx, R->P.ny, R->P.nz };
   for (int d = 0; d < 3; d++) {
      cor[d][0] -= f * (((double) sz) / ((double) n[d])) / ((double) R->mesh_size[d]);
      cor[d][1] += f * (((double) sz) / ((double) n[d])) / ((double) R->mesh_size[d]);
   }
}

/* ref/move.c:65-141 */
void amr_check_objects(amr_run *R)
{
   int na = amr_n_active(R);
   for (int in = 0; in < na; in++) {
      amr_block *b = &R->blocks[R->sorted_slot[in]];
      int sz = R->p2[R->P.num_refine - b->level];
      double cor[3][2];
      for (int d = 0; d < 3; d++) {
         cor[d][0] = ((double) (b->cen[d] - sz)) / ((double) R->mesh_size[d]);
         cor[d][1] = ((double) (b->cen[d] + sz)) / ((double) R->mesh_size[d]);
      }
      if (R->P.refine_ghost)
         grow(R, cor, sz, 2.0);
      if (amr_check_block(R, cor))
         b->refine = 1;
      else if (R->P.refine_ghost && b->level) {
         /* would the parent refine again?  then leave the block alone */
         grow(R, cor, R->p2[R->P.num_refine - b->level + 1], 1.0);
         if (amr_check_block(R, cor))
            b->refine = 0;
      }
      if ((b->level == R->P.num_refine && b->refine == 1) || !b->refine) {
         b->refine = 0;
         if (b->parent != -1) {
            amr_parent *pp = &R->parents[b->parent];
            pp->refine = 0;
            for (int c = 0; c < 8; c++)
               if (pp->child[c] >= 0 && R->blocks[pp->child[c]].refine == -1)
                  R->blocks[pp->child[c]].refine = 0;
         }
      }
      if (!b->level && b->refine == -1)
         b->refine = 0;
   }
}
