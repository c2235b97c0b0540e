/*
 * geometry.h -- block/object intersection (ref/move.c:143-623), shared by
 * the host driver (objects.c) and the device marking kernel (marking.hip).
 *
 * Refinement decisions must match the reference bit for bit, so every
 * comparison below evaluates the same IEEE expression as the reference:
 * squared terms are ((corner - centre) / size) squared and summed in axis
 * order, single-axis tests use fabs(corner - centre) < size.  Both builds
 * compile this one source without contraction (-ffp-contract=off) and
 * FP64 +, -, *, / and comparisons are correctly rounded on both, so host
 * and device decide alike.  The quirks of the reference are kept (they
 * decide which blocks refine):
 *   - rectangle surface (type 0): when the block overlaps in x but neither
 *     straddles in x nor overlaps in y, the reference still refines when z
 *     overlaps and straddles (ref/move.c:183-196);
 *   - cylinder surfaces (20/22/24): the "straddles an end" test always uses
 *     the x extent whatever the axis (ref/move.c:517-521).
 */
#ifndef AMR_GEOMETRY_H
#define AMR_GEOMETRY_H

#include <math.h>

#ifdef __HIPCC__
#define AMR_GEO __host__ __device__ static inline
#else
#define AMR_GEO static inline
#endif

/* ((cor - c) / s)^2 */
AMR_GEO double amr_geo_sq(double cor, double c, double s)
{
   double t = (cor - c) / s;
   return t * t;
}

/* Where the object's centre lies relative to the block on one axis:
 * inside (*in = 1) or on the low/high side (*v = 0/1).  With `mid`, an
 * inside centre also picks the nearer half (ref/move.c:271-281). */
AMR_GEO void amr_geo_locate(const double cor[2], double c, int mid, int *in, int *v)
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
AMR_GEO int amr_geo_near(const double cor[3][2], const double *c, const double *s,
                         const int *ax, const int *v, int na)
{
   if (na == 0)
      return 1;
   if (na == 1)
      return fabs(cor[ax[0]][v[ax[0]]] - c[ax[0]]) < s[ax[0]];
   double acc = amr_geo_sq(cor[ax[0]][v[ax[0]]], c[ax[0]], s[ax[0]]);
   for (int i = 1; i < na; i++)
      acc = acc + amr_geo_sq(cor[ax[i]][v[ax[i]]], c[ax[i]], s[ax[i]]);
   return acc < 1.0;
}

AMR_GEO int amr_geo_spheroid_surface(const double *c, const double *s, const double cor[3][2])
{
   int in[3], v[3], ax[3], na = 0;
   for (int d = 0; d < 3; d++) {
      amr_geo_locate(cor[d], c[d], 1, &in[d], &v[d]);
      if (!in[d])
         ax[na++] = d;
   }
   double far = amr_geo_sq(cor[0][1 - v[0]], c[0], s[0]) +
                amr_geo_sq(cor[1][1 - v[1]], c[1], s[1]) +
                amr_geo_sq(cor[2][1 - v[2]], c[2], s[2]);
   if (na == 0)
      return far > 1.0;
   return amr_geo_near(cor, c, s, ax, v, na) && far > 1.0;
}

AMR_GEO int amr_geo_spheroid_solid(const double *c, const double *s, const double cor[3][2])
{
   int in[3], v[3], ax[3], na = 0;
   for (int d = 0; d < 3; d++) {
      amr_geo_locate(cor[d], c[d], 0, &in[d], &v[d]);
      if (!in[d])
         ax[na++] = d;
   }
   return amr_geo_near(cor, c, s, ax, v, na);
}

/* Half-space cut that turns a spheroid into a hemispheroid
 * (ref/move.c:356-370 and 450-464): returns 1 if the block lies entirely
 * in the excluded half.  k = (type - 2) / 2 for surfaces, (type - 3) / 2
 * for solids: 1 = below x, 2 = above x, 3/4 = y, 5/6 = z. */
AMR_GEO int amr_geo_hemi_excluded(int k, const double *c, const double cor[3][2])
{
   if (k <= 0)
      return 0;
   int d = (k - 1) / 2;
   if (k % 2)
      return cor[d][1] < c[d];
   return cor[d][0] > c[d];
}

AMR_GEO int amr_geo_overlap(const double cor[2], double c, double s)
{
   return cor[1] > (c - s) && cor[0] < (c + s);
}

AMR_GEO int amr_geo_straddle(const double cor[2], double c, double s)
{
   return (cor[0] < (c - s) && cor[1] < (c + s)) ||
          (cor[0] > (c - s) && cor[1] > (c + s));
}

AMR_GEO int amr_geo_rectangle_surface(const double *c, const double *s, const double cor[3][2])
{
   if (!amr_geo_overlap(cor[0], c[0], s[0]))
      return 0;
   if (amr_geo_straddle(cor[0], c[0], s[0]))
      return amr_geo_overlap(cor[1], c[1], s[1]) && amr_geo_overlap(cor[2], c[2], s[2]);
   if (amr_geo_overlap(cor[1], c[1], s[1]))
      return amr_geo_straddle(cor[1], c[1], s[1]) && amr_geo_overlap(cor[2], c[2], s[2]);
   return amr_geo_overlap(cor[2], c[2], s[2]) && amr_geo_straddle(cor[2], c[2], s[2]);
}

AMR_GEO int amr_geo_rectangle_solid(const double *c, const double *s, const double cor[3][2])
{
   return amr_geo_overlap(cor[0], c[0], s[0]) && amr_geo_overlap(cor[1], c[1], s[1]) &&
          amr_geo_overlap(cor[2], c[2], s[2]);
}

/* Cylinders of type 20 + 2w / 21 + 2w: axis w, cross-section axes w+1 and
 * w+2 (mod 3) (ref/move.c:477-620). */
AMR_GEO int amr_geo_cylinder(int type, const double *c, const double *s, const double cor[3][2],
                             int surface)
{
   int which = (type - 20) / 2;
   int ca = which, c1 = (which + 1) % 3, c2 = (which + 2) % 3;
   if (!amr_geo_overlap(cor[ca], c[ca], s[ca]))
      return 0;
   int in[3] = { 0, 0, 0 }, v[3] = { 0, 0, 0 }, ax[2], na = 0;
   amr_geo_locate(cor[c1], c[c1], surface, &in[c1], &v[c1]);
   amr_geo_locate(cor[c2], c[c2], surface, &in[c2], &v[c2]);
   if (!in[c1]) ax[na++] = c1;
   if (!in[c2]) ax[na++] = c2;
   int near = amr_geo_near(cor, c, s, ax, v, na);
   if (!surface)
      return near;
   if (amr_geo_straddle(cor[0], c[0], s[0]))
      return near;
   double far = amr_geo_sq(cor[c1][1 - v[c1]], c[c1], s[c1]) +
                amr_geo_sq(cor[c2][1 - v[c2]], c[c2], s[c2]);
   return near && far > 1.0;
}

/* Whether an object type is one check_block knows. */
AMR_GEO int amr_geo_known_type(int t)
{
   return (t >= 0 && t <= 15) || (t >= 20 && t <= 25);
}

/* One object against a block extent (the body of ref/move.c:150-621's
 * loop); the type must be a known one. */
AMR_GEO int amr_geo_object(int t, const double *c, const double *s, const double cor[3][2])
{
   if (t == 0)
      return amr_geo_rectangle_surface(c, s, cor);
   if (t == 1)
      return amr_geo_rectangle_solid(c, s, cor);
   if (t >= 2 && t <= 14 && !(t % 2))
      return amr_geo_spheroid_surface(c, s, cor) && !amr_geo_hemi_excluded((t - 2) / 2, c, cor);
   if (t >= 3 && t <= 15 && (t % 2))
      return amr_geo_spheroid_solid(c, s, cor) && !amr_geo_hemi_excluded((t - 3) / 2, c, cor);
   if (t == 20 || t == 22 || t == 24)
      return amr_geo_cylinder(t, c, s, cor, 1);
   return amr_geo_cylinder(t, c, s, cor, 0);
}

/* Block extent in the unit cube (ref/move.c:79-101): half-size sz in
 * finest half-cell units, grown by `f` ghost cells of a block of half-size
 * gsz when f > 0.  n[] is the block size in cells. */
AMR_GEO void amr_geo_extent(double cor[3][2], const int cen[3], int sz, const int mesh[3])
{
   for (int d = 0; d < 3; d++) {
      cor[d][0] = ((double) (cen[d] - sz)) / ((double) mesh[d]);
      cor[d][1] = ((double) (cen[d] + sz)) / ((double) mesh[d]);
   }
}

AMR_GEO void amr_geo_grow(double cor[3][2], int gsz, double f, const int n[3], const int mesh[3])
{
   for (int d = 0; d < 3; d++) {
      cor[d][0] -= f * (((double) gsz) / ((double) n[d])) / ((double) mesh[d]);
      cor[d][1] += f * (((double) gsz) / ((double) n[d])) / ((double) mesh[d]);
   }
}

#endif /* AMR_GEOMETRY_H */
