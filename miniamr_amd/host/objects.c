/*
 * objects.c -- object motion and the host's refinement marking
 * (check_objects); the intersection tests live in geometry.h.
 *
 * A quirk of the reference is kept: move() reflects only for bounce == 1;
 * the bounce == 2 branch of ref/move.c:37-63 is unreachable (it is the
 * else of the inner if), so bounce 2 moves like bounce 0.
 */
#include <math.h>
#include "amr.h"
#include "geometry.h"

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

/* ref/move.c:143-623 */
int amr_check_block(const amr_run *R, double cor[3][2])
{
   int intersect = 0;
   for (int i = 0; i < R->P.num_objects; i++) {
      const amr_object *o = &R->P.objects[i];
      if (intersect > 0 || o->size[0] < 0.0 || o->size[1] < 0.0 || o->size[2] < 0.0)
         continue;
      if (amr_geo_known_type(o->type))
         intersect = amr_geo_object(o->type, o->cen, o->size, (const double (*)[2]) cor);
      else
         printf("undefined object %d\n", o->type);
   }
   return intersect;
}

static void grow(const amr_run *R, double cor[3][2], int sz, double f)
{
   const int n[3] = { R->P.nx, R->P.ny, R->P.nz };
   amr_geo_grow(cor, sz, f, n, R->mesh_size);
}

/* ref/move.c:65-141 */
void amr_check_objects(amr_run *R)
{
   int na = amr_n_active(R);
   for (int in = 0; in < na; in++) {
      amr_block *b = &R->blocks[R->sorted_slot[in]];
      int sz = R->p2[R->P.num_refine - b->level];
      double cor[3][2];
      amr_geo_extent(cor, b->cen, sz, R->mesh_size);
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
