/*
 * refine.c -- refinement decisions and the topology half of split and
 * consolidate, for one rank.
 *
 * Decisions (ref/refine.c:38-542): reset, mark by objects, propagate the
 * 2:1 balance level by level to a fixpoint, split, then coarsen.  On one
 * rank every neighbour is local, so all branches of the reference that
 * handle off-rank neighbours, message lists, redistribution and load
 * balance are no-ops and are left out; what remains must make the same
 * choices in the same order, because the slot each new block lands in
 * decides which stale ghost values a 27-point stencil later reads.
 *
 * Slot allocation follows the reference exactly:
 *   split:        children take the first free slots scanning upward from a
 *                 cursor that only moves back when a parent's own slot is
 *                 freed (ref/block.c:116-127, 304-307);
 *   consolidate:  the merged block takes the lowest free slot
 *                 (ref/block.c:370-377);
 *   parent records: lowest free record (ref/block.c:82-90).
 * The cell data of each split or merge is queued as a device job and
 * flushed to the device before the next step that reads cells.
 */
#include <string.h>
#include "amr.h"

static int opposite(int f) { return (f / 2) * 2 + (f + 1) % 2; }

/* children on each side of a parent, ref/block.c:43-45 */
static const int side[6][2][2] = { { {0, 2}, {4, 6} }, { {1, 3}, {5, 7} },
                                   { {0, 1}, {4, 5} }, { {2, 3}, {6, 7} },
                                   { {0, 1}, {2, 3} }, { {4, 5}, {6, 7} } };
/* octant offset to the sibling across face c, ref/block.c:46 */
static const int across[6] = { 1, -1, 2, -2, 4, -4 };

static void relax_children(amr_run *R, amr_parent *pp, int *changes)
{
   for (int c = 0; c < 8; c++)
      if (pp->child[c] >= 0 && R->blocks[pp->child[c]].refine == -1) {
         R->blocks[pp->child[c]].refine = 0;
         if (changes)
            (*changes)++;
      }
}

/* ref/refine.c:500-527 */
static void reset_all(amr_run *R)
{
   int na = amr_n_active(R);
   for (int in = 0; in < na; in++)
      R->blocks[R->sorted_slot[in]].refine = -1;
   for (int p = 0; p < R->max_active_parent; p++) {
      amr_parent *pp = &R->parents[p];
      if (pp->number < 0)
         continue;
      pp->refine = -1;
      for (int c = 0; c < 8; c++)
         if (pp->child[c] < 0)
            pp->refine = 0;   /* a child is itself a parent */
      if (pp->refine == 0)
         relax_children(R, pp, NULL);
   }
}

/* 2:1 balance fixpoint, ref/refine.c:325-496.  Returns blocks to split. */
int amr_refine_level(amr_run *R)
{
   for (int level = R->cur_max_level; level >= 0; level--) {
      int changes;
      /* blocks that refine: same-level neighbours may not coarsen, coarser
       * neighbours must refine */
      do {
         changes = 0;
         for (int in = R->sorted_index[level]; in < R->sorted_index[level + 1]; in++) {
            amr_block *b = &R->blocks[R->sorted_slot[in]];
            if (b->level != level)
               continue;
            if (b->refine == 1) {
               if (b->parent != -1) {
                  amr_parent *pp = &R->parents[b->parent];
                  if (pp->refine == -1)
                     pp->refine = 0;
                  relax_children(R, pp, &changes);
               }
               for (int f = 0; f < 6; f++) {
                  if (b->nei_level[f] == level) {
                     amr_block *nb = &R->blocks[b->nei[f][0][0]];
                     if (nb->refine == -1) {
                        nb->refine = 0;
                        changes++;
                        if (nb->parent != -1) {
                           amr_parent *pp = &R->parents[nb->parent];
                           if (pp->refine == -1)
                              pp->refine = 0;
                           relax_children(R, pp, &changes);
                        }
                     }
                  } else if (b->nei_level[f] == level - 1) {
                     amr_block *nb = &R->blocks[b->nei[f][0][0]];
                     if (nb->refine != 1) {
                        nb->refine = 1;
                        changes++;
                     }
                  }
               }
            } else if (b->refine == -1) {
               /* a finer neighbour forbids coarsening */
               for (int f = 0; f < 6; f++)
                  if (b->nei_level[f] == level + 1) {
                     b->refine = 0;
                     changes++;
                     if (b->parent != -1) {
                        amr_parent *pp = &R->parents[b->parent];
                        if (pp->refine == -1)
                           pp->refine = 0;
                        relax_children(R, pp, NULL);
                     }
                  }
            }
         }
      } while (changes);

      /* blocks that stay: coarser neighbours may not coarsen */
      do {
         changes = 0;
         for (int in = R->sorted_index[level]; in < R->sorted_index[level + 1]; in++) {
            amr_block *b = &R->blocks[R->sorted_slot[in]];
            if (b->level != level || b->refine != 0)
               continue;
            for (int f = 0; f < 6; f++)
               if (b->nei_level[f] == level - 1) {
                  amr_block *nb = &R->blocks[b->nei[f][0][0]];
                  if (nb->refine == -1) {
                     nb->refine = 0;
                     changes++;
                     if (nb->parent != -1) {
                        amr_parent *pp = &R->parents[nb->parent];
                        if (pp->refine == -1) {
                           pp->refine = 0;
                           relax_children(R, pp, NULL);
                        }
                     }
                  }
               }
         }
      } while (changes);
   }

   int n = 0, na = amr_n_active(R);
   for (int in = 0; in < na; in++)
      if (R->blocks[R->sorted_slot[in]].refine == 1)
         n++;
   return n;
}

/* reset_all + check_objects (or the uniform rule) + refine_level on the
 * device (amr_dev_mark_refine): the same flags, one launch.  Returns 0 when
 * the device declines (mesh too large for its local memory, an undefined
 * object type, or a device library without it) and marking is not
 * required; the caller then marks on the host. */
static int device_mark(amr_run *R)
{
   const amr_params *P = &R->P;
   if (P->device_marking == 0 || (P->device_marking < 0 && R->loop.overlapped))
      return 0;
   const int na = amr_n_active(R), np = R->max_active_parent;
   amr_mark_block *mb = (amr_mark_block *) amr_xmalloc(sizeof(amr_mark_block) * (na ? na : 1));
   amr_mark_parent *mp = (amr_mark_parent *) amr_xmalloc(sizeof(amr_mark_parent) * (np ? np : 1));
   amr_mark_object *mo =
      (amr_mark_object *) amr_xmalloc(sizeof(amr_mark_object) * (P->num_objects ? P->num_objects : 1));
   int32_t *index = (int32_t *) amr_xmalloc(sizeof(int32_t) * P->max_num_blocks);
   int32_t *ls = (int32_t *) amr_xmalloc(sizeof(int32_t) * (P->num_refine + 2));
   int8_t *bf = (int8_t *) amr_xmalloc(na ? na : 1), *pf = (int8_t *) amr_xmalloc(np ? np : 1);
   for (int in = 0; in < na; in++)
      index[R->sorted_slot[in]] = in;
   for (int in = 0; in < na; in++) {
      const amr_block *b = &R->blocks[R->sorted_slot[in]];
      amr_mark_block *k = &mb[in];
      k->level = b->level;
      k->parent = b->parent;
      for (int d = 0; d < 3; d++)
         k->cen[d] = b->cen[d];
      for (int f = 0; f < 6; f++) {
         k->nei_level[f] = b->nei_level[f];
         k->nei[f] = (b->nei_level[f] == b->level || b->nei_level[f] == b->level - 1)
                        ? index[b->nei[f][0][0]] : -1;
      }
   }
   for (int p = 0; p < np; p++) {
      const amr_parent *pp = &R->parents[p];
      mp[p].number = pp->number < 0 ? -1 : 0;
      for (int c = 0; c < 8; c++)
         mp[p].child[c] = (pp->number >= 0 && pp->child[c] >= 0) ? index[pp->child[c]] : -1;
   }
   for (int i = 0; i < P->num_objects; i++) {
      const amr_object *o = &P->objects[i];
      mo[i].type = o->type;
      mo[i].pad = 0;
      for (int d = 0; d < 3; d++) {
         mo[i].cen[d] = o->cen[d];
         mo[i].size[d] = o->size[d];
      }
   }
   for (int l = 0; l <= P->num_refine + 1; l++)
      ls[l] = R->sorted_index[l];
   amr_mark_params mk;
   mk.num_refine = P->num_refine;
   mk.cur_max_level = R->cur_max_level;
   mk.uniform_refine = P->uniform_refine;
   mk.refine_ghost = P->refine_ghost;
   for (int d = 0; d < 3; d++)
      mk.mesh_size[d] = R->mesh_size[d];
   mk.n_objects = P->num_objects;
   int rc = amr_dev_mark_refine(R->dev, &mk, mb, na, ls, mp, np, mo, bf, pf);
   if (rc == AMR_OK) {
      for (int in = 0; in < na; in++)
         R->blocks[R->sorted_slot[in]].refine = bf[in];
      for (int p = 0; p < np; p++)
         if (R->parents[p].number >= 0)
            R->parents[p].refine = pf[p];
   } else if (P->device_marking == 1) {
      amr_fatal("ERROR: device mark_refine failed (%d): %s\n", rc, amr_dev_last_error(R->dev));
   }
   free(mb); free(mp); free(mo); free(index); free(ls); free(bf); free(pf);
   return rc == AMR_OK;
}

/* Split every block marked 1, lowest level first (ref/block.c:41-343). */
void amr_split_blocks(amr_run *R)
{
   const amr_params *P = &R->P;
   int cursor = 0;   /* first-fit cursor for child slots */
   int pcursor = 0;  /* parents below it are in use (none is freed here) */

   for (int level = 0; level <= R->cur_max_level; level++) {
      /* the slot range grows while we go; new children are one level up */
      for (int n = 0; n < R->max_active_block; n++) {
         amr_block *b = &R->blocks[n];
         if (b->number < 0 || b->level != level || b->refine != 1)
            continue;
         int64_t nl = b->number - R->block_start[level];
         int64_t lx = (int64_t) R->p2[level] * R->init_x;
         int64_t ly = (int64_t) R->p2[level] * R->init_y;
         int64_t zp = nl / (lx * ly);
         int64_t yp = (nl % (lx * ly)) / lx;
         int64_t xp = nl % lx;
         if (R->num_active + 8 > P->max_num_blocks)
            amr_fatal("ERROR: Need more blocks %d %d on %d\n", R->num_active,
                      P->max_num_blocks, 0);
         if (R->num_active + 8 > R->local_max_b)
            R->local_max_b = R->num_active + 8;
         R->num_active += 7;
         R->num_blocks[level]--;
         R->num_blocks[level + 1] += 8;

         int p;
         for (p = pcursor; p < R->max_active_parent; p++)
            if (R->parents[p].number < 0)
               break;
         pcursor = p + 1;
         if (p == R->max_num_parents)
            amr_fatal("ERROR: Need more parents\n");
         if (p == R->max_active_parent)
            R->max_active_parent++;
         R->num_parents++;
         R->num_refined++;
         amr_parent *pp = &R->parents[p];
         pp->number = b->number;
         pp->level = b->level;
         pp->parent = b->parent;
         pp->child_number = b->child_number;
         if (pp->parent >= 0)
            R->parents[pp->parent].child[pp->child_number] = -1 - p;
         pp->refine = 0;
         memcpy(pp->cen, b->cen, sizeof(pp->cen));

         int sib[8];
         int half = R->p2[P->num_refine - level - 1];
         int64_t cx = (int64_t) R->p2[level + 1] * R->init_x;
         int64_t cy = (int64_t) R->p2[level + 1] * R->init_y;
         for (int o = 0; o < 8; o++) {
            while (cursor < P->max_num_blocks && R->blocks[cursor].number >= 0)
               cursor++;
            if (cursor == P->max_num_blocks)
               amr_fatal("Error: No inactive blocks available %d %d %d\n", cursor,
                         R->num_active, P->max_num_blocks);
            int m = cursor;
            if (m + 1 > R->max_active_block)
               R->max_active_block = m + 1;
            amr_block *c = &R->blocks[m];
            amr_sorted_note(R, m);
            int ox = o % 2, oy = (o / 2) % 2, oz = o / 4;
            sib[o] = m;
            pp->child[o] = m;
            c->refine = 0;
            c->level = level + 1;
            c->parent = p;
            c->child_number = o;
            c->number = ((2 * zp + oz) * cy + (2 * yp + oy)) * cx + 2 * xp + ox +
                        R->block_start[level + 1];
            c->cen[0] = b->cen[0] + (2 * ox - 1) * half;
            c->cen[1] = b->cen[1] + (2 * oy - 1) * half;
            c->cen[2] = b->cen[2] + (2 * oz - 1) * half;
            amr_split_job job = { n, m, o, 0 };
            amr_vec_push(&R->split_jobs, &job);
         }

         for (int f = 0; f < 6; f++) {
            /* siblings facing each other across the parent's interior */
            for (int i = 0; i < 2; i++)
               for (int j = 0; j < 2; j++) {
                  amr_block *c = &R->blocks[sib[side[f][i][j] + across[f]]];
                  c->nei_level[f] = level + 1;
                  c->nei[f][0][0] = sib[side[f][i][j]];
               }
            int g = opposite(f);
            if (b->nei_level[f] == AMR_BOUNDARY) {
               for (int i = 0; i < 2; i++)
                  for (int j = 0; j < 2; j++) {
                     amr_block *c = &R->blocks[sib[side[f][i][j]]];
                     c->nei_level[f] = AMR_BOUNDARY;
                     c->nei[f][0][0] = 0;
                  }
            } else if (b->nei_level[f] == level - 1) {
               amr_fatal("%d ERROR: internal misconnect block %lld c %d\n", 0,
                         (long long) b->number, f);
            } else if (b->nei_level[f] == level) {
               int other = b->nei[f][0][0];
               R->blocks[other].nei_level[g] = level + 1;
               for (int i = 0; i < 2; i++)
                  for (int j = 0; j < 2; j++) {
                     amr_block *c = &R->blocks[sib[side[f][i][j]]];
                     c->nei_level[f] = level;
                     c->nei[f][0][0] = other;
                     R->blocks[other].nei[g][i][j] = sib[side[f][i][j]];
                  }
            } else if (b->nei_level[f] == level + 1) {
               for (int i = 0; i < 2; i++)
                  for (int j = 0; j < 2; j++) {
                     int other = b->nei[f][i][j];
                     amr_block *c = &R->blocks[sib[side[f][i][j]]];
                     c->nei_level[f] = level + 1;
                     c->nei[f][0][0] = other;
                     R->blocks[other].nei_level[g] = level + 1;
                     R->blocks[other].nei[g][0][0] = sib[side[f][i][j]];
                  }
            } else {
               amr_fatal("%d ERROR: misconnected b %d %lld l %d nei[%d] %d\n", 0, n,
                         (long long) b->number, level, f, b->nei_level[f]);
            }
         }
         b->number = -1;
         if (n < cursor)
            cursor = n;
      }
   }
   if (R->split_jobs.n) {
      amr_sorted_rebuild(R);
      amr_dev_check(R, amr_dev_split(R->dev, (const amr_split_job *) R->split_jobs.v,
                                     R->split_jobs.n), "split");
      R->split_jobs.n = 0;
      R->mesh_dirty = 1;
   }
}

/* Merge the children of every parent marked -1, finest level first
 * (ref/block.c:349-608). */
void amr_consolidate_blocks(amr_run *R)
{
   const amr_params *P = &R->P;
   int cursor = 0;   /* every slot below it is in use: first fit from slot 0 */
   for (int level = R->cur_max_level; level >= 0; level--)
      for (int p = 0; p < R->max_active_parent; p++) {
         amr_parent *pp = &R->parents[p];
         if (pp->number < 0 || pp->level != level || pp->refine != -1)
            continue;
         int n;
         for (n = cursor; n < P->max_num_blocks; n++)
            if (R->blocks[n].number < 0)
               break;
         if (n == P->max_num_blocks)
            amr_fatal("Out of free blocks in consolidate_blocks %d\n", 0);
         cursor = n + 1;
         amr_block *b = &R->blocks[n];
         amr_sorted_note(R, n);
         if (n + 1 > R->max_active_block)
            R->max_active_block = n + 1;
         if (R->num_active + 1 > R->local_max_b)
            R->local_max_b = R->num_active + 1;
         R->num_active -= 7;
         R->num_reformed++;
         R->num_blocks[level]++;
         R->num_blocks[level + 1] -= 8;
         b->number = pp->number;
         pp->number = -1;
         R->num_parents--;
         b->level = pp->level;
         b->parent = pp->parent;
         b->child_number = pp->child_number;
         if (b->level)
            R->parents[b->parent].child[b->child_number] = n;
         b->refine = 0;
         memcpy(b->cen, pp->cen, sizeof(b->cen));
         for (int o = 0; o < 8; o++) {
            amr_block *c = &R->blocks[pp->child[o]];
            c->number = -1;
            if (pp->child[o] < cursor)
               cursor = pp->child[o];
            amr_merge_job job = { n, pp->child[o], o, 0 };
            amr_vec_push(&R->merge_jobs, &job);
         }
         for (int f = 0; f < 6; f++) {
            int g = opposite(f);
            const amr_block *first = &R->blocks[pp->child[side[f][0][0]]];
            if (first->nei_level[f] == AMR_BOUNDARY) {
               b->nei_level[f] = AMR_BOUNDARY;
            } else if (first->nei_level[f] == level) {
               int nb = first->nei[f][0][0];
               b->nei[f][0][0] = nb;
               b->nei_level[f] = level;
               R->blocks[nb].nei[g][0][0] = n;
               R->blocks[nb].nei_level[g] = level;
            } else {
               for (int i = 0; i < 2; i++)
                  for (int j = 0; j < 2; j++) {
                     const amr_block *c = &R->blocks[pp->child[side[f][i][j]]];
                     if (c->nei_level[f] == level + 2)
                        amr_fatal("%d ERROR: %d con %d block %d c %d wrong level %d\n", 0,
                                  p, n, pp->child[side[f][i][j]], f, level);
                     int nb = c->nei[f][0][0];
                     b->nei[f][i][j] = nb;
                     b->nei_level[f] = level + 1;
                     R->blocks[nb].nei[g][0][0] = n;
                     R->blocks[nb].nei_level[g] = level;
                  }
            }
         }
      }
   if (R->merge_jobs.n) {
      amr_sorted_rebuild(R);
      amr_dev_check(R, amr_dev_consolidate(R->dev, (const amr_merge_job *) R->merge_jobs.v,
                                           R->merge_jobs.n), "consolidate");
      R->merge_jobs.n = 0;
      R->mesh_dirty = 1;
   }
}

/* ref/refine.c:38-323 for one rank.  Prints the per-level block counts
 * (report_perf & 8) exactly as the reference does. */
void amr_refine(amr_run *R, int ts)
{
   const amr_params *P = &R->P;
   int steps = ts ? P->block_change : P->num_refine;
   R->nrs++;

   for (int step = 0; step < steps; step++) {
      for (int j = P->num_refine; j >= 0; j--)
         if (R->num_blocks[j]) {
            R->cur_max_level = j;
            break;
         }
      double t0 = amr_now(), t1 = t0;
      if (!device_mark(R)) {
         reset_all(R);
         if (P->uniform_refine) {
            int na = amr_n_active(R);
            for (int in = 0; in < na; in++) {
               amr_block *b = &R->blocks[R->sorted_slot[in]];
               b->refine = (b->level < P->num_refine) ? 1 : 0;
            }
         } else {
            amr_check_objects(R);
         }
         t1 = amr_now();
         amr_refine_level(R);
      }
      double t2 = amr_now();
      amr_split_blocks(R);
      double t3 = amr_now();
      amr_consolidate_blocks(R);
      double t4 = amr_now();
      R->t_ref_compare += t1 - t0;
      R->t_ref_mark += t2 - t1;
      R->t_ref_split += t3 - t2;
      R->t_ref_coarsen += t4 - t3;
   }
   if (R->num_active > R->local_max_b)
      R->local_max_b = R->num_active;
   R->global_active = 0;
   for (int j = 0; j <= P->num_refine; j++) {
      R->global_active += R->num_blocks[j];
      if (P->report_perf & 8)
         printf("Number of blocks at level %d at timestep %d is %lld\n", j, ts,
                (long long) R->num_blocks[j]);
   }
   if (P->report_perf & 8) {
      printf("Total number of blocks at timestep %d is %lld\n", ts,
             (long long) R->global_active);
      /* one rank has no communication partners */
      static const char *dn[5] = { "x", "y", "z", "total", "unique" };
      for (int i = 0; i < 5; i++)
         printf("comm partners %s ave %6.2lf min %d max %d\n%s", dn[i], 0.0, 0, 0,
                i == 4 ? "\n" : "");
   }
   amr_sync_mesh(R);
}
