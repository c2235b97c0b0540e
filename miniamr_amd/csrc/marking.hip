// marking.hip -- refinement marking on the device: reset_all, check_objects
// (or the uniform rule) and the 2:1 balance fixpoint refine_level of one
// refinement step (ref/refine.c:66-82, 325-527; ref/move.c:65-623).
//
// One workgroup does the whole step.  The refine flags of blocks and
// parent records live in LDS, so each pass of the fixpoint ends at a
// barrier instead of a host round trip; topology (levels, face neighbours,
// parents) and geometry come from global memory.  The block/object tests
// are geometry.h, the same source the host compiles.
//
// The reference walks the blocks in sorted order and updates flags in
// place; here a pass looks at all blocks of a level at once.  Both reach
// the same flags: every rule only moves a flag up the order -1 -> 0 -> 1
// (coarsen -> stay -> refine), and no move switches off a rule for the
// flags it writes:
//   check_objects: a block is 1 when an object crosses it below the top
//     level; it is a "zero source" when it would refine at the top level,
//     when its parent's grown extent is crossed (refine_ghost) or when it
//     was 0 already; every zero source sets its parent to 0 and relaxes its
//     -1 siblings, whichever of them the reference visits first; level-0
//     blocks left at -1 become 0.
//   refine_level, level L from the top down, first loop: blocks at L marked
//     1 relax same-level neighbours, parents and siblings and raise coarser
//     neighbours to 1; blocks at L marked -1 with a finer neighbour become 0
//     (with parent and siblings).  Flags at L only move -1 -> 0 and flags at
//     L-1 only to 1, so the set of 1s at L is fixed and racing writers
//     store the same value.  Second loop: blocks at L marked 0 relax their
//     coarser neighbours (-1 -> 0) with those neighbours' parents and
//     siblings.
// Whenever a parent becomes 0 its -1 children are relaxed, so "parent 0 =>
// no child at -1" holds from then on, which the second loop relies on.
#include <string.h>

#include "amr_internal.h"
#include "../host/geometry.h"

namespace {

constexpr int kMarkThreads = 1024;
constexpr int kMarkLds = 65536 - 256;   // bytes of flags (blocks + parents) in 64 KiB of LDS

struct MarkIn {
   amr_mark_params p;
   const amr_mark_block *blocks;
   const amr_mark_parent *parents;
   const amr_mark_object *objects;
   const int32_t *level_start;
   int nb, np, nx, ny, nz;
   int8_t *out_blocks, *out_parents;
};

__device__ int crosses(const MarkIn &m, const double cor[3][2])
{
   // ref/move.c:143-623: the first object that crosses decides; objects
   // with a negative size are skipped
   for (int i = 0; i < m.p.n_objects; i++) {
      const amr_mark_object &o = m.objects[i];
      if (o.size[0] < 0.0 || o.size[1] < 0.0 || o.size[2] < 0.0)
         continue;
      if (amr_geo_object(o.type, o.cen, o.size, cor))
         return 1;
   }
   return 0;
}

__device__ void relax_children(const MarkIn &m, int8_t *bf, int p, int &changed)
{
   for (int c = 0; c < 8; c++) {
      const int ch = m.parents[p].child[c];
      if (ch >= 0 && bf[ch] == -1) {
         bf[ch] = 0;
         changed = 1;
      }
   }
}

// a parent becomes 0 and its -1 children follow
__device__ void settle_parent(const MarkIn &m, int8_t *bf, int8_t *pf, int p, int &changed)
{
   if (pf[p] == -1)
      pf[p] = 0;
   relax_children(m, bf, p, changed);
}

__global__ __launch_bounds__(kMarkThreads) void k_mark_refine(MarkIn m)
{
   extern __shared__ int8_t flags[];
   int8_t *bf = flags, *pf = flags + m.nb;
   __shared__ int changed_any;
   const int tid = threadIdx.x, nt = blockDim.x;
   int dummy = 0;

   // reset_all (ref/refine.c:500-527): blocks -1; a parent with a child
   // that is itself a parent is 0 and relaxes its children
   for (int b = tid; b < m.nb; b += nt)
      bf[b] = -1;
   for (int p = tid; p < m.np; p += nt) {
      int8_t r = -1;
      if (m.parents[p].number < 0)
         r = 0;
      else
         for (int c = 0; c < 8; c++)
            if (m.parents[p].child[c] < 0)
               r = 0;
      pf[p] = r;
   }
   __syncthreads();
   for (int p = tid; p < m.np; p += nt)
      if (m.parents[p].number >= 0 && pf[p] == 0)
         relax_children(m, bf, p, dummy);
   __syncthreads();

   if (m.p.uniform_refine) {
      // ref/refine.c:67-74
      for (int b = tid; b < m.nb; b += nt)
         bf[b] = m.blocks[b].level < m.p.num_refine ? 1 : 0;
   } else {
      // check_objects (ref/move.c:65-141)
      const int n[3] = {m.nx, m.ny, m.nz};
      for (int b = tid; b < m.nb; b += nt) {
         const amr_mark_block &k = m.blocks[b];
         const int sz = 1 << (m.p.num_refine - k.level);
         double cor[3][2];
         amr_geo_extent(cor, k.cen, sz, m.p.mesh_size);
         if (m.p.refine_ghost)
            amr_geo_grow(cor, sz, 2.0, n, m.p.mesh_size);
         int8_t s = bf[b];
         if (crosses(m, cor)) {
            s = 1;
         } else if (m.p.refine_ghost && k.level) {
            amr_geo_grow(cor, 2 * sz, 1.0, n, m.p.mesh_size);
            if (crosses(m, cor))
               s = 0;
         }
         if ((k.level == m.p.num_refine && s == 1) || s == 0) {
            s = 0;
            if (k.parent != -1)
               pf[k.parent] = 0;   // marks the group; children relaxed below
         }
         if (!k.level && s == -1)
            s = 0;
         bf[b] = s;
      }
      __syncthreads();
      for (int p = tid; p < m.np; p += nt)
         if (m.parents[p].number >= 0 && pf[p] == 0)
            relax_children(m, bf, p, dummy);
   }
   __syncthreads();

   // refine_level (ref/refine.c:325-496)
   for (int L = m.p.cur_max_level; L >= 0; L--) {
      const int b0 = m.level_start[L], b1 = m.level_start[L + 1];
      for (int loop = 0; loop < 2; loop++) {
         while (true) {
            if (tid == 0)
               changed_any = 0;
            __syncthreads();
            int changed = 0;
            for (int b = b0 + tid; b < b1; b += nt) {
               const amr_mark_block &k = m.blocks[b];
               if (k.level != L)
                  continue;
               const int8_t r = bf[b];
               if (loop == 0 && r == 1) {
                  if (k.parent != -1)
                     settle_parent(m, bf, pf, k.parent, changed);
                  for (int f = 0; f < 6; f++) {
                     const int nb = k.nei[f];
                     if (k.nei_level[f] == L) {
                        if (bf[nb] == -1) {
                           bf[nb] = 0;
                           changed = 1;
                           if (m.blocks[nb].parent != -1)
                              settle_parent(m, bf, pf, m.blocks[nb].parent, changed);
                        }
                     } else if (k.nei_level[f] == L - 1) {
                        if (bf[nb] != 1) {
                           bf[nb] = 1;
                           changed = 1;
                        }
                     }
                  }
               } else if (loop == 0 && r == -1) {
                  for (int f = 0; f < 6; f++)
                     if (k.nei_level[f] == L + 1) {
                        bf[b] = 0;
                        changed = 1;
                        if (k.parent != -1)
                           settle_parent(m, bf, pf, k.parent, changed);
                        break;
                     }
               } else if (loop == 1 && r == 0) {
                  for (int f = 0; f < 6; f++)
                     if (k.nei_level[f] == L - 1) {
                        const int nb = k.nei[f];
                        if (bf[nb] == -1) {
                           bf[nb] = 0;
                           changed = 1;
                           if (m.blocks[nb].parent != -1)
                              settle_parent(m, bf, pf, m.blocks[nb].parent, changed);
                        }
                     }
               }
            }
            if (changed)
               changed_any = 1;
            __syncthreads();
            const int again = changed_any;
            __syncthreads();
            if (!again)
               break;
         }
      }
   }

   for (int b = tid; b < m.nb; b += nt)
      m.out_blocks[b] = bf[b];
   for (int p = tid; p < m.np; p += nt)
      m.out_parents[p] = pf[p];
}

template <typename X>
int upload(amr_dev *d, void **buf, size_t *cap, const X *src, size_t n, size_t n_max)
{
   // sized once for the most the context can hold: freeing a device buffer
   // waits for the whole device, stages still queued included
   const size_t bytes = sizeof(X) * std::max<size_t>(std::max<size_t>(n, n_max), 1);
   if (sizeof(X) * n > *cap) {
      hipFree(*buf);
      *buf = nullptr;
      *cap = 0;
      HIP_TRY(d, hipMalloc(buf, bytes));
      *cap = bytes;
   }
   if (n)
      HIP_TRY(d, hipMemcpyAsync(*buf, src, sizeof(X) * n, hipMemcpyHostToDevice,
                                d->mark_stream));
   return AMR_OK;
}

}  // namespace

extern "C" int amr_dev_mark_refine(amr_dev *d, const amr_mark_params *p,
                                   const amr_mark_block *blocks, int n_blocks,
                                   const int32_t *level_start, const amr_mark_parent *parents,
                                   int n_parents, const amr_mark_object *objects,
                                   int8_t *block_refine, int8_t *parent_refine)
{
   if (!d || !p || n_blocks < 0 || n_parents < 0 || (n_blocks && (!blocks || !block_refine)) ||
       (n_parents && (!parents || !parent_refine)) || !level_start ||
       (p->n_objects && !objects) || p->num_refine < 0 || p->num_refine > 30 ||
       p->cur_max_level < 0 || p->cur_max_level > p->num_refine)
      return amr_set_error(d, AMR_ERR_ARG, "mark_refine: bad arguments");
   if (n_blocks + n_parents > kMarkLds)
      return amr_set_error(d, AMR_ERR_ARG, "mark_refine: %d flags exceed the kernel's %d",
                           n_blocks + n_parents, kMarkLds);
   if (!n_blocks)
      return AMR_OK;
   for (int i = 0; i < p->n_objects; i++)
      if (!amr_geo_known_type(objects[i].type))
         return amr_set_error(d, AMR_ERR_ARG, "mark_refine: undefined object %d",
                              objects[i].type);
   // every index the kernel follows must stay inside the arrays
   for (int l = 0; l <= p->num_refine; l++)
      if (level_start[l] < 0 || level_start[l] > level_start[l + 1] || level_start[l + 1] > n_blocks)
         return amr_set_error(d, AMR_ERR_ARG, "mark_refine: level_start not ordered");
   for (int b = 0; b < n_blocks; b++) {
      const amr_mark_block &k = blocks[b];
      if (k.level < 0 || k.level > p->num_refine || k.parent < -1 || k.parent >= n_parents)
         return amr_set_error(d, AMR_ERR_ARG, "mark_refine: block %d out of range", b);
      for (int f = 0; f < 6; f++)
         if ((k.nei_level[f] == k.level || k.nei_level[f] == k.level - 1) &&
             (k.nei[f] < 0 || k.nei[f] >= n_blocks))
            return amr_set_error(d, AMR_ERR_ARG, "mark_refine: neighbour of block %d", b);
   }
   for (int q = 0; q < n_parents; q++)
      for (int c = 0; c < 8; c++)
         if (parents[q].child[c] >= n_blocks)
            return amr_set_error(d, AMR_ERR_ARG, "mark_refine: child of parent %d", q);

   static_assert(sizeof(amr_mark_block) % 4 == 0, "block record");
   // Marking reads no cell data: it runs on a stream of its own, beside the
   // stages that may still be queued on d->stream (batch members, which
   // share the batch's stream, mark on that stream).
   if (!d->mark_stream) {
      if (d->owns_stream)
         HIP_TRY(d, hipStreamCreateWithFlags(&d->mark_stream, hipStreamNonBlocking));
      else
         d->mark_stream = d->stream;
   }
   int rc;
   const size_t mb = (size_t) d->c.max_blocks;
   if ((rc = upload(d, &d->mk_blocks, &d->mk_blocks_cap, blocks, n_blocks, mb)) ||
       (rc = upload(d, &d->mk_parents, &d->mk_parents_cap, parents, n_parents, mb)) ||
       (rc = upload(d, &d->mk_objects, &d->mk_objects_cap, objects, p->n_objects, 64)) ||
       (rc = upload(d, &d->mk_levels, &d->mk_levels_cap, level_start, p->num_refine + 2, 64)))
      return rc;
   const size_t flag_bytes = (size_t) n_blocks + n_parents;
   int8_t *out = nullptr;
   {
      void *o = d->mk_out;
      size_t cap = d->mk_out_cap;
      if (flag_bytes > cap) {
         hipFree(o);
         o = nullptr;
         d->mk_out = nullptr;
         d->mk_out_cap = 0;
         const size_t want = std::max(flag_bytes, 2 * mb);
         HIP_TRY(d, hipMalloc(&o, want));
         d->mk_out = o;
         d->mk_out_cap = want;
      }
      out = (int8_t *) d->mk_out;
   }
   MarkIn m;
   m.p = *p;
   m.blocks = (const amr_mark_block *) d->mk_blocks;
   m.parents = (const amr_mark_parent *) d->mk_parents;
   m.objects = (const amr_mark_object *) d->mk_objects;
   m.level_start = (const int32_t *) d->mk_levels;
   m.nb = n_blocks;
   m.np = n_parents;
   m.nx = d->g.nx;
   m.ny = d->g.ny;
   m.nz = d->g.nz;
   m.out_blocks = out;
   m.out_parents = out + n_blocks;
   hipLaunchKernelGGL(k_mark_refine, dim3(1), dim3(kMarkThreads), flag_bytes, d->mark_stream, m);
   HIP_TRY(d, hipGetLastError());
   if (n_blocks)
      HIP_TRY(d, hipMemcpyAsync(block_refine, out, n_blocks, hipMemcpyDeviceToHost,
                                d->mark_stream));
   if (n_parents)
      HIP_TRY(d, hipMemcpyAsync(parent_refine, out + n_blocks, n_parents, hipMemcpyDeviceToHost,
                                d->mark_stream));
   HIP_TRY(d, hipStreamSynchronize(d->mark_stream));
   return AMR_OK;
}
