// kern_exchange.h -- device body of one face job of the ghost exchange,
// shared by the single-instance kernel (exchange.hip) and the batch
// persistent kernel (batch.hip).  See exchange.hip for the rules.
#pragma once

#include "amr_internal.h"

namespace amrk {

// canonical coordinates of in-plane point (a, b) on plane p of direction dir
__device__ __forceinline__ void canon(int dir, int p, int a, int b, int &i, int &j, int &k)
{
   if (dir == 0) { i = p; j = a; k = b; }
   else if (dir == 1) { i = a; j = p; k = b; }
   else { i = a; j = b; k = p; }
}

constexpr int kVars = 8;   // default variables per item, loads issued together

// Two adjacent cells as one vector (the prolongation's paired stores).
template <typename T> struct VecOf2 {
   typedef T __attribute__((ext_vector_type(2))) type;
};

// Thread groups a job's variables are split over: the largest power of two
// G <= KV with G * ncell <= nthr and whole wavefronts per group, so that a
// job with few cells (a restriction quarter, a small block's face) still
// occupies every thread; group g takes variables g, g + G, ...  (A group's
// variable, hence its record base, is then the same across a wavefront.)
template <int KV>
__device__ __forceinline__ int var_groups(int ncell, int nthr)
{
   int G = 1;
   while (2 * G <= KV && 2 * G * ncell <= nthr && 2 * G * 64 <= nthr)
      G *= 2;
   return G;
}

// Fill the ghost plane of face job `jb` for variables v0 .. v0+nv-1
// (nv <= KV) of the pool `pool` (the whole [var][slot][record] pool),
// with threads tid = 0 .. nthr-1 of the item (a workgroup or a wavefront;
// no synchronisation is needed: every thread writes its own cells).
// A thread takes up to CP cells per pass and issues the loads of all its
// cells and variables before the first store.
template <typename T, int KV = kVars>
__device__ __forceinline__ void exchange_job(T *__restrict__ pool, const amr_face_job &jb,
                                             const Lay &L, int v0, int nv, int tid, int nthr)
{
   typedef typename Acc<T>::type A;
   constexpr int CP = KV * sizeof(T) <= 64 ? 2 : 1;   // cells per pass (copies)
   const int dir = jb.face >> 1, hi = jb.face & 1;
   const int n = dir == 0 ? L.nx : dir == 1 ? L.ny : L.nz;
   const int na = dir == 0 ? L.ny : L.nx;   // slower in-plane axis
   const int nb = dir == 2 ? L.ny : L.nz;   // faster in-plane axis
   const int ghost = hi ? n + 1 : 0;
   const int plane = hi ? 1 : n;            // neighbour plane next to the shared face
   const gptr<T> dst = global_view(pool) + (int64_t) v0 * L.var_stride + (int64_t) jb.dst * L.rec;
   const gptr<const T> src =
      global_view((const T *) pool) + (int64_t) v0 * L.var_stride + (int64_t) jb.src * L.rec;
   const int64_t vs = L.var_stride;
   const int kind = jb.kind & 3;                 // AMR_FACE_SAME .. AMR_FACE_BC
   const bool whole = jb.kind & AMR_FACE_GHOSTS;  // --code 1/2: ghost ring included

   if (kind == AMR_FACE_PROLONG && !whole && !(nb & 1)) {
      // Prolongation of the face interior, two fine cells per thread: the
      // ghost cells (a, 2m+1) and (a, 2m+2) take the same coarse cell
      // (ao + (a+1)/2, bo + m+1) and are adjacent in the ghost plane (b is
      // its fastest axis, and the plane and every record start on an even
      // element), so a pair is one load and one two-element store per
      // variable: half the memory instructions of a cell per thread.  The
      // values are the generic path's, coarse / 4.
      typedef typename VecOf2<T>::type V2;
      const int hb = nb / 2, npair = na * hb;
      const int ao = jb.jq * (na / 2), bo = jb.iq * hb;
      const int G = var_groups<KV>(npair, nthr), cw = nthr / G;
      const int g = __builtin_amdgcn_readfirstlane(tid / cw), c0 = tid - g * cw;
      for (int t = c0; t < npair; t += cw) {
         const int a = 1 + t / hb, m = t - (a - 1) * hb, b = 2 * m + 1;
         int i, j, k, i1, j1, k1, si, sj, sk;
         canon(dir, plane, ((a + 1) >> 1) + ao, m + 1 + bo, si, sj, sk);
         canon(dir, ghost, a, b, i, j, k);
         canon(dir, ghost, a, b + 1, i1, j1, k1);
         const int so = dir == 2 ? zcopy(L, si, sj, sk) : home(L, si, sj, sk);
         const int dh = home(L, i, j, k);   // (a, b + 1) is dh + 1
         const int dc0 = (k == 1 || k == L.nz) ? zcopy(L, i, j, k) : -1;
         const int dc1 = (k1 == 1 || k1 == L.nz) ? zcopy(L, i1, j1, k1) : -1;
         T v[KV];
#pragma unroll
         for (int q = 0; q < KV; q++)
            if (q < nv && (q & (G - 1)) == g)
               v[q] = ld(uniform_base(src + q * vs), so);
#pragma unroll
         for (int q = 0; q < KV; q++)
            if (q < nv && (q & (G - 1)) == g) {
               const gptr<T> dq = uniform_base(dst + q * vs);
               const T o = from_acc<T>(to_acc(v[q]) / A(4));
               V2 w;
               w[0] = o;
               w[1] = o;
               *(gptr<V2>) ((gptr<char>) dq + (uint32_t) dh * (uint32_t) sizeof(T)) = w;
               if (dc0 >= 0) st(dq, dc0, o);
               if (dc1 >= 0) st(dq, dc1, o);
            }
      }
      return;
   }

   if (kind != AMR_FACE_RESTRICT) {
      // SAME / BC / PROLONG: one source cell per ghost cell
      int ext_a = 0, ext_b = 0;
      if (whole) {
         ext_a = ext_b = 1;
      } else if (L.stencil != 7 && kind != AMR_FACE_PROLONG) {   // 27 and 0 extend
         if (kind == AMR_FACE_BC) ext_a = ext_b = 1;
         else if (dir == 1) ext_a = 1;
         else if (dir == 2) ext_a = ext_b = 1;
      }
      const int a0 = 1 - ext_a, b0 = 1 - ext_b;
      const int wa = na + 2 * ext_a, wb = nb + 2 * ext_b, ncell = wa * wb;
      // BC reads the block's own plane next to the face
      const int sp = kind == AMR_FACE_BC ? (hi ? n : 1) : plane;
      // prolongation over the whole plane: fine (a, b) <- coarse
      // (ao + (a+1)/2, bo + (b+1)/2), which also maps the fine ring rows
      // 0 and n+1 to coarse rows ao and ao + n/2 + 1
      const bool prolong = kind == AMR_FACE_PROLONG;
      // prolongation: fine cell (a, b) of the whole face <- coarse cell of the quarter
      const int ao = jb.jq * (na / 2), bo = jb.iq * (nb / 2);
      const int G = var_groups<KV>(ncell, nthr), cw = nthr / G;
      const int g = __builtin_amdgcn_readfirstlane(tid / cw), c0 = tid - g * cw;
      for (int t0 = c0; t0 < ncell; t0 += CP * cw) {
         int so[CP], dh[CP], dc[CP];
#pragma unroll
         for (int h = 0; h < CP; h++) {
            const int t = min(t0 + h * cw, ncell - 1);
            const int a = a0 + t / wb, b = b0 + t % wb;
            int i, j, k, si, sj, sk;
            if (prolong)
               canon(dir, sp, ((a + 1) >> 1) + ao, ((b + 1) >> 1) + bo, si, sj, sk);
            else
               canon(dir, sp, a, b, si, sj, sk);
            canon(dir, ghost, a, b, i, j, k);
            so[h] = dir == 2 ? zcopy(L, si, sj, sk) : home(L, si, sj, sk);
            dh[h] = home(L, i, j, k);
            dc[h] = (k == 1 || k == L.nz) ? zcopy(L, i, j, k) : -1;
         }
         T v[CP][KV];
#pragma unroll
         for (int q = 0; q < KV; q++) {
            const gptr<const T> sq = uniform_base(src + q * vs);
#pragma unroll
            for (int h = 0; h < CP; h++)
               if (t0 + h * cw < ncell && q < nv && (q & (G - 1)) == g)
                  v[h][q] = ld(sq, so[h]);
         }
#pragma unroll
         for (int q = 0; q < KV; q++) {
            const gptr<T> dq = uniform_base(dst + q * vs);
#pragma unroll
            for (int h = 0; h < CP; h++)
               if (t0 + h * cw < ncell && q < nv && (q & (G - 1)) == g) {
                  const T o = prolong ? from_acc<T>(to_acc(v[h][q]) / A(4)) : v[h][q];
                  st(dq, dh[h], o);
                  if (dc[h] >= 0) st(dq, dc[h], o);
               }
         }
      }
   } else if (!whole && !(nb & 1)) {
      // RESTRICT of the face interior with paired loads: a coarse cell's
      // terms in the reference's order are (fa, fb), (fa, fb+1), (fa+1, fb),
      // (fa+1, fb+1); each row's two are adjacent in the source plane (b is
      // its fastest axis, and fb - 1 = 2(b - bo) - 2 is even, as is every
      // part and record start), so one two-element load gives terms 0 and
      // 1 and another terms 2 and 3, summed in the same order.
      typedef typename VecOf2<T>::type V2;
      constexpr int RV = CP * KV / 4 < 1 ? 1 : CP * KV / 4;
      const int ha = na / 2, hb = nb / 2, ncell = ha * hb;
      const int ao = jb.jq * ha, bo = jb.iq * hb;
      const int G = var_groups<KV>(ncell, nthr), cw = nthr / G;
      const int g = __builtin_amdgcn_readfirstlane(tid / cw), c0 = tid - g * cw;
      for (int t = c0; t < ncell; t += cw) {
         const int a = 1 + ao + t / hb, b = 1 + bo + t % hb;   // coarse ghost cell
         const int fa = 2 * (a - ao) - 1, fb = 2 * (b - bo) - 1;
         int si, sj, sk;
         canon(dir, plane, fa, fb, si, sj, sk);
         const int s0 = dir == 2 ? zcopy(L, si, sj, sk) : home(L, si, sj, sk);
         canon(dir, plane, fa + 1, fb, si, sj, sk);
         const int s1 = dir == 2 ? zcopy(L, si, sj, sk) : home(L, si, sj, sk);
         int i, j, k;
         canon(dir, ghost, a, b, i, j, k);
         const int dh = home(L, i, j, k);
         const int dc = (k == 1 || k == L.nz) ? zcopy(L, i, j, k) : -1;
         for (int q0 = g; q0 < nv; q0 += RV * G) {
            V2 v[RV][2];
#pragma unroll
            for (int m = 0; m < RV; m++)
               if (q0 + m * G < nv) {
                  const gptr<const T> sq = uniform_base(src + (int64_t) (q0 + m * G) * vs);
                  v[m][0] = *(gptr<const V2>) ((gptr<const char>) sq +
                                               (uint32_t) s0 * (uint32_t) sizeof(T));
                  v[m][1] = *(gptr<const V2>) ((gptr<const char>) sq +
                                               (uint32_t) s1 * (uint32_t) sizeof(T));
               }
#pragma unroll
            for (int m = 0; m < RV; m++)
               if (q0 + m * G < nv) {
                  const gptr<T> dq = uniform_base(dst + (int64_t) (q0 + m * G) * vs);
                  A s = to_acc((T) v[m][0][0]);
                  s = s + to_acc((T) v[m][0][1]);
                  s = s + to_acc((T) v[m][1][0]);
                  s = s + to_acc((T) v[m][1][1]);
                  const T o = from_acc<T>(s);
                  st(dq, dh, o);
                  if (dc >= 0) st(dq, dc, o);
               }
         }
      }
   } else {  // RESTRICT: coarse quarter cell <- sum of 4 fine cells
      // RV of the thread's variables per pass (4 loads each)
      constexpr int RV = CP * KV / 4 < 1 ? 1 : CP * KV / 4;
      const int ha = na / 2, hb = nb / 2;
      const int ao = jb.jq * ha, bo = jb.iq * hb;
      // whole planes: the quarter's rows/columns grow by the ring row or
      // column on its side (coarse 0 for the low quarter, n+1 for the high
      // one); a ring row takes the fine ring row, one cell per coarse cell
      const int wa = ha + whole, wb = hb + whole, ncell = wa * wb;
      const int a_lo = whole ? (jb.jq ? ha + 1 : 0) : 1 + ao;
      const int b_lo = whole ? (jb.iq ? hb + 1 : 0) : 1 + bo;
      const int G = var_groups<KV>(ncell, nthr), cw = nthr / G;
      const int g = __builtin_amdgcn_readfirstlane(tid / cw), c0 = tid - g * cw;
      for (int t = c0; t < ncell; t += cw) {
         const int a = a_lo + t / wb, b = b_lo + t % wb;   // coarse ghost cell
         // fine rows (and columns) summed: 2a'-1, 2a' (a' = a - ao) for an
         // interior coarse row, the fine ring row itself for a ring row
         const bool ra = a == 0 || a == na + 1, rb = b == 0 || b == nb + 1;
         const int fa = ra ? a : 2 * (a - ao) - 1, fb = rb ? b : 2 * (b - bo) - 1;
         const int nfb = rb ? 1 : 2, cnt = (ra ? 1 : 2) * nfb;
         int so[4];
#pragma unroll
         for (int r = 0; r < 4; r++) {
            // term r of the sum in the reference's order (row-major over
            // the fine rows, then columns); terms past cnt repeat term 0
            const int rr = r < cnt ? r : 0;
            int si, sj, sk;
            canon(dir, plane, fa + rr / nfb, fb + rr % nfb, si, sj, sk);
            so[r] = dir == 2 ? zcopy(L, si, sj, sk) : home(L, si, sj, sk);
         }
         int i, j, k;
         canon(dir, ghost, a, b, i, j, k);
         const int dh = home(L, i, j, k);
         const int dc = (k == 1 || k == L.nz) ? zcopy(L, i, j, k) : -1;
         // this thread's variables are g, g + G, ...
         for (int q0 = g; q0 < nv; q0 += RV * G) {
            T v[RV][4];
#pragma unroll
            for (int m = 0; m < RV; m++)
               if (q0 + m * G < nv) {
                  const gptr<const T> sq = uniform_base(src + (int64_t) (q0 + m * G) * vs);
#pragma unroll
                  for (int r = 0; r < 4; r++)
                     v[m][r] = ld(sq, so[r]);
               }
#pragma unroll
            for (int m = 0; m < RV; m++)
               if (q0 + m * G < nv) {
                  const gptr<T> dq = uniform_base(dst + (int64_t) (q0 + m * G) * vs);
                  A s = to_acc(v[m][0]);
                  if (cnt > 1) s = s + to_acc(v[m][1]);
                  if (cnt > 2) {
                     s = s + to_acc(v[m][2]);
                     s = s + to_acc(v[m][3]);
                  }
                  const T o = from_acc<T>(s);
                  st(dq, dh, o);
                  if (dc >= 0) st(dq, dc, o);
               }
         }
      }
   }
}

// One face job of a fused stage (FusedJob).  Prolongation and restriction
// are the plain fill.  Same-level and boundary jobs fill their cells as the
// plain fill does, except that an FJ_EXT_ONLY job fills only the cells
// outside the face interior (two rows for y, the ring of the plane for z),
// and a read of a source ghost row that is read in place (alt[] >= 0) goes
// to the cell that row would hold: the x = 0 row of the source is the
// interior x = nx plane of the block alt[0], and so on.  A z job reads the
// four corners of the source's ring (cells of its edge lines along z) where
// the source's edge entries say (edges, kEdgeSrcs per slot).  The source
// block then varies per cell, so reads use 64-bit offsets from the
// variable's base.
template <typename T, int KV = kVars>
__device__ __forceinline__ void exchange_job_fused(T *__restrict__ pool, const FusedJob &fj,
                                                   const EdgeSrc *__restrict__ edges,
                                                   const Lay &L, int v0, int nv, int tid,
                                                   int nthr)
{
   const amr_face_job &jb = fj.jb;
   if (jb.kind == AMR_FACE_PROLONG || jb.kind == AMR_FACE_RESTRICT ||
       (jb.kind & AMR_FACE_GHOSTS)) {
      exchange_job<T, KV>(pool, jb, L, v0, nv, tid, nthr);
      return;
   }
   const int dir = jb.face >> 1, hi = jb.face & 1;
   const int n = dir == 0 ? L.nx : dir == 1 ? L.ny : L.nz;
   const int na = dir == 0 ? L.ny : L.nx;
   const int nb = dir == 2 ? L.ny : L.nz;
   const int ghost = hi ? n + 1 : 0;
   int ext_a = 0, ext_b = 0;
   if (L.stencil != 7) {
      if (jb.kind == AMR_FACE_BC) ext_a = ext_b = 1;
      else if (dir == 1) ext_a = 1;
      else if (dir == 2) ext_a = ext_b = 1;
   }
   const int a0 = 1 - ext_a, b0 = 1 - ext_b;
   const int wa = na + 2 * ext_a, wb = nb + 2 * ext_b;
   const int sp = jb.kind == AMR_FACE_BC ? (hi ? n : 1) : (hi ? 1 : n);
   const int64_t vs = L.var_stride;
   const gptr<T> base = global_view(pool) + (int64_t) v0 * vs;
   const int64_t dst0 = (int64_t) jb.dst * L.rec;
   bool ext_only = fj.mode == FJ_EXT_ONLY;
   if (!ext_only && !(nb & 1) && (ext_a || !ext_b)) {
      // A full job's face-interior cells in pairs along b.  Their sources
      // are cells of the source block's own plane (no ring row, so no
      // alt[] block and no edge entry), adjacent like the ghost cells they
      // fill, and pairs start on even elements: one two-element load and
      // store per variable.  The ring cells then go one by one, as in an
      // ext-only job.
      typedef typename VecOf2<T>::type V2;
      const int hb = nb / 2, npair = na * hb;
      const int Gp = var_groups<KV>(npair, nthr), cwp = nthr / Gp;
      const int gp = __builtin_amdgcn_readfirstlane(tid / cwp), cp = tid - gp * cwp;
      const int64_t src0 = (int64_t) jb.src * L.rec;
      for (int t = cp; t < npair; t += cwp) {
         const int a = 1 + t / hb, m = t - (a - 1) * hb, b = 2 * m + 1;
         int i, j, k, i1, j1, k1, si, sj, sk;
         canon(dir, sp, a, b, si, sj, sk);
         canon(dir, ghost, a, b, i, j, k);
         canon(dir, ghost, a, b + 1, i1, j1, k1);
         const int64_t src = src0 + (dir == 2 ? zcopy(L, si, sj, sk) : home(L, si, sj, sk));
         const int64_t dh = dst0 + home(L, i, j, k);   // (a, b + 1) is dh + 1
         const int64_t dc0 = (k == 1 || k == L.nz) ? dst0 + zcopy(L, i, j, k) : -1;
         const int64_t dc1 = (k1 == 1 || k1 == L.nz) ? dst0 + zcopy(L, i1, j1, k1) : -1;
         V2 v[KV];
#pragma unroll
         for (int q = 0; q < KV; q++)
            if (q < nv && (q & (Gp - 1)) == gp)
               v[q] = *(gptr<const V2>) (base + (q * vs + src));
#pragma unroll
         for (int q = 0; q < KV; q++)
            if (q < nv && (q & (Gp - 1)) == gp) {
               *(gptr<V2>) (base + (q * vs + dh)) = v[q];
               if (dc0 >= 0) base[q * vs + dc0] = v[q][0];
               if (dc1 >= 0) base[q * vs + dc1] = v[q][1];
            }
      }
      if (!ext_a)
         return;   // no ring: an x face, or the 7-point stencil
      ext_only = true;
   }
   // ext-only: rows ai = 0 and wa-1 (all bi), then, with ext_b, columns
   // bi = 0 and wb-1 of the rows between
   const int ncell = ext_only ? 2 * wb + (ext_b ? 2 * (wa - 2) : 0) : wa * wb;
   const int G = var_groups<KV>(ncell, nthr), cw = nthr / G;
   const int g = __builtin_amdgcn_readfirstlane(tid / cw), c0 = tid - g * cw;
   for (int t = c0; t < ncell; t += cw) {
      int ai, bi;
      if (!ext_only) {
         ai = t / wb;
         bi = t - ai * wb;
      } else if (t < 2 * wb) {
         ai = t < wb ? 0 : wa - 1;
         bi = t < wb ? t : t - wb;
      } else {
         const int u = t - 2 * wb;
         bi = u < wa - 2 ? 0 : wb - 1;
         ai = 1 + (u < wa - 2 ? u : u - (wa - 2));
      }
      const int a = a0 + ai, b = b0 + bi;
      int i, j, k, si, sj, sk;
      canon(dir, sp, a, b, si, sj, sk);
      canon(dir, ghost, a, b, i, j, k);
      int slot = jb.src;
      int so = dir == 2 ? zcopy(L, si, sj, sk) : home(L, si, sj, sk);
      int64_t src = -1;
      if (dir == 1) {
         if (si == 0 && fj.alt[0] >= 0) {
            slot = fj.alt[0];
            so = home(L, L.nx, sj, sk);
         } else if (si == L.nx + 1 && fj.alt[1] >= 0) {
            slot = fj.alt[1];
            so = home(L, 1, sj, sk);
         }
      } else if (dir == 2) {
         const bool xin = si >= 1 && si <= L.nx, yin = sj >= 1 && sj <= L.ny;
         if (si == 0 && yin && fj.alt[0] >= 0) {
            slot = fj.alt[0];
            so = zcopy(L, L.nx, sj, sk);
         } else if (si == L.nx + 1 && yin && fj.alt[1] >= 0) {
            slot = fj.alt[1];
            so = zcopy(L, 1, sj, sk);
         } else if (sj == 0 && xin && fj.alt[2] >= 0) {
            slot = fj.alt[2];
            so = zcopy(L, si, L.ny, sk);
         } else if (sj == L.ny + 1 && xin && fj.alt[3] >= 0) {
            slot = fj.alt[3];
            so = zcopy(L, si, 1, sk);
         } else if (!xin && !yin) {
            src = edge_at(edges[(int64_t) kEdgeSrcs * jb.src + (si != 0) * 2 + (sj != 0)], L,
                          sk - 1);
         }
      }
      if (src < 0)
         src = (int64_t) slot * L.rec + so;
      const int64_t dh = dst0 + home(L, i, j, k);
      const int64_t dc = (k == 1 || k == L.nz) ? dst0 + zcopy(L, i, j, k) : -1;
      T v[KV];
#pragma unroll
      for (int q = 0; q < KV; q++)
         if (q < nv && (q & (G - 1)) == g)
            v[q] = base[q * vs + src];
#pragma unroll
      for (int q = 0; q < KV; q++)
         if (q < nv && (q & (G - 1)) == g) {
            base[q * vs + dh] = v[q];
            if (dc >= 0) base[q * vs + dc] = v[q];
         }
   }
}

}  // namespace amrk
