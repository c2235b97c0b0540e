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

// Fill the ghost plane of face job `jb` for variables v0 .. v0+nv-1
// (nv <= KV) of the pool `pool` (the whole [var][slot][record] pool),
// with threads tid = 0 .. nthr-1 of the item (a workgroup or a wavefront;
// no synchronisation is needed: every thread writes its own cells).
template <typename T, int KV = kVars>
__device__ __forceinline__ void exchange_job(T *__restrict__ pool, const amr_face_job &jb,
                                             const Lay &L, int v0, int nv, int tid, int nthr)
{
   typedef typename Acc<T>::type A;
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

   if (jb.kind == AMR_FACE_SAME || jb.kind == AMR_FACE_BC) {
      int ext_a = 0, ext_b = 0;
      if (L.stencil == 27) {
         if (jb.kind == AMR_FACE_BC) ext_a = ext_b = 1;
         else if (dir == 1) ext_a = 1;
         else if (dir == 2) ext_a = ext_b = 1;
      }
      const int a0 = 1 - ext_a, b0 = 1 - ext_b;
      const int wa = na + 2 * ext_a, wb = nb + 2 * ext_b;
      // BC reads the block's own plane next to the face
      const int sp = jb.kind == AMR_FACE_BC ? (hi ? n : 1) : plane;
      for (int t = tid; t < wa * wb; t += nthr) {
         const int a = a0 + t / wb, b = b0 + t % wb;
         int i, j, k, si, sj, sk;
         canon(dir, sp, a, b, si, sj, sk);
         canon(dir, ghost, a, b, i, j, k);
         const int so = dir == 2 ? zcopy(L, si, sj, sk) : home(L, si, sj, sk);
         const int dh = home(L, i, j, k);
         const int dc = (k == 1 || k == L.nz) ? zcopy(L, i, j, k) : -1;
         T v[KV];
#pragma unroll
         for (int q = 0; q < KV; q++)
            if (q < nv) v[q] = src[q * vs + so];
#pragma unroll
         for (int q = 0; q < KV; q++)
            if (q < nv) {
               dst[q * vs + dh] = v[q];
               if (dc >= 0) dst[q * vs + dc] = v[q];
            }
      }
   } else if (jb.kind == AMR_FACE_PROLONG) {
      // fine cell (a, b) of the whole face <- coarse cell of the quarter
      const int ao = jb.jq * (na / 2), bo = jb.iq * (nb / 2);
      for (int t = tid; t < na * nb; t += nthr) {
         const int a = 1 + t / nb, b = 1 + t % nb;
         const int ca = ((a + 1) >> 1) + ao, cb = ((b + 1) >> 1) + bo;
         int i, j, k, si, sj, sk;
         canon(dir, plane, ca, cb, si, sj, sk);
         canon(dir, ghost, a, b, i, j, k);
         const int so = dir == 2 ? zcopy(L, si, sj, sk) : home(L, si, sj, sk);
         const int dh = home(L, i, j, k);
         const int dc = (k == 1 || k == L.nz) ? zcopy(L, i, j, k) : -1;
         T v[KV];
#pragma unroll
         for (int q = 0; q < KV; q++)
            if (q < nv) v[q] = src[q * vs + so];
#pragma unroll
         for (int q = 0; q < KV; q++)
            if (q < nv) {
               const T o = from_acc<T>(to_acc(v[q]) / A(4));
               dst[q * vs + dh] = o;
               if (dc >= 0) dst[q * vs + dc] = o;
            }
      }
   } else {  // RESTRICT: coarse quarter cell <- sum of 4 fine cells
      const int ha = na / 2, hb = nb / 2;
      const int ao = jb.jq * ha, bo = jb.iq * hb;
      for (int t = tid; t < ha * hb; t += nthr) {
         const int a = 1 + t / hb, b = 1 + t % hb;
         int so[4];
#pragma unroll
         for (int r = 0; r < 4; r++) {
            int si, sj, sk;
            canon(dir, plane, 2 * a - 1 + (r >> 1), 2 * b - 1 + (r & 1), si, sj, sk);
            so[r] = dir == 2 ? zcopy(L, si, sj, sk) : home(L, si, sj, sk);
         }
         int i, j, k;
         canon(dir, ghost, a + ao, b + bo, i, j, k);
         const int dh = home(L, i, j, k);
         const int dc = (k == 1 || k == L.nz) ? zcopy(L, i, j, k) : -1;
         T v[KV][4];
#pragma unroll
         for (int q = 0; q < KV; q++)
            if (q < nv) {
#pragma unroll
               for (int r = 0; r < 4; r++)
                  v[q][r] = src[q * vs + so[r]];
            }
#pragma unroll
         for (int q = 0; q < KV; q++)
            if (q < nv) {
               A s = to_acc(v[q][0]);
               s = s + to_acc(v[q][1]);
               s = s + to_acc(v[q][2]);
               s = s + to_acc(v[q][3]);
               const T o = from_acc<T>(s);
               dst[q * vs + dh] = o;
               if (dc >= 0) dst[q * vs + dc] = o;
            }
      }
   }
}

}  // namespace amrk
