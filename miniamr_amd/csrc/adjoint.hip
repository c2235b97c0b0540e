// adjoint.hip -- tape, replay, checkpoints and the adjoint kernels
// (include/amr_adjoint.h).
//
// The forward program, at the level of pool elements, is a sequence of
// assignments "targets := linear function of sources" (each kernel reads
// everything it needs before its writes can matter: exchange phases read
// cells no fill of the phase writes; the stencil reads a block's old
// values and overwrites its interior; split and merge snapshot what they
// read).  Its transpose, applied in reverse order, is per assignment
//     mu := adjoint of the targets;  targets' adjoints := 0;
//     sources' adjoints += (d targets / d sources)^T mu,
// and that is what each adjoint kernel computes, element by element:
//   * a cell stored with store_cell() has two elements (its home and, at
//     z = 1 / z = nz, its copy): mu is the sum of both, and both clear;
//   * sources are the elements the forward read (homes, or for the z
//     exchange the copies), so the copies' adjoints flow back through the
//     writes that filled them.
// Within one exchange phase no element is both read and written, and no
// two split (merge) jobs of one reverse segment touch a common slot; the
// "+=" onto sources are atomic adds (FP64 and FP32 storage: reverse mode
// does not support BF16).
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <map>
#include <set>

#include "amr_adjoint.h"
#include "amr_internal.h"
#include "kern_exchange.h"
#include "kern_stencil.h"

void amr_tape_record(amr_dev *d, int kind, int a1, int a2, int a3, int a4)
{
   if (!d || !d->tape || !d->tape->on || d->replaying)
      return;
   std::array<int32_t, 8> op = {kind, a1, a2, a3, a4, 0, 0, d->tape->cur_plan};
   d->tape->ops.push_back(op);
}

void amr_tape_free(amr_dev *d)
{
   delete d->tape;
   d->tape = nullptr;
}

struct amr_adj {
   amr_dev *d = nullptr;
   void *pool = nullptr;            // adjoint of every pool element
   size_t bytes = 0;
   int32_t *d_sorted = nullptr;     // plan in force for the current op
   amr_face_job *d_jobs = nullptr;
   int jobs_cap = 0, slots_cap = 0;
   int plan = -2, n_active = 0;
   int32_t count[3] = {0, 0, 0};
   bool exclusive[3] = {false, false, false};   // no element read by two fills
   void *d_jobbuf = nullptr;        // split/merge jobs of one reverse segment
   size_t jobbuf_bytes = 0;
   double *d_stage = nullptr;       // one slot, canonical FP64
   char err[512] = {0};
};

namespace {

using namespace amrk;

int aerr(amr_adj *a, int rc, const char *fmt, ...)
{
   if (a) {
      va_list ap;
      va_start(ap, fmt);
      vsnprintf(a->err, sizeof(a->err), fmt, ap);
      va_end(ap);
   }
   return rc;
}

#define AHIP(a, call)                                                                   \
   do {                                                                                 \
      hipError_t e_ = (call);                                                           \
      if (e_ != hipSuccess)                                                             \
         return aerr((a), e_ == hipErrorOutOfMemory ? AMR_ERR_NOMEM : AMR_ERR_HIP,      \
                     "%s: %s", #call, hipGetErrorString(e_));                           \
   } while (0)

// ---------------------------------------------------------------- stencil
// One workgroup per (active block, variable), streaming the x planes like
// the forward kernel.  With t = mu / DIV on the interior and 0 elsewhere
// (mu = the adjoint of an interior cell: home + z-plane copy), the
// transposed stencil at any cell c of the (nx+2)(ny+2)(nz+2) box is the sum
// of t over c's stencil neighbourhood (the 7/27-point pattern is
// symmetric).  For the 27-point stencil that is S(q-1) + S(q) + S(q+1),
// S being the 9-term sums of a t plane taken at every (j, k) of the plane,
// ghost positions included; for the 7-point stencil the centre, its four
// in-plane and two out-of-plane neighbours.  Interior cells take the value
// (the forward overwrote them) and their copies clear; ghost cells add it
// (7-point: face ghosts only, nothing reads edges and corners).
//
// Each thread owns plane positions r = tid + q * 256 (q < NP), the ny*nz
// interior positions first in record order, then the ghost ring (as the
// forward kernel, so a wavefront's loads cover whole cache lines); only the
// first IQ slots can hold interior positions.  t planes sit in a 4-deep LDS
// ring with a zero border, so neighbour reads need no bounds tests; the
// ghost cells' old adjoints are kept from their load.  NZ, when not 0, is
// the block's nz at compile time.  t = mu / DIV uses the same division as
// the forward kernel (IEEE in FP64, a corrected reciprocal product in FP32).
template <typename T, int ST, int NP, int IQ, int NZ>
__global__ __launch_bounds__(256) void k_adj_stencil(T *__restrict__ lam,
                                                      const int32_t *__restrict__ sorted, Lay L,
                                                      int var0)
{
   typedef typename Acc<T>::type A;
   extern __shared__ __align__(16) unsigned char smem[];
   const int nz = NZ ? NZ : L.nz, gz = nz + 2, gy = L.gy, plane = gy * gz;
   const int ncol = L.ny * nz;
   const int bw = gz + 2, bplane = (gy + 2) * bw;     // bordered plane
   A *ring = (A *) smem;                              // 4 bordered planes
   const gptr<T> rec = uniform_base(global_view(lam) + (int64_t) (var0 + blockIdx.y) * L.var_stride +
                                    (int64_t) sorted[blockIdx.x] * L.rec);
   const int tid = threadIdx.x;
   for (int e = tid; e < 4 * bplane; e += 256)
      ring[e] = A(0);

   int off0[NP], offN[NP], off1[NP], stp[NP], zc1[IQ], bpos[NP];
   bool live[NP];
#pragma unroll
   for (int q = 0; q < NP; q++) {
      const int r = min(tid + q * 256, plane - 1);
      int j, k;
      if (r < ncol) {
         j = r / nz + 1;
         k = r - (j - 1) * nz + 1;
      } else {
         const int g = r - ncol;
         if (g < 2 * gz) {
            j = g < gz ? 0 : L.ny + 1;
            k = g < gz ? g : g - gz;
         } else {
            const int h = g - 2 * gz;
            j = (h < L.ny ? h : h - L.ny) + 1;
            k = h < L.ny ? 0 : nz + 1;
         }
      }
      live[q] = tid + q * 256 < plane;
      off0[q] = home(L, 0, j, k);
      offN[q] = home(L, L.nx + 1, j, k);
      off1[q] = home(L, 1, j, k);
      stp[q] = L.nx > 1 ? home(L, 2, j, k) - off1[q] : 0;
      if (q < IQ)
         zc1[q] = r < ncol && (k == 1 || k == nz) ? zcopy(L, 1, j, k) : -1;
      bpos[q] = (j + 1) * bw + (k + 1);
      if (ST == 7 && r >= ncol && (j == 0 || j == gy - 1) && (k == 0 || k == gz - 1))
         live[q] = false;   // 7-point: edge lines of interior planes are not touched
   }
   // interior positions: the first ncol of the enumeration
   auto inner = [&](int q) { return q < IQ && tid + q * 256 < ncol; };
   auto offset = [&](int q, int p) {
      return p == 0 ? off0[q] : p == L.nx + 1 ? offN[q] : off1[q] + (p - 1) * stp[q];
   };
   const int np = L.nx + 2;
   A s_prev[NP], s_cur[NP], old_prev[NP], old_cur[NP];
#pragma unroll
   for (int q = 0; q < NP; q++)
      s_prev[q] = s_cur[q] = old_prev[q] = old_cur[q] = A(0);
   // planes of loads in flight: the adjoint of every cell of the plane and,
   // for interior rows at z = 1 / z = nz, of its copy
   constexpr int D = kDepth;
   T pre[D][NP], zpre[D][IQ];
   auto fetch = [&](int d, int p) {
#pragma unroll
      for (int q = 0; q < NP; q++)
         if (tid + q * 256 < plane) {
            pre[d][q] = ld(rec, offset(q, p));
            if (q < IQ && p >= 1 && p <= L.nx && zc1[q < IQ ? q : 0] >= 0)
               zpre[d][q < IQ ? q : 0] = ld(rec, zc1[q < IQ ? q : 0] + (p - 1) * L.ny);
         }
   };
#pragma unroll
   for (int d = 0; d < D; d++)
      if (d < np)
         fetch(d, d);
   __syncthreads();
   for (int p0 = 0; p0 <= np; p0 += D) {
#pragma unroll
   for (int d = 0; d < D; d++) {
      const int p = p0 + d;
      if (p > np)
         break;
      // park t(p); keep the old adjoints of plane p for its ghost cells
      A *slot = ring + (p & 3) * bplane;
#pragma unroll
      for (int q = 0; q < NP; q++) {
         A old = A(0), t = A(0);
         if (p < np && tid + q * 256 < plane) {
            old = to_acc(pre[d][q]);
            if (p >= 1 && p <= L.nx && inner(q)) {
               A mu = old;
               if (zc1[q < IQ ? q : 0] >= 0)
                  mu = mu + to_acc(zpre[d][q < IQ ? q : 0]);
               t = divide<A, ST>(mu);
            }
         }
         if (tid + q * 256 < plane)
            slot[bpos[q]] = t;
         old_prev[q] = old_cur[q];
         old_cur[q] = old;
      }
      if (p + D < np)
         fetch(d, p + D);
      __syncthreads();
      if (p == 0)
         continue;
      const int qo = p - 1;   // output plane
      const A *tp = ring + (p & 3) * bplane;             // plane p = qo + 1
      const A *tc = ring + ((p + 3) & 3) * bplane;       // plane qo
      const A *tm = ring + ((p + 2) & 3) * bplane;       // plane qo - 1 (zeros at qo = 0)
#pragma unroll
      for (int q = 0; q < NP; q++) {
         const int c = bpos[q];
         A v;
         if (ST == 27) {
            A sn = tp[c - bw - 1] + tp[c - bw];
            sn = sn + tp[c - bw + 1];
            sn = sn + tp[c - 1];
            sn = sn + tp[c];
            sn = sn + tp[c + 1];
            sn = sn + tp[c + bw - 1];
            sn = sn + tp[c + bw];
            sn = sn + tp[c + bw + 1];
            v = (s_prev[q] + s_cur[q]) + sn;
            s_prev[q] = s_cur[q];
            s_cur[q] = sn;
         } else {
            v = tm[c] + tc[c - bw];
            v = v + tc[c - 1];
            v = v + tc[c];
            v = v + tc[c + 1];
            v = v + tc[c + bw];
            v = v + tp[c];
            (void) s_prev;
         }
         if (!live[q])
            continue;
         const bool in_plane = qo >= 1 && qo <= L.nx;
         if (ST == 7 && !in_plane && !inner(q))
            continue;   // 7-point: edges and corners of the ghost planes
         const int o = offset(q, qo);
         if (in_plane && inner(q)) {
            st(rec, o, from_acc<T>(v));
            if (zc1[q < IQ ? q : 0] >= 0)
               st(rec, zc1[q < IQ ? q : 0] + (qo - 1) * L.ny, from_acc<T>(A(0)));
         } else {
            st(rec, o, from_acc<T>(old_prev[q] + v));
         }
      }
   }
   }
}

// --------------------------------------------------------------- exchange
// One workgroup per (face job, variable), the forward fill's cells.
template <typename T>
__device__ __forceinline__ typename Acc<T>::type take(T *rec, int h, int c)
{
   typedef typename Acc<T>::type A;
   A mu = to_acc(rec[h]);
   rec[h] = from_acc<T>(A(0));
   if (c >= 0) {
      mu = mu + to_acc(rec[c]);
      rec[c] = from_acc<T>(A(0));
   }
   return mu;
}

// Accumulate into a source element.  A real mesh never lets two fills of
// one phase read the same element; the host checks that per plan and
// direction (exclusive_reads) and otherwise uses atomic adds, so any plan
// the forward accepts is transposed correctly.
template <typename T, bool ATOMIC>
__device__ __forceinline__ void give(T *rec, int o, typename Acc<T>::type v)
{
   if (ATOMIC)
      atomicAdd(&rec[o], v);
   else
      rec[o] = from_acc<T>(to_acc(rec[o]) + v);
}

// The transpose of one face job's fill for variables v0 .. v0+nv-1
// (nv <= KV), threads tid of nthr: every thread issues all its loads (the
// ghost adjoints and, without atomics, the source adjoints) for the KV
// variables before any store, as the forward fill does.
template <typename T, int KV, bool ATOMIC>
__device__ __forceinline__ void adj_exchange_job(T *__restrict__ lam, const amr_face_job &jb,
                                                 const Lay &L, int v0, int nv, int tid,
                                                 int nthr)
{
   typedef typename Acc<T>::type A;
   const int dir = jb.face >> 1, hi = jb.face & 1;
   const int n = dir == 0 ? L.nx : dir == 1 ? L.ny : L.nz;
   const int na = dir == 0 ? L.ny : L.nx;
   const int nb = dir == 2 ? L.ny : L.nz;
   const int ghost = hi ? n + 1 : 0;
   const int plane = hi ? 1 : n;
   const int64_t vs = L.var_stride;
   T *dst = lam + (int64_t) v0 * vs + (int64_t) jb.dst * L.rec;
   T *src = lam + (int64_t) v0 * vs + (int64_t) jb.src * L.rec;
   // variable q's records as scalar bases (the atomic path keeps addresses)
   auto dq = [&](int q) { return uniform_base(global_view(dst) + q * vs); };
   auto sq = [&](int q) { return uniform_base(global_view(src) + q * vs); };
   auto srcoff = [&](int si, int sj, int sk) {
      return dir == 2 ? zcopy(L, si, sj, sk) : home(L, si, sj, sk);
   };
   auto dstcopy = [&](int i, int j, int k) {
      return (k == 1 || k == L.nz) ? zcopy(L, i, j, k) : -1;
   };
   const T zero = from_acc<T>(A(0));
   const int kind = jb.kind & 3;                 // AMR_FACE_SAME .. AMR_FACE_BC
   const bool whole = jb.kind & AMR_FACE_GHOSTS;  // --code 1/2: ghost ring included
   if (kind == AMR_FACE_SAME || kind == AMR_FACE_BC) {
      int ext_a = 0, ext_b = 0;
      if (whole) {
         ext_a = ext_b = 1;
      } else if (L.stencil == 27) {
         if (kind == AMR_FACE_BC) ext_a = ext_b = 1;
         else if (dir == 1) ext_a = 1;
         else if (dir == 2) ext_a = ext_b = 1;
      }
      const int a0 = 1 - ext_a, b0 = 1 - ext_b;
      const int wa = na + 2 * ext_a, wb = nb + 2 * ext_b;
      const int sp = kind == AMR_FACE_BC ? (hi ? n : 1) : plane;
      for (int t = tid; t < wa * wb; t += nthr) {
         const int a = a0 + t / wb, b = b0 + t % wb;
         int i, j, k, si, sj, sk;
         canon(dir, sp, a, b, si, sj, sk);
         canon(dir, ghost, a, b, i, j, k);
         const int so = srcoff(si, sj, sk), dh = home(L, i, j, k), dc = dstcopy(i, j, k);
         A mu[KV], sv[KV];
#pragma unroll
         for (int q = 0; q < KV; q++)
            if (q < nv) {
               mu[q] = to_acc(ld(dq(q), dh));
               if (dc >= 0) mu[q] = mu[q] + to_acc(ld(dq(q), dc));
               if (!ATOMIC) sv[q] = to_acc(ld(sq(q), so));
            }
#pragma unroll
         for (int q = 0; q < KV; q++)
            if (q < nv) {
               st(dq(q), dh, zero);
               if (dc >= 0) st(dq(q), dc, zero);
               if (ATOMIC) atomicAdd(&src[q * vs + so], mu[q]);
               else st(sq(q), so, from_acc<T>(sv[q] + mu[q]));
            }
      }
   } else if (kind == AMR_FACE_PROLONG && whole) {
      // whole fine plane: coarse cell (ca, cb) of [ao, ao+ha+1] x
      // [bo, bo+hb+1] fed the fine ghosts (a, b) with ao + (a+1)/2 == ca:
      // fine row 0 for ca == ao, na+1 for ca == ao+ha+1, else 2(ca-ao)-1
      // and 2(ca-ao); columns alike.  Neighbouring quarters share the ring
      // rows of the coarse plane, so such plans take the atomic path.
      const int ha = na / 2, hb = nb / 2;
      const int ao = jb.jq * ha, bo = jb.iq * hb;
      const int wb = hb + 2;
      for (int t = tid; t < (ha + 2) * wb; t += nthr) {
         const int ca = ao + t / wb, cb = bo + t % wb;
         const int ra = ca - ao, rb = cb - bo;      // 0 .. h+1
         const int fa0 = ra == 0 ? 0 : ra == ha + 1 ? na + 1 : 2 * ra - 1;
         const int fb0 = rb == 0 ? 0 : rb == hb + 1 ? nb + 1 : 2 * rb - 1;
         const int nfa = (ra == 0 || ra == ha + 1) ? 1 : 2, nfb = (rb == 0 || rb == hb + 1) ? 1 : 2;
         int dh[4], dc[4];
#pragma unroll
         for (int r = 0; r < 4; r++) {
            const int rr = (r >> 1) < nfa && (r & 1) < nfb ? r : 0;
            int i, j, k;
            canon(dir, ghost, fa0 + (rr >> 1), fb0 + (rr & 1), i, j, k);
            dh[r] = (r == rr) ? home(L, i, j, k) : -1;
            dc[r] = (r == rr) ? dstcopy(i, j, k) : -1;
         }
         int si, sj, sk;
         canon(dir, plane, ca, cb, si, sj, sk);
         const int so = srcoff(si, sj, sk);
         A mu[KV], sv[KV];
#pragma unroll
         for (int q = 0; q < KV; q++)
            if (q < nv) {
               A m = 0;
#pragma unroll
               for (int r = 0; r < 4; r++) {
                  if (dh[r] < 0)
                     continue;
                  m = m + to_acc(ld(dq(q), dh[r]));
                  if (dc[r] >= 0) m = m + to_acc(ld(dq(q), dc[r]));
               }
               mu[q] = m / A(4);
               if (!ATOMIC) sv[q] = to_acc(ld(sq(q), so));
            }
#pragma unroll
         for (int q = 0; q < KV; q++)
            if (q < nv) {
#pragma unroll
               for (int r = 0; r < 4; r++) {
                  if (dh[r] < 0)
                     continue;
                  st(dq(q), dh[r], zero);
                  if (dc[r] >= 0) st(dq(q), dc[r], zero);
               }
               if (ATOMIC) atomicAdd(&src[q * vs + so], mu[q]);
               else st(sq(q), so, from_acc<T>(sv[q] + mu[q]));
            }
      }
   } else if (kind == AMR_FACE_RESTRICT && whole) {
      // the quarter plus its ring share: each coarse ghost was the sum of
      // 4 fine cells (interior), 2 fine ring cells (ring edge) or 1 (corner)
      const int ha = na / 2, hb = nb / 2;
      const int ao = jb.jq * ha, bo = jb.iq * hb;
      const int a_lo = jb.jq ? ha + 1 : 0, b_lo = jb.iq ? hb + 1 : 0, wb = hb + 1;
      for (int t = tid; t < (ha + 1) * wb; t += nthr) {
         const int a = a_lo + t / wb, b = b_lo + t % wb;
         const bool ra = a == 0 || a == na + 1, rb = b == 0 || b == nb + 1;
         const int fa = ra ? a : 2 * (a - ao) - 1, fb = rb ? b : 2 * (b - bo) - 1;
         const int nfb = rb ? 1 : 2, cnt = (ra ? 1 : 2) * nfb;
         int i, j, k;
         canon(dir, ghost, a, b, i, j, k);
         const int dh = home(L, i, j, k), dc = dstcopy(i, j, k);
         int so[4];
#pragma unroll
         for (int r = 0; r < 4; r++) {
            const int rr = r < cnt ? r : 0;
            int si, sj, sk;
            canon(dir, plane, fa + rr / nfb, fb + rr % nfb, si, sj, sk);
            so[r] = srcoff(si, sj, sk);
         }
         A mu[KV], sv[KV][4];
#pragma unroll
         for (int q = 0; q < KV; q++)
            if (q < nv) {
               mu[q] = to_acc(ld(dq(q), dh));
               if (dc >= 0) mu[q] = mu[q] + to_acc(ld(dq(q), dc));
               if (!ATOMIC) {
#pragma unroll
                  for (int r = 0; r < 4; r++)
                     if (r < cnt) sv[q][r] = to_acc(ld(sq(q), so[r]));
               }
            }
#pragma unroll
         for (int q = 0; q < KV; q++)
            if (q < nv) {
               st(dq(q), dh, zero);
               if (dc >= 0) st(dq(q), dc, zero);
#pragma unroll
               for (int r = 0; r < 4; r++) {
                  if (r >= cnt)
                     continue;
                  if (ATOMIC) atomicAdd(&src[q * vs + so[r]], mu[q]);
                  else st(sq(q), so[r], from_acc<T>(sv[q][r] + mu[q]));
               }
            }
      }
   } else if (kind == AMR_FACE_PROLONG) {
      // each coarse cell of the quarter fed 4 fine ghosts with value / 4
      const int ha = na / 2, hb = nb / 2;
      const int ao = jb.jq * ha, bo = jb.iq * hb;
      for (int t = tid; t < ha * hb; t += nthr) {
         const int ca = 1 + t / hb, cb = 1 + t % hb;
         int dh[4], dc[4];
#pragma unroll
         for (int r = 0; r < 4; r++) {
            int i, j, k;
            canon(dir, ghost, 2 * ca - 1 + (r >> 1), 2 * cb - 1 + (r & 1), i, j, k);
            dh[r] = home(L, i, j, k);
            dc[r] = dstcopy(i, j, k);
         }
         int si, sj, sk;
         canon(dir, plane, ca + ao, cb + bo, si, sj, sk);
         const int so = srcoff(si, sj, sk);
         A mu[KV], sv[KV];
#pragma unroll
         for (int q = 0; q < KV; q++)
            if (q < nv) {
               A m = 0;
#pragma unroll
               for (int r = 0; r < 4; r++) {
                  m = m + to_acc(ld(dq(q), dh[r]));
                  if (dc[r] >= 0) m = m + to_acc(ld(dq(q), dc[r]));
               }
               mu[q] = m / A(4);
               if (!ATOMIC) sv[q] = to_acc(ld(sq(q), so));
            }
#pragma unroll
         for (int q = 0; q < KV; q++)
            if (q < nv) {
#pragma unroll
               for (int r = 0; r < 4; r++) {
                  st(dq(q), dh[r], zero);
                  if (dc[r] >= 0) st(dq(q), dc[r], zero);
               }
               if (ATOMIC) atomicAdd(&src[q * vs + so], mu[q]);
               else st(sq(q), so, from_acc<T>(sv[q] + mu[q]));
            }
      }
   } else {   // RESTRICT: each coarse ghost was the sum of 4 fine cells
      const int ha = na / 2, hb = nb / 2;
      const int ao = jb.jq * ha, bo = jb.iq * hb;
      for (int t = tid; t < ha * hb; t += nthr) {
         const int a = 1 + t / hb, b = 1 + t % hb;
         int i, j, k;
         canon(dir, ghost, a + ao, b + bo, i, j, k);
         const int dh = home(L, i, j, k), dc = dstcopy(i, j, k);
         int so[4];
#pragma unroll
         for (int r = 0; r < 4; r++) {
            int si, sj, sk;
            canon(dir, plane, 2 * a - 1 + (r >> 1), 2 * b - 1 + (r & 1), si, sj, sk);
            so[r] = srcoff(si, sj, sk);
         }
         A mu[KV], sv[KV][4];
#pragma unroll
         for (int q = 0; q < KV; q++)
            if (q < nv) {
               mu[q] = to_acc(ld(dq(q), dh));
               if (dc >= 0) mu[q] = mu[q] + to_acc(ld(dq(q), dc));
               if (!ATOMIC) {
#pragma unroll
                  for (int r = 0; r < 4; r++)
                     sv[q][r] = to_acc(ld(sq(q), so[r]));
               }
            }
#pragma unroll
         for (int q = 0; q < KV; q++)
            if (q < nv) {
               st(dq(q), dh, zero);
               if (dc >= 0) st(dq(q), dc, zero);
#pragma unroll
               for (int r = 0; r < 4; r++) {
                  if (ATOMIC) atomicAdd(&src[q * vs + so[r]], mu[q]);
                  else st(sq(q), so[r], from_acc<T>(sv[q][r] + mu[q]));
               }
            }
      }
   }
}

// One workgroup per (face job, group of up to 8 variables).
template <typename T, bool ATOMIC>
__global__ __launch_bounds__(256) void k_adj_exchange(T *__restrict__ lam,
                                                       const amr_face_job *__restrict__ jobs,
                                                       Lay L, int var0, int var_end)
{
   const amr_face_job jb = jobs[blockIdx.x];
   const int v0 = var0 + blockIdx.y * 8;
   adj_exchange_job<T, 8, ATOMIC>(lam, jb, L, v0, min(8, var_end - v0), threadIdx.x, 256);
}

// ------------------------------------------------------- split and merge
// split: child cell := parent cell / 8 (interior, store_cell).  Adjoint per
// parent cell of the octant: mu = sum over its 8 child cells (home + copy,
// cleared), parent home += mu / 8.
template <typename T>
__global__ void k_adj_split(T *__restrict__ lam, const amr_split_job *__restrict__ jobs, Lay L)
{
   typedef typename Acc<T>::type A;
   const amr_split_job jb = jobs[blockIdx.x];
   const int var = blockIdx.y;
   const int hx = L.nx / 2, hy = L.ny / 2, hz = L.nz / 2;
   const int i1 = (jb.octant & 1) * hx, j1 = ((jb.octant >> 1) & 1) * hy,
             k1 = (jb.octant >> 2) * hz;
   T *c = lam + (int64_t) var * L.var_stride + (int64_t) jb.child * L.rec;
   T *p = lam + (int64_t) var * L.var_stride + (int64_t) jb.parent * L.rec;
   for (int t = threadIdx.x; t < hx * hy * hz; t += blockDim.x) {
      const int i = t / (hy * hz), r = t - i * hy * hz, j = r / hz, k = r - j * hz;
      A mu = 0;
      for (int q = 0; q < 8; q++) {
         const int ci = 2 * i + 1 + (q & 1), cj = 2 * j + 1 + ((q >> 1) & 1),
                   ck = 2 * k + 1 + (q >> 2);
         mu = mu + take(c, home(L, ci, cj, ck), (ck == 1 || ck == L.nz) ? zcopy(L, ci, cj, ck) : -1);
      }
      give<T, false>(p, home(L, i + 1 + i1, j + 1 + j1, k + 1 + k1), mu / A(8));
   }
}

// merge: merged cell := sum of its 8 child cells (store_cell).  Adjoint:
// mu of the merged cell (cleared) added to each of the 8 child homes.
template <typename T>
__global__ void k_adj_merge(T *__restrict__ lam, const amr_merge_job *__restrict__ jobs, Lay L)
{
   typedef typename Acc<T>::type A;
   const amr_merge_job jb = jobs[blockIdx.x];
   const int var = blockIdx.y;
   const int hx = L.nx / 2, hy = L.ny / 2, hz = L.nz / 2;
   const int i1 = (jb.octant & 1) * hx, j1 = ((jb.octant >> 1) & 1) * hy,
             k1 = (jb.octant >> 2) * hz;
   T *c = lam + (int64_t) var * L.var_stride + (int64_t) jb.child * L.rec;
   T *p = lam + (int64_t) var * L.var_stride + (int64_t) jb.dst * L.rec;
   for (int t = threadIdx.x; t < hx * hy * hz; t += blockDim.x) {
      const int i = t / (hy * hz), r = t - i * hy * hz, j = r / hz, k = r - j * hz;
      const int pi = i + 1 + i1, pj = j + 1 + j1, pk = k + 1 + k1;
      const A mu = take(p, home(L, pi, pj, pk), (pk == 1 || pk == L.nz) ? zcopy(L, pi, pj, pk) : -1);
      for (int q = 0; q < 8; q++)
         give<T, false>(c, home(L, 2 * i + 1 + (q & 1), 2 * j + 1 + ((q >> 1) & 1),
                                2 * k + 1 + (q >> 2)), mu);
   }
}

// Canonical FP64 [var][gx][gy][gz] <-> adjoint record: get sums home and
// copy; set writes the home and clears the copy.
template <typename T>
__global__ void k_adj_get(const T *__restrict__ lam, double *__restrict__ out, Lay L, int slot)
{
   const int var = blockIdx.y, cells = L.gx * L.gy * L.gz;
   const T *rec = lam + (int64_t) var * L.var_stride + (int64_t) slot * L.rec;
   for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < cells; t += gridDim.x * blockDim.x) {
      const int i = t / (L.gy * L.gz), r = t - i * L.gy * L.gz, j = r / L.gz, k = r - j * L.gz;
      double v = (double) to_acc(rec[home(L, i, j, k)]);
      if (k == 1 || k == L.nz)
         v += (double) to_acc(rec[zcopy(L, i, j, k)]);
      out[(int64_t) var * cells + t] = v;
   }
}

template <typename T>
__global__ void k_adj_set(T *__restrict__ lam, const double *__restrict__ in, Lay L, int slot)
{
   typedef typename Acc<T>::type A;
   const int var = blockIdx.y, cells = L.gx * L.gy * L.gz;
   T *rec = lam + (int64_t) var * L.var_stride + (int64_t) slot * L.rec;
   for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < cells; t += gridDim.x * blockDim.x) {
      const int i = t / (L.gy * L.gz), r = t - i * L.gy * L.gz, j = r / L.gz, k = r - j * L.gz;
      rec[home(L, i, j, k)] = from_acc<T>((A) in[(int64_t) var * cells + t]);
      if (k == 1 || k == L.nz)
         rec[zcopy(L, i, j, k)] = from_acc<T>(A(0));
   }
}

// Zero one slot's record (all parts) for every variable.
template <typename T>
__global__ void k_clear_slot(T *__restrict__ pool, Lay L, int slot)
{
   T *rec = pool + (int64_t) blockIdx.y * L.var_stride + (int64_t) slot * L.rec;
   for (int64_t t = blockIdx.x * blockDim.x + threadIdx.x; t < L.rec; t += gridDim.x * blockDim.x)
      rec[t] = T(0);
}

// ------------------------------------------------------------ host side
template <typename F>
int by_precision(int prec, F f)
{
   switch (prec) {
   case AMR_PREC_F64: return f((double) 0);
   case AMR_PREC_F32: return f((float) 0);
   default: return f((uint16_t) 0);
   }
}

// Adjoint contexts exist for FP64 and FP32 storage only.
template <typename F>
int adj_precision(int prec, F f)
{
   return prec == AMR_PREC_F64 ? f((double) 0) : f((float) 0);
}

// Per direction: do the plan's fills read disjoint source cells?  Every
// fill reads one plane of its source block -- plane 1 or n of the
// direction -- whole (same level, boundary, restriction, with the 27-point
// extended ranges staying in that plane) or one quarter (prolongation), so
// they are disjoint unless two fills read the same plane of the same block
// and one of them reads all of it or both the same quarter.
void exclusive_reads(amr_adj *a, const TapePlan &pl)
{
   int off = 0;
   for (int dir = 0; dir < 3; dir++) {
      std::map<std::pair<int, int>, int> seen;   // (slot, plane side) -> quarter mask, 16 = whole
      bool ok = true;
      for (int t = off; t < off + pl.count[dir] && ok; t++) {
         const amr_face_job &jb = pl.jobs[t];
         const int hi = jb.face & 1, kind = jb.kind & 3;
         int slot, side, mask;
         if (kind == AMR_FACE_BC) {
            slot = jb.dst;
            side = hi;        // the block's own plane next to the face
            mask = 16;
         } else {
            slot = jb.src;
            side = !hi;       // the neighbour's plane next to the shared face
            // a whole-plane (--code 1/2) prolongation reads beyond its
            // quarter (the ring rows shared with the neighbouring quarters)
            mask = jb.kind == AMR_FACE_PROLONG ? 1 << (jb.iq * 2 + jb.jq) : 16;
         }
         int &m = seen[{slot, side}];
         if ((m & mask) || (m && mask == 16) || (m & 16))
            ok = false;
         m |= mask;
      }
      a->exclusive[dir] = ok;
      off += pl.count[dir];
   }
}

int load_plan(amr_adj *a, int plan)
{
   if (plan == a->plan)
      return AMR_OK;
   a->plan = plan;
   a->n_active = 0;
   a->count[0] = a->count[1] = a->count[2] = 0;
   if (plan < 0)
      return AMR_OK;
   const TapePlan &pl = a->d->tape->plans[plan];
   const int nj = (int) pl.jobs.size(), na = (int) pl.sorted.size();
   AHIP(a, hipStreamSynchronize(a->d->stream));
   if (nj > a->jobs_cap) {
      hipFree(a->d_jobs);
      a->d_jobs = nullptr;
      a->jobs_cap = 0;
      AHIP(a, hipMalloc(&a->d_jobs, sizeof(amr_face_job) * nj));
      a->jobs_cap = nj;
   }
   if (na > a->slots_cap) {
      hipFree(a->d_sorted);
      a->d_sorted = nullptr;
      a->slots_cap = 0;
      AHIP(a, hipMalloc(&a->d_sorted, sizeof(int32_t) * na));
      a->slots_cap = na;
   }
   if (nj)
      AHIP(a, hipMemcpy(a->d_jobs, pl.jobs.data(), sizeof(amr_face_job) * nj,
                        hipMemcpyHostToDevice));
   if (na)
      AHIP(a, hipMemcpy(a->d_sorted, pl.sorted.data(), sizeof(int32_t) * na,
                        hipMemcpyHostToDevice));
   a->n_active = na;
   for (int i = 0; i < 3; i++)
      a->count[i] = pl.count[i];
   exclusive_reads(a, pl);
   return AMR_OK;
}

int adj_stencil(amr_adj *a, int start, int num)
{
   amr_dev *d = a->d;
   if (!a->n_active || !num)
      return AMR_OK;
   const int plane = d->g.gy * d->g.gz;
   if (plane > 3 * 256)
      return aerr(a, AMR_ERR_ARG, "adjoint stencil: block face %dx%d too large", d->g.ny, d->g.nz);
   return adj_precision(d->c.precision, [&](auto z) -> int {
      typedef decltype(z) T;
      typedef typename Acc<T>::type A;
      const size_t lds = sizeof(A) * 4 * (size_t) (d->g.gy + 2) * (d->g.gz + 2);
      dim3 grid(a->n_active, num);
      const int ncol = d->g.ny * d->g.nz;
      // NP plane positions per thread, IQ of them possibly interior; a
      // compile-time nz for the 16^3 and 8^3 shapes
#define GO(ST, NP, IQ, NZ) \
   hipLaunchKernelGGL((k_adj_stencil<T, ST, NP, IQ, NZ>), grid, dim3(256), lds, d->stream, \
                      (T *) a->pool, a->d_sorted, d->g, start)
#define BY_SHAPE(ST)                                                                   \
   if (d->g.nz == 16 && d->g.ny == 16) GO(ST, 2, 1, 16);                              \
   else if (d->g.nz == 8 && d->g.ny == 8) GO(ST, 1, 1, 8);                            \
   else if (plane <= 256) GO(ST, 1, 1, 0);                                            \
   else if (plane <= 512) { if (ncol <= 256) GO(ST, 2, 1, 0); else GO(ST, 2, 2, 0); } \
   else { if (ncol <= 256) GO(ST, 3, 1, 0); else if (ncol <= 512) GO(ST, 3, 2, 0);     \
          else GO(ST, 3, 3, 0); }
      if (d->g.stencil == 27) { BY_SHAPE(27) } else { BY_SHAPE(7) }
#undef BY_SHAPE
#undef GO
      AHIP(a, hipGetLastError());
      return AMR_OK;
   });
}

int adj_comm(amr_adj *a, int start, int num, int stage, int permute)
{
   static const int perm[6][3] = { {0, 1, 2}, {1, 2, 0}, {2, 0, 1},
                                   {0, 2, 1}, {1, 0, 2}, {2, 1, 0} };
   amr_dev *d = a->d;
   for (int o = 2; o >= 0; o--) {
      const int dir = permute ? perm[stage % 6][o] : o;
      const int off = dir == 0 ? 0 : dir == 1 ? a->count[0] : a->count[0] + a->count[1];
      const int n = a->count[dir];
      if (!n || !num)
         continue;
      int rc = adj_precision(d->c.precision, [&](auto z) -> int {
         typedef decltype(z) T;
         auto k = a->exclusive[dir] ? k_adj_exchange<T, false> : k_adj_exchange<T, true>;
         hipLaunchKernelGGL(k, dim3(n, (num + 7) / 8), dim3(256), 0, d->stream,
                            (T *) a->pool, a->d_jobs + off, d->g, start, start + num);
         AHIP(a, hipGetLastError());
         return AMR_OK;
      });
      if (rc)
         return rc;
   }
   return AMR_OK;
}

int ensure_jobbuf(amr_adj *a, size_t bytes)
{
   if (bytes <= a->jobbuf_bytes)
      return AMR_OK;
   AHIP(a, hipStreamSynchronize(a->d->stream));
   hipFree(a->d_jobbuf);
   a->d_jobbuf = nullptr;
   a->jobbuf_bytes = 0;
   AHIP(a, hipMalloc(&a->d_jobbuf, bytes));
   a->jobbuf_bytes = bytes;
   return AMR_OK;
}

// Reverse a split/merge job list: walk it last job first, cutting it into
// segments in which no slot is written by one job and read or written by
// another (a parent/merged slot reused as a child); each segment is one
// launch, segments in reverse order.
template <typename J, bool SPLIT>
int adj_refine(amr_adj *a, const J *jobs, int n)
{
   amr_dev *d = a->d;
   std::vector<J> seg;
   std::set<int> rd, wr;   // adjoint reads (children / merged dst) and writes
   auto flush = [&]() -> int {
      if (seg.empty())
         return AMR_OK;
      int rc = ensure_jobbuf(a, sizeof(J) * seg.size());
      if (rc)
         return rc;
      AHIP(a, hipStreamSynchronize(d->stream));
      AHIP(a, hipMemcpy(a->d_jobbuf, seg.data(), sizeof(J) * seg.size(), hipMemcpyHostToDevice));
      rc = adj_precision(d->c.precision, [&](auto z) -> int {
         typedef decltype(z) T;
         dim3 grid((unsigned) seg.size(), d->c.num_vars);
         if (SPLIT)
            hipLaunchKernelGGL(k_adj_split<T>, grid, dim3(256), 0, d->stream, (T *) a->pool,
                               (const amr_split_job *) a->d_jobbuf, d->g);
         else
            hipLaunchKernelGGL(k_adj_merge<T>, grid, dim3(256), 0, d->stream, (T *) a->pool,
                               (const amr_merge_job *) a->d_jobbuf, d->g);
         AHIP(a, hipGetLastError());
         return AMR_OK;
      });
      seg.clear();
      rd.clear();
      wr.clear();
      return rc;
   };
   for (int t = n - 1; t >= 0; t--) {
      int r, w;   // slot whose adjoint is taken, slot whose adjoint grows
      if (SPLIT) {
         r = ((const amr_split_job *) (const void *) &jobs[t])->child;
         w = ((const amr_split_job *) (const void *) &jobs[t])->parent;
      } else {
         r = ((const amr_merge_job *) (const void *) &jobs[t])->dst;
         w = ((const amr_merge_job *) (const void *) &jobs[t])->child;
      }
      // same taken slot (a merged block's 8 jobs) is fine: disjoint octants
      if (wr.count(r) || rd.count(w) || (SPLIT && rd.count(r)) || (!SPLIT && wr.count(w))) {
         int rc = flush();
         if (rc)
            return rc;
      }
      seg.push_back(jobs[t]);
      rd.insert(r);
      wr.insert(w);
   }
   return flush();
}

}  // namespace

extern "C" {

int amr_dev_tape(amr_dev *d, int on)
{
   if (!d)
      return AMR_ERR_ARG;
   if (on) {
      if (d->g.stencil == 0)
         return amr_set_error(d, AMR_ERR_ARG, "tape: the adjoint covers the 7/27-point stencils");
      delete d->tape;
      d->tape = new Tape();
      d->tape->on = true;
   } else if (d->tape) {
      d->tape->on = false;
   }
   return AMR_OK;
}

int amr_dev_tape_length(const amr_dev *d, int *n)
{
   if (!d || !n)
      return AMR_ERR_ARG;
   *n = d->tape ? (int) d->tape->ops.size() : 0;
   return AMR_OK;
}

int amr_dev_tape_op(const amr_dev *d, int i, int32_t op[8])
{
   if (!d || !op || !d->tape || i < 0 || i >= (int) d->tape->ops.size())
      return AMR_ERR_ARG;
   memcpy(op, d->tape->ops[i].data(), sizeof(int32_t) * 8);
   return AMR_OK;
}

int amr_dev_tape_plan(const amr_dev *d, int plan, int32_t *n_active, int32_t *sorted,
                      int32_t count[3], amr_face_job *jobs)
{
   if (!d || !d->tape || plan < 0 || plan >= (int) d->tape->plans.size())
      return AMR_ERR_ARG;
   const TapePlan &pl = d->tape->plans[plan];
   if (n_active)
      *n_active = (int32_t) pl.sorted.size();
   if (count)
      memcpy(count, pl.count, sizeof(pl.count));
   if (sorted)
      memcpy(sorted, pl.sorted.data(), sizeof(int32_t) * pl.sorted.size());
   if (jobs)
      memcpy(jobs, pl.jobs.data(), sizeof(amr_face_job) * pl.jobs.size());
   return AMR_OK;
}

int amr_dev_tape_split_jobs(const amr_dev *d, int i, amr_split_job *jobs)
{
   if (!d || !jobs || !d->tape || i < 0 || i >= (int) d->tape->ops.size() ||
       d->tape->ops[i][0] != AMR_OP_SPLIT)
      return AMR_ERR_ARG;
   const auto &op = d->tape->ops[i];
   memcpy(jobs, &d->tape->splits[op[1]], sizeof(amr_split_job) * op[2]);
   return AMR_OK;
}

int amr_dev_tape_merge_jobs(const amr_dev *d, int i, amr_merge_job *jobs)
{
   if (!d || !jobs || !d->tape || i < 0 || i >= (int) d->tape->ops.size() ||
       d->tape->ops[i][0] != AMR_OP_MERGE)
      return AMR_ERR_ARG;
   const auto &op = d->tape->ops[i];
   memcpy(jobs, &d->tape->merges[op[1]], sizeof(amr_merge_job) * op[2]);
   return AMR_OK;
}

int amr_dev_replay(amr_dev *d, int op0, int op1)
{
   if (!d || !d->tape || op0 < 0 || op1 < op0 || op1 > (int) d->tape->ops.size())
      return amr_set_error(d, AMR_ERR_ARG, "replay: bad operation range");
   Tape *tp = d->tape;
   // the plan in force at op0
   int rc = AMR_OK;
   d->replaying = true;
   const int plan0 = op0 < (int) tp->ops.size() ? tp->ops[op0][7] : -1;
   if (plan0 >= 0) {
      const TapePlan &pl = tp->plans[plan0];
      rc = amr_set_mesh_impl(d, pl.sorted.data(), (int) pl.sorted.size(), pl.jobs.data(),
                             pl.count);
   }
   for (int i = op0; i < op1 && !rc; i++) {
      const auto &op = tp->ops[i];
      switch (op[0]) {
      case AMR_OP_SET_MESH: {
         const TapePlan &pl = tp->plans[op[1]];
         rc = amr_set_mesh_impl(d, pl.sorted.data(), (int) pl.sorted.size(), pl.jobs.data(),
                                pl.count);
         break;
      }
      case AMR_OP_SPLIT: rc = amr_split_impl(d, &tp->splits[op[1]], op[2]); break;
      case AMR_OP_MERGE: rc = amr_merge_impl(d, &tp->merges[op[1]], op[2]); break;
      case AMR_OP_COMM: rc = amr_comm_impl(d, op[1], op[2], op[3], op[4]); break;
      case AMR_OP_STENCIL: rc = amr_stencil_impl(d, op[1], 0); break;
      case AMR_OP_STAGE: rc = amr_stage_impl(d, op[1], op[2], op[3], op[4], -1); break;
      default: break;   // uploads: the caller provides the values
      }
   }
   d->replaying = false;
   return rc;
}

int amr_dev_state_bytes(const amr_dev *d, size_t *bytes)
{
   if (!d || !bytes)
      return AMR_ERR_ARG;
   *bytes = (size_t) d->g.var_stride * d->c.num_vars * d->elem_bytes;
   return AMR_OK;
}

int amr_dev_save_state(amr_dev *d, void *dst)
{
   if (d && amr_require_fresh(d, 0, d->c.num_vars, "save_state"))
      return AMR_ERR_STATE;
   size_t bytes = 0;
   if (!d || !dst || amr_dev_state_bytes(d, &bytes))
      return amr_set_error(d, AMR_ERR_ARG, "save_state: bad arguments");
   HIP_TRY(d, hipMemcpyAsync(dst, d->pool, bytes, hipMemcpyDefault, d->stream));
   HIP_TRY(d, hipStreamSynchronize(d->stream));
   return AMR_OK;
}

int amr_dev_load_state(amr_dev *d, const void *src)
{
   size_t bytes = 0;
   if (!d || !src || amr_dev_state_bytes(d, &bytes))
      return amr_set_error(d, AMR_ERR_ARG, "load_state: bad arguments");
   HIP_TRY(d, hipStreamSynchronize(d->stream));
   HIP_TRY(d, hipMemcpyAsync(d->pool, src, bytes, hipMemcpyDefault, d->stream));
   HIP_TRY(d, hipStreamSynchronize(d->stream));
   return AMR_OK;
}

int amr_dev_clear_block(amr_dev *d, int slot)
{
   if (!d || slot < 0 || slot >= d->c.max_blocks)
      return amr_set_error(d, AMR_ERR_ARG, "clear_block: slot %d out of range", slot);
   return by_precision(d->c.precision, [&](auto z) -> int {
      typedef decltype(z) T;
      hipLaunchKernelGGL(k_clear_slot<T>, dim3((unsigned) ((d->g.rec + 255) / 256), d->c.num_vars),
                         dim3(256), 0, d->stream, (T *) d->pool, d->g, slot);
      HIP_TRY(d, hipGetLastError());
      return AMR_OK;
   });
}

int amr_adj_create(amr_dev *d, amr_adj **out)
{
   if (!out)
      return AMR_ERR_ARG;
   amr_adj *a = new amr_adj();
   *out = a;
   if (!d)
      return aerr(a, AMR_ERR_ARG, "no forward context");
   a->d = d;
   if (d->c.precision == AMR_PREC_BF16)
      return aerr(a, AMR_ERR_ARG, "reverse mode supports FP64 and FP32 storage, not BF16");
   amr_dev_state_bytes(d, &a->bytes);
   AHIP(a, hipSetDevice(d->c.device));
   AHIP(a, hipMalloc(&a->pool, a->bytes));
   AHIP(a, hipMemsetAsync(a->pool, 0, a->bytes, d->stream));
   AHIP(a, hipMalloc(&a->d_stage, sizeof(double) * (size_t) d->g.gx * d->g.gy * d->g.gz *
                                       d->c.num_vars));
   AHIP(a, hipStreamSynchronize(d->stream));
   return AMR_OK;
}

int amr_adj_destroy(amr_adj *a)
{
   if (!a)
      return AMR_OK;
   if (a->d)
      hipStreamSynchronize(a->d->stream);
   hipFree(a->pool);
   hipFree(a->d_sorted);
   hipFree(a->d_jobs);
   hipFree(a->d_jobbuf);
   hipFree(a->d_stage);
   delete a;
   return AMR_OK;
}

const char *amr_adj_last_error(const amr_adj *a)
{
   return a ? a->err : "no adjoint context";
}

int amr_adj_zero(amr_adj *a)
{
   if (!a || !a->pool)
      return AMR_ERR_ARG;
   AHIP(a, hipMemsetAsync(a->pool, 0, a->bytes, a->d->stream));
   return AMR_OK;
}

int amr_adj_load(amr_adj *a, const void *src)
{
   if (!a || !a->pool || !src)
      return AMR_ERR_ARG;
   AHIP(a, hipStreamSynchronize(a->d->stream));
   AHIP(a, hipMemcpyAsync(a->pool, src, a->bytes, hipMemcpyDefault, a->d->stream));
   AHIP(a, hipStreamSynchronize(a->d->stream));
   return AMR_OK;
}

int amr_adj_save(amr_adj *a, void *dst)
{
   if (!a || !a->pool || !dst)
      return AMR_ERR_ARG;
   AHIP(a, hipMemcpyAsync(dst, a->pool, a->bytes, hipMemcpyDefault, a->d->stream));
   AHIP(a, hipStreamSynchronize(a->d->stream));
   return AMR_OK;
}

int amr_adj_get_block(amr_adj *a, int slot, double *cells)
{
   if (!a || !a->pool || !cells || slot < 0 || slot >= a->d->c.max_blocks)
      return aerr(a, AMR_ERR_ARG, "get_block: bad arguments");
   amr_dev *d = a->d;
   const int ncell = d->g.gx * d->g.gy * d->g.gz;
   int rc = adj_precision(d->c.precision, [&](auto z) -> int {
      typedef decltype(z) T;
      hipLaunchKernelGGL(k_adj_get<T>, dim3((ncell + 255) / 256, d->c.num_vars), dim3(256), 0,
                         d->stream, (const T *) a->pool, a->d_stage, d->g, slot);
      AHIP(a, hipGetLastError());
      return AMR_OK;
   });
   if (rc)
      return rc;
   AHIP(a, hipMemcpyAsync(cells, a->d_stage, sizeof(double) * ncell * d->c.num_vars,
                          hipMemcpyDeviceToHost, d->stream));
   AHIP(a, hipStreamSynchronize(d->stream));
   return AMR_OK;
}

int amr_adj_set_block(amr_adj *a, int slot, const double *cells)
{
   if (!a || !a->pool || !cells || slot < 0 || slot >= a->d->c.max_blocks)
      return aerr(a, AMR_ERR_ARG, "set_block: bad arguments");
   amr_dev *d = a->d;
   const int ncell = d->g.gx * d->g.gy * d->g.gz;
   AHIP(a, hipMemcpyAsync(a->d_stage, cells, sizeof(double) * ncell * d->c.num_vars,
                          hipMemcpyHostToDevice, d->stream));
   int rc = adj_precision(d->c.precision, [&](auto z) -> int {
      typedef decltype(z) T;
      hipLaunchKernelGGL(k_adj_set<T>, dim3((ncell + 255) / 256, d->c.num_vars), dim3(256), 0,
                         d->stream, (T *) a->pool, (const double *) a->d_stage, d->g, slot);
      AHIP(a, hipGetLastError());
      return AMR_OK;
   });
   if (rc)
      return rc;
   AHIP(a, hipStreamSynchronize(d->stream));
   return AMR_OK;
}

int amr_adj_run(amr_adj *a, int op0, int op1)
{
   if (!a || !a->d || !a->d->tape || op0 < 0 || op1 < op0 ||
       op1 > (int) a->d->tape->ops.size())
      return aerr(a, AMR_ERR_ARG, "adj_run: bad operation range");
   Tape *tp = a->d->tape;
   for (int i = op1 - 1; i >= op0; i--) {
      const auto &op = tp->ops[i];
      int rc = load_plan(a, op[7]);
      if (rc)
         return rc;
      switch (op[0]) {
      case AMR_OP_SPLIT:
         rc = adj_refine<amr_split_job, true>(a, &tp->splits[op[1]], op[2]);
         break;
      case AMR_OP_MERGE:
         rc = adj_refine<amr_merge_job, false>(a, &tp->merges[op[1]], op[2]);
         break;
      case AMR_OP_COMM:
         rc = adj_comm(a, op[1], op[2], op[3], op[4]);
         break;
      case AMR_OP_STENCIL:
         rc = adj_stencil(a, op[1], 1);
         break;
      case AMR_OP_STAGE:
         rc = adj_stencil(a, op[1], op[2]);
         if (!rc)
            rc = adj_comm(a, op[1], op[2], op[3], op[4]);
         break;
      default:
         break;   // uploads and mesh changes move no values
      }
      if (rc)
         return rc;
   }
   return AMR_OK;
}

int amr_adj_sync(amr_adj *a)
{
   if (!a || !a->d)
      return AMR_ERR_ARG;
   AHIP(a, hipStreamSynchronize(a->d->stream));
   return AMR_OK;
}

}  // extern "C"
