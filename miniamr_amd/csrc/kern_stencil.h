// kern_stencil.h -- device body of the 7/27-point stencil on one block
// record, shared by the single-instance kernel (stencil.hip) and the batch
// persistent kernel (batch.hip).  See stencil.hip for the algorithm.
#pragma once

#include <type_traits>

#include "amr_internal.h"

namespace amrk {

constexpr int kThreads = 256;
constexpr int kCols = 4;       // columns per thread: ny*nz <= 1024
constexpr int kPer = 8;        // plane elements per thread: gy*gz <= 2048
constexpr int kDepth = 4;      // planes of loads in flight per workgroup

// Workgroups are dispatched round-robin over the device's 8 XCDs in
// flattened-id order, and each XCD has its own L2.  xcd_remap(id, total)
// turns the flattened id into a work index such that each XCD runs one
// contiguous share of the work (XCD x gets indices [x*per, (x+1)*per),
// with the remainder spread over the first XCDs): consecutive items --
// neighbouring blocks in sorted order -- then run on the same XCD and the
// face planes one reads from the other's record are found in that L2.
__device__ __forceinline__ int xcd_remap(int id, int total)
{
   const int x = id & 7, p = id >> 3, per = total >> 3, rem = total & 7;
   return x < rem ? x * (per + 1) + p : rem * (per + 1) + (x - rem) * per + p;
}

// Synchronisation and reduction over the NT threads that share one work
// item: the whole workgroup (NT == kThreads) or one wavefront (NT == 64).
// A wavefront's LDS operations complete in issue order, so within a wave
// a compiler barrier between the ring writes and reads is enough.
template <int NT>
__device__ __forceinline__ void item_sync()
{
   if constexpr (NT == 64) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
   } else {
      __syncthreads();
   }
}

__device__ __forceinline__ double wave_reduce(double v)
{
   for (int o = 32; o > 0; o >>= 1)
      v += __shfl_down(v, o, 64);
   return v;   // valid in lane 0
}

__device__ __forceinline__ double block_reduce(double v, double *red)
{
   for (int o = 32; o > 0; o >>= 1)
      v += __shfl_down(v, o, 64);
   const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
   __syncthreads();
   if (lane == 0)
      red[wave] = v;
   __syncthreads();
   double s = 0.0;
   if (threadIdx.x == 0)
      for (int w = 0; w < (int) (blockDim.x >> 6); w++)
         s += red[w];
   return s;   // valid in thread 0
}

template <typename A>
__device__ __forceinline__ A sum9(const A *pl, int c, int gz)
{
   A a = pl[c - gz - 1] + pl[c - gz];
   a = a + pl[c - gz + 1];
   a = a + pl[c - 1];
   a = a + pl[c];
   a = a + pl[c + 1];
   a = a + pl[c + gz - 1];
   a = a + pl[c + gz];
   a = a + pl[c + gz + 1];
   return a;
}

// Division of the 27- and 7-point sums.  FP64: IEEE division, so the path
// is bit-identical to the reference.  FP32/BF16 forms: division by a known
// constant with one FMA correction of the reciprocal product (Markstein's
// method for a fixed divisor): q = x*r, e = fma(-q, d, x), q' = fma(e, r, q)
// with r = RN(1/d).  Unlike a bare x*r it carries no systematic bias (RN(1/7)
// alone is 4.5e-8 high, which would drift conserved sums by ~1e-5 over 400
// stages); on 2e6 random operands it agreed with IEEE division everywhere.
template <typename A, int DIV>
__device__ __forceinline__ A divide(A w)
{
   if (sizeof(A) == 8)
      return w / A(DIV);
   const float r = 1.0f / (float) DIV;
   const float q = (float) w * r;
   const float e = __fmaf_rn(-q, (float) DIV, (float) w);
   return (A) __fmaf_rn(e, r, q);
}

// A plane element's load: 32-bit offset from a record, or (fused stages)
// 64-bit offset from a variable's base.
template <typename T> __device__ __forceinline__ T load_at(gptr<T> b, int o)
{
   return ld_stream(b, (uint32_t) o);
}
template <typename T> __device__ __forceinline__ T load_at(gptr<T> b, int64_t o)
{
   return ld_stream64(b, o);
}

// One (block, up to VPW variables) item, computed by NT threads: the whole
// workgroup (NT == kThreads) or one wavefront (NT == 64).  rec points at
// the block's record of the first variable; `ring` is 3*VPW planes of LDS
// owned by the item; `red` 4 doubles of LDS (workgroup items only);
// `partials`, when not null, receives the FP64 sum of the stored values of
// variable u at partials[u * part_stride].  Requires ny*nz <= NCOL*NT and
// (ny+2)*(nz+2) <= NPER*NT.  NZ, when not 0, is the block's nz known at
// compile time (the LDS neighbour offsets and the column decomposition
// then become constants); 0 reads it from L.  D planes of loads are kept
// in flight (small blocks can afford all of theirs).
//
// FUSED (a stage that reads same-level faces in place): the block's slot
// bslot, redir[6], the slots its six faces are read from (-1: its own ghost
// plane), and es[kEdgeSrcs], where its ghost edge lines and corners are
// read (EdgeSrc).  The face interior of a redirected face is read from the
// neighbour's interior (its z-plane copy for z faces) in the copy L
// describes, and the outputs go to the copy Lout describes, so no
// neighbour's cells change while they are read.  Plane offsets are then
// 64-bit offsets from the variable's base, the source block varying per
// element.
template <typename T, int ST, int NCOL, int NPER, int VPW, int NT = kThreads, int NZ = 0,
          int D = kDepth, bool FUSED = false>
__device__ __forceinline__ void stencil_block(T *__restrict__ rec, const Lay &L, int nv,
                                              typename Acc<T>::type *ring, double *red,
                                              double *__restrict__ partials, int part_stride,
                                              const Lay *Lout = nullptr, int bslot = 0,
                                              const int32_t *__restrict__ redir = nullptr,
                                              const EdgeSrc *__restrict__ es = nullptr)
{
   // interior cells are read with L and written with Lo: the same interior
   // copy (in place), or the record's other copy
   const Lay &Lo = Lout ? *Lout : L;
   typedef typename std::conditional<FUSED, int64_t, int>::type Off;
   typedef typename Acc<T>::type A;
   const int nz = NZ ? NZ : L.nz, gz = nz + 2;
   const int plane = L.gy * gz, ncol = L.ny * nz, nplanes = L.nx + 2;
   const int64_t vs = L.var_stride;
   const int tid = NT == kThreads ? (int) threadIdx.x : (int) (threadIdx.x & (NT - 1));
   // the item's records, one per variable, as scalar bases
   gptr<T> ub[VPW], lb[VPW];   // output bases; load bases (FUSED: the variables')
#pragma unroll
   for (int u = 0; u < VPW; u++) {
      ub[u] = uniform_base(global_view(rec) + u * vs);
      lb[u] = FUSED ? uniform_base(global_view(rec) - (int64_t) bslot * L.rec + u * vs) : ub[u];
   }
   int rdr[6] = {-1, -1, -1, -1, -1, -1};
   if constexpr (FUSED) {
#pragma unroll
      for (int f = 0; f < 6; f++)
         rdr[f] = __builtin_amdgcn_readfirstlane(redir[f]);
   }

   // This thread's plane elements r = tid + q*NT: the ny*nz interior cells
   // first, in record order (a wavefront's loads then cover whole cache
   // lines), then the ghost ring: rows j = 0 and j = ny+1, columns k = 0
   // and k = nz+1.  pos is the element's place in an LDS plane; the record
   // offsets are those of plane 0, plane nx+1 and interior planes p, where
   // the offset is affine in p.
   Off off0[NPER], offN[NPER], off1[NPER];
   int stp[NPER], pos[NPER];
#pragma unroll
   for (int q = 0; q < NPER; q++) {
      const int r = min(tid + q * NT, plane - 1);
      int jj, kk;
      if (r < ncol) {
         jj = r / nz + 1;
         kk = r - (jj - 1) * nz + 1;
      } else {
         const int g = r - ncol;
         if (g < 2 * gz) {
            jj = g < gz ? 0 : L.ny + 1;
            kk = g < gz ? g : g - gz;
         } else {
            const int h = g - 2 * gz;
            jj = (h < L.ny ? h : h - L.ny) + 1;
            kk = h < L.ny ? 0 : nz + 1;
         }
      }
      pos[q] = jj * gz + kk;
      if constexpr (!FUSED) {
         off0[q] = home(L, 0, jj, kk);
         offN[q] = home(L, L.nx + 1, jj, kk);
         off1[q] = home(L, 1, jj, kk);
         stp[q] = L.nx > 1 ? home(L, 2, jj, kk) - off1[q] : 0;
      } else {
         const int64_t own = (int64_t) bslot * L.rec;
         const bool jin = jj >= 1 && jj <= L.ny, kin = kk >= 1 && kk <= nz;
         if (jin && kin) {
            off0[q] = rdr[0] >= 0 ? (int64_t) rdr[0] * L.rec + home(L, L.nx, jj, kk)
                                  : own + home(L, 0, jj, kk);
            offN[q] = rdr[1] >= 0 ? (int64_t) rdr[1] * L.rec + home(L, 1, jj, kk)
                                  : own + home(L, L.nx + 1, jj, kk);
         } else {
            // an edge line or corner of the x = 0 / x = nx+1 planes (the
            // x = nx+1 entry follows the x = 0 one at +2, corners at +4)
            int e, t;
            if (kin) { e = jj != 0; t = kk - 1; }
            else if (jin) { e = 4 + (kk != 0); t = jj - 1; }
            else { e = 12 + (jj != 0) * 2 + (kk != 0); t = 0; }
            off0[q] = edge_at(es[e], L, t);
            offN[q] = edge_at(es[e + (e < 12 ? 2 : 4)], L, t);
         }
         // interior planes: a y / z face row from the neighbour's interior
         // (z: its z-plane copy), an edge line along x from its source, else
         // the block's own record
         int o1 = home(L, 1, jj, kk), o2 = L.nx > 1 ? home(L, 2, jj, kk) : o1;
         int64_t b = own;
         if (!jin && !kin) {
            const EdgeSrc &x = es[8 + (jj != 0) * 2 + (kk != 0)];
            b = edge_at(x, L, 0);
            o1 = 0;
            o2 = x.stride;
         } else if (kin && jj == 0 && rdr[2] >= 0) {
            b = (int64_t) rdr[2] * L.rec;
            o1 = home(L, 1, L.ny, kk);
            o2 = L.nx > 1 ? home(L, 2, L.ny, kk) : o1;
         } else if (kin && jj == L.ny + 1 && rdr[3] >= 0) {
            b = (int64_t) rdr[3] * L.rec;
            o1 = home(L, 1, 1, kk);
            o2 = L.nx > 1 ? home(L, 2, 1, kk) : o1;
         } else if (jin && kk == 0 && rdr[4] >= 0) {
            b = (int64_t) rdr[4] * L.rec;
            o1 = zcopy(L, 1, jj, nz);
            o2 = L.nx > 1 ? zcopy(L, 2, jj, nz) : o1;
         } else if (jin && kk == nz + 1 && rdr[5] >= 0) {
            b = (int64_t) rdr[5] * L.rec;
            o1 = zcopy(L, 1, jj, 1);
            o2 = L.nx > 1 ? zcopy(L, 2, jj, 1) : o1;
         }
         off1[q] = b + o1;
         stp[q] = o2 - o1;
      }
   }
   // Per output column col = (j-1)*nz + (k-1): its index in an LDS plane
   // (-1: no such column) and the offset of its z-plane copy at x = 1 (-1:
   // k is not 1 or nz).  The interior offset at x = i is (i-1)*ncol + col
   // and the copy's is zc1 + (i-1)*ny.
   int cpl[NCOL], zc1[NCOL];
#pragma unroll
   for (int q = 0; q < NCOL; q++) {
      const int col = tid + q * NT;
      cpl[q] = -1;
      zc1[q] = -1;
      if (col < ncol) {
         const int j = col / nz + 1, k = col - (j - 1) * nz + 1;
         cpl[q] = j * gz + k;
         if (k == 1 || k == nz)
            zc1[q] = zcopy(Lo, 1, j, k);
      }
   }
   T pre[D][VPW][NPER];
#define PREFETCH(slot, p)                                                  \
   _Pragma("unroll") for (int q = 0; q < NPER; q++) {                      \
      if (tid + q * NT < plane) {                                          \
         const Off o = (p) == 0 ? off0[q] : (p) == L.nx + 1 ? offN[q]       \
                                         : off1[q] + ((p) - 1) * stp[q];    \
         _Pragma("unroll") for (int u = 0; u < VPW; u++)                   \
            if (u < nv) pre[slot][u][q] = load_at(lb[u], o);               \
      }                                                                    \
   }

   A s_prev[VPW][NCOL], s_cur[VPW][NCOL], c_prev[VPW][NCOL];
   double bsum[VPW];
#pragma unroll
   for (int u = 0; u < VPW; u++)
      bsum[u] = 0.0;
#pragma unroll
   for (int d = 0; d < D; d++)
      if (d < nplanes) { PREFETCH(d, d); }

   for (int p0 = 0; p0 < nplanes; p0 += D) {
#pragma unroll
      for (int d = 0; d < D; d++) {
         const int p = p0 + d;
         if (p >= nplanes)
            break;
         // park plane p (ring slot p % 3 was last read two planes ago)
#pragma unroll
         for (int u = 0; u < VPW; u++) {
            A *slot = ring + (u * 3 + p % 3) * plane;
#pragma unroll
            for (int q = 0; q < NPER; q++) {
               if (tid + q * NT < plane && u < nv)
                  slot[pos[q]] = to_acc(pre[d][u][q]);
            }
         }
         if (p + D < nplanes) { PREFETCH(d, p + D); }
         item_sync<NT>();
         const bool out_now = p >= 2;
         const int i = p - 1;                  // output plane when out_now
#pragma unroll
         for (int u = 0; u < VPW; u++) {
            if (u >= nv)
               continue;
            const A *slot = ring + (u * 3 + p % 3) * plane;
            const A *pc = ring + (u * 3 + (p + 2) % 3) * plane;   // plane p - 1
#pragma unroll
            for (int q = 0; q < NCOL; q++) {
               const int c = cpl[q];
               if (c < 0)
                  continue;
               A w = 0;
               if (ST == 27) {
                  const A s_new = sum9(slot, c, gz);
                  if (out_now) {
                     w = (s_prev[u][q] + s_cur[u][q]) + s_new;
                     w = divide<A, 27>(w);
                  }
                  s_prev[u][q] = s_cur[u][q];
                  s_cur[u][q] = s_new;
               } else {
                  if (out_now) {
                     const A cc = pc[c];
                     w = c_prev[u][q] + pc[c - gz];
                     w = w + pc[c - 1];
                     w = w + cc;
                     w = w + pc[c + 1];
                     w = w + pc[c + gz];
                     w = w + slot[c];
                     w = divide<A, 7>(w);
                     c_prev[u][q] = cc;
                  } else if (p == 0) {
                     c_prev[u][q] = slot[c];
                  }
               }
               if (out_now) {
                  const T out = from_acc<T>(w);
                  st_stream(ub[u], Lo.oI + (i - 1) * ncol + tid + q * NT, out);
                  if (zc1[q] >= 0)
                     st_stream(ub[u], zc1[q] + (i - 1) * L.ny, out);
                  if (partials)
                     bsum[u] += (double) to_acc(out);
               }
            }
         }
      }
   }
#undef PREFETCH
   if (partials) {
#pragma unroll
      for (int u = 0; u < VPW; u++) {
         double s;
         if constexpr (NT == 64)
            s = wave_reduce(bsum[u]);
         else
            s = block_reduce(bsum[u], red);
         if (tid == 0 && u < nv)
            global_view(partials)[(int64_t) u * part_stride] = s;
      }
   }
   item_sync<NT>();   // the ring (and red[]) are reused by the next item
}

}  // namespace amrk
