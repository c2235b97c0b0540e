// kern_stencil_vec.h -- the 27-point stencil of stencil_block (kern_stencil.h)
// for blocks with 16 x 16 faces, in a form that issues about half the vector
// memory instructions: each thread owns C adjacent z columns of one
// variable (one 16-byte vector of FP64 or FP32 cells, 8 bytes of BF16) and
// loads, computes and stores them together.
//
// On the bandwidth-scale workload the scalar form keeps the per-CU vector
// memory pipeline (texture address and data units) ~80-90% busy: per plane
// and variable it issues 6 loads and 8 stores, four of them z-plane-copy
// stores with 8 active lanes.  Here a plane of V variables takes one vector
// load per thread for the interiors, the 68 ring cells of each variable
// spread one per thread, one vector store per thread and one z-plane-copy
// store per thread whose columns hold k = 1 or k = nz.
//
// The arithmetic is the scalar form's, term for term: each output is
// ((S(i-1) + S(i)) + S(i+1)) / 27 with S the 9-term plane sum taken row by
// row, left to right (ref/stencil.c:102-144), so the cells are bit-identical
// to it.  The per-block checksum partials are summed in another order (per
// thread over its columns and planes, then over the wavefronts of the
// variable in a fixed order); plain and fused stages both use this form, so
// they agree with each other bit for bit.
#pragma once

#include "kern_stencil.h"

namespace amrk {

// Columns per thread.
template <typename T> struct VecCols { static constexpr int c = sizeof(T) == 8 ? 2 : 4; };
// Variables per workgroup of 256 threads for 16 x 16 faces: 256 / (256 / C).
template <typename T> struct VecVars { static constexpr int v = VecCols<T>::c; };

// Planes of loads in flight per workgroup of the vector form.  Measured on
// the 16^3 workload, ms per fused launch for 2 / 4 / 6 / 8 planes: FP64
// 2.853 / 2.844 / 2.894 / 2.894, FP32 1.614 / 1.599 / 1.865 / 1.825, BF16
// 1.065 / 1.192 / 1.440 / 1.191.  More planes cost registers, hence
// workgroups per CU (BF16 at 2 planes: 64 VGPRs, 7 waves per SIMD; at 4:
// 82, 5); the loads in flight per CU stay about the same.
template <typename T> struct VecDepth { static constexpr int d = sizeof(T) == 2 ? 2 : 4; };

template <typename T, int C> struct VecOf {
   typedef T __attribute__((ext_vector_type(C))) type;
};

template <typename T, typename Off>
__device__ __forceinline__ typename VecOf<T, VecCols<T>::c>::type ld_vec(gptr<T> b, Off o)
{
   typedef typename VecOf<T, VecCols<T>::c>::type VT;
   return *(gptr<const VT>) ((gptr<const char>) b + (int64_t) o * (int64_t) sizeof(T));
}

// Vector stores: nontemporal in FP64, as the scalar form's st_stream.
template <typename T>
__device__ __forceinline__ void st_vec(gptr<T> b, uint32_t o,
                                       typename VecOf<T, VecCols<T>::c>::type v)
{
   typedef typename VecOf<T, VecCols<T>::c>::type VT;
   const gptr<VT> p = (gptr<VT>) ((gptr<char>) b + o * (uint32_t) sizeof(T));
   if constexpr (sizeof(T) == 8)
      __builtin_nontemporal_store(v, p);
   else
      *p = v;
}

// One (block, up to V variables) item on the 256 threads of a workgroup.
// Arguments as stencil_block's (27-point, NT = kThreads); requires
// ny = nz = 16 and every interior, x-face and record offset a multiple of C
// elements (16-byte aligned vectors).
template <typename T, int D, bool FUSED>
__device__ __forceinline__ void stencil27_v16(T *__restrict__ rec, const Lay &L, int nv,
                                              typename Acc<T>::type *ring, double *red,
                                              double *__restrict__ partials, int part_stride,
                                              const Lay *Lout, int bslot,
                                              const int32_t *__restrict__ redir,
                                              const EdgeSrc *__restrict__ es)
{
   typedef typename Acc<T>::type A;
   constexpr int C = VecCols<T>::c;
   typedef typename VecOf<T, C>::type VT;
   constexpr int N = 16, G = N + 2, PL = G * G, NCOL = N * N;
   constexpr int TPV = NCOL / C;               // threads per variable (>= 64)
   constexpr int V = kThreads / TPV;           // variables per item
   constexpr int NR = 4 * G - 4;               // ring cells per plane and variable
   constexpr int RS = (NR * V + kThreads - 1) / kThreads;   // ring cells per thread
   typedef typename std::conditional<FUSED, int64_t, int>::type Off;
   static_assert(TPV % 64 == 0, "a wavefront holds one variable");
   const Lay &Lo = Lout ? *Lout : L;
   const int tid = threadIdx.x;
   const int u = tid / TPV, r = tid - u * TPV;  // variable (uniform per wavefront), column group
   const int j = r / (N / C) + 1, k0 = (r - (j - 1) * (N / C)) * C + 1;
   const int nplanes = L.nx + 2;
   const int64_t vs = L.var_stride;
   const bool act = u < nv;
   const int uu = act ? u : 0;
   const gptr<T> rec0 = global_view(rec);
   // output base of this thread's variable; load bases (FUSED: the
   // variables' bases, offsets from there)
   const gptr<T> ub = uniform_base(rec0 + (int64_t) uu * vs);
   const gptr<T> lb = FUSED ? uniform_base(rec0 - (int64_t) bslot * L.rec + (int64_t) uu * vs) : ub;
   const gptr<T> rb = FUSED ? uniform_base(rec0 - (int64_t) bslot * L.rec) : uniform_base(rec0);
   int rdr[6] = {-1, -1, -1, -1, -1, -1};
   if constexpr (FUSED) {
#pragma unroll
      for (int f = 0; f < 6; f++)
         rdr[f] = __builtin_amdgcn_readfirstlane(redir[f]);
   }
   const int64_t own = FUSED ? (int64_t) bslot * L.rec : 0;

   // the thread's interior vector: plane 0, plane nx+1, plane 1 (then +NCOL)
   Off iv0, ivN, iv1;
   if constexpr (FUSED) {
      iv0 = rdr[0] >= 0 ? (int64_t) rdr[0] * L.rec + home(L, L.nx, j, k0) : own + home(L, 0, j, k0);
      ivN = rdr[1] >= 0 ? (int64_t) rdr[1] * L.rec + home(L, 1, j, k0)
                        : own + home(L, L.nx + 1, j, k0);
   } else {
      iv0 = home(L, 0, j, k0);
      ivN = home(L, L.nx + 1, j, k0);
   }
   iv1 = own + home(L, 1, j, k0);

   // ring cells e = tid + q*256 of the item: variable e / NR, ring index g:
   // rows j = 0 and j = 17 (all k), then columns k = 0 and k = 17 (j = 1..16)
   int64_t r0[RS], rN[RS], r1[RS];
   int rstp[RS], rpos[RS];
   bool rval[RS];
#pragma unroll
   for (int q = 0; q < RS; q++) {
      const int e = tid + q * kThreads;
      const int ru = min(e / NR, V - 1), g = e - (e / NR) * NR;
      rval[q] = e < NR * V && ru < nv;
      int jj, kk;
      if (g < 2 * G) {
         jj = g < G ? 0 : G - 1;
         kk = g < G ? g : g - G;
      } else {
         const int h = g - 2 * G;
         jj = (h < N ? h : h - N) + 1;
         kk = h < N ? 0 : G - 1;
      }
      rpos[q] = ru * 3 * PL + jj * G + kk;
      const int64_t vo = (int64_t) ru * vs;
      if constexpr (!FUSED) {
         r0[q] = vo + home(L, 0, jj, kk);
         rN[q] = vo + home(L, L.nx + 1, jj, kk);
         r1[q] = vo + home(L, 1, jj, kk);
         rstp[q] = L.nx > 1 ? home(L, 2, jj, kk) - home(L, 1, jj, kk) : 0;
      } else {
         const bool jin = jj >= 1 && jj <= N, kin = kk >= 1 && kk <= N;
         // x = 0 / x = nx+1 planes: the ring is edge lines and corners
         int e0, t;
         if (kin) { e0 = jj != 0; t = kk - 1; }
         else if (jin) { e0 = 4 + (kk != 0); t = jj - 1; }
         else { e0 = 12 + (jj != 0) * 2 + (kk != 0); t = 0; }
         r0[q] = vo + edge_at(es[e0], L, t);
         rN[q] = vo + edge_at(es[e0 + (e0 < 12 ? 2 : 4)], L, t);
         // interior planes: face rows from the neighbour's interior (z:
         // its z-plane copy) or the block's own record, edge lines along x
         // from their sources
         int o1 = home(L, 1, jj, kk), o2 = L.nx > 1 ? home(L, 2, jj, kk) : o1;
         int64_t b = own;
         if (!jin && !kin) {
            const EdgeSrc &x = es[8 + (jj != 0) * 2 + (kk != 0)];
            b = edge_at(x, L, 0);
            o1 = 0;
            o2 = x.stride;
         } else if (jj == 0 && rdr[2] >= 0) {
            b = (int64_t) rdr[2] * L.rec;
            o1 = home(L, 1, N, kk);
            o2 = L.nx > 1 ? home(L, 2, N, kk) : o1;
         } else if (jj == G - 1 && rdr[3] >= 0) {
            b = (int64_t) rdr[3] * L.rec;
            o1 = home(L, 1, 1, kk);
            o2 = L.nx > 1 ? home(L, 2, 1, kk) : o1;
         } else if (kk == 0 && rdr[4] >= 0) {
            b = (int64_t) rdr[4] * L.rec;
            o1 = zcopy(L, 1, jj, N);
            o2 = L.nx > 1 ? zcopy(L, 2, jj, N) : o1;
         } else if (kk == G - 1 && rdr[5] >= 0) {
            b = (int64_t) rdr[5] * L.rec;
            o1 = zcopy(L, 1, jj, 1);
            o2 = L.nx > 1 ? zcopy(L, 2, jj, 1) : o1;
         }
         r1[q] = vo + b + o1;
         rstp[q] = o2 - o1;
      }
   }
   // z-plane copy written by this thread (its columns hold k = 1 or k = 16)
   const int zk = k0 == 1 ? 1 : k0 + C - 1 == N ? N : 0;
   const int zc1 = zk ? zcopy(Lo, 1, j, zk) : 0;   // at x = 1; x = i adds (i-1)*ny
   const int ipos = j * G + k0;                    // first column's place in a plane
   A *const mine = ring + uu * 3 * PL;

   VT pv[D];
   T pr[D][RS];
#define VPREFETCH(s, p)                                                           \
   {                                                                              \
      if (act)                                                                    \
         pv[s] = ld_vec(lb, (p) == 0 ? iv0 : (p) == L.nx + 1 ? ivN                \
                                     : iv1 + (Off) ((p) - 1) * NCOL);             \
      _Pragma("unroll") for (int q = 0; q < RS; q++)                              \
         if (rval[q])                                                             \
            pr[s][q] = ld_stream64(rb, (p) == 0 ? r0[q] : (p) == L.nx + 1 ? rN[q] \
                                       : r1[q] + (int64_t) ((p) - 1) * rstp[q]);  \
   }

   A s_prev[C], s_cur[C];
   double bsum = 0.0;
#pragma unroll
   for (int d = 0; d < D; d++)
      if (d < nplanes) VPREFETCH(d, d);

   for (int p0 = 0; p0 < nplanes; p0 += D) {
#pragma unroll
      for (int d = 0; d < D; d++) {
         const int p = p0 + d;
         if (p >= nplanes)
            break;
         const int sl = (p % 3) * PL;
         // park plane p (ring slot p % 3 was last read two planes ago)
         if (act) {
#pragma unroll
            for (int m = 0; m < C; m++)
               mine[sl + ipos + m] = to_acc(pv[d][m]);
         }
#pragma unroll
         for (int q = 0; q < RS; q++)
            if (rval[q])
               ring[rpos[q] + sl] = to_acc(pr[d][q]);
         if (p + D < nplanes) VPREFETCH(d, p + D);
         __syncthreads();
         if (!act)
            continue;
         const bool out_now = p >= 2;
         const int i = p - 1;                  // output plane when out_now
         const A *pl = mine + sl + ipos - G - 1;
         A v[3][C + 2];
#pragma unroll
         for (int rr = 0; rr < 3; rr++)
#pragma unroll
            for (int c = 0; c < C + 2; c++)
               v[rr][c] = pl[rr * G + c];
         VT out;
#pragma unroll
         for (int m = 0; m < C; m++) {
            A s = v[0][m] + v[0][m + 1];
            s = s + v[0][m + 2];
            s = s + v[1][m];
            s = s + v[1][m + 1];
            s = s + v[1][m + 2];
            s = s + v[2][m];
            s = s + v[2][m + 1];
            s = s + v[2][m + 2];
            if (out_now) {
               A w = (s_prev[m] + s_cur[m]) + s;
               w = divide<A, 27>(w);
               out[m] = from_acc<T>(w);
            }
            s_prev[m] = s_cur[m];
            s_cur[m] = s;
         }
         if (out_now) {
            st_vec<T>(ub, (uint32_t) (Lo.oI + (i - 1) * NCOL + (j - 1) * N + (k0 - 1)), out);
            if (zk)
               st_stream(ub, (uint32_t) (zc1 + (i - 1) * N), zk == 1 ? out[0] : out[C - 1]);
            if (partials) {
#pragma unroll
               for (int m = 0; m < C; m++)
                  bsum += (double) to_acc(out[m]);
            }
         }
      }
   }
#undef VPREFETCH
   if (partials) {
      const double s = wave_reduce(bsum);
      const int wave = tid >> 6;
      if ((tid & 63) == 0)
         red[wave] = s;
      __syncthreads();
      if (r == 0 && act) {
         double t = 0.0;
         for (int w = u * (TPV / 64); w < (u + 1) * (TPV / 64); w++)
            t += red[w];
         global_view(partials)[(int64_t) u * part_stride] = t;
      }
   }
   __syncthreads();   // the ring (and red[]) are reused by the next item
}

}  // namespace amrk
