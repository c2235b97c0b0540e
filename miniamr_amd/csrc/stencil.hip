// stencil.hip -- 7/27-point averaging stencil, checksum reductions and
// conversions between canonical cell arrays and records.
//
// Stencil: one workgroup per (active block, variable), one thread per
// (j, k) column (up to kCols columns per thread).  The workgroup walks the
// x planes 0..nx+1 once: each canonical plane is gathered from the record
// (interior rows, face planes, edge lines) into registers kDepth planes
// ahead (so several planes of loads are in flight per workgroup), parked in
// a 3-deep LDS ring, and every thread derives from it what its columns
// need:
//   27-point: the plane's 9-term sum S(p) at (j, k), summed left to right
//             over (j-1..j+1) x (k-1..k+1) -- exactly the reference's
//             sb/sm/sf for the outputs at x = p+1, p, p-1
//             (ref/stencil.c:102-144) -- so each output costs one 9-term
//             sum and w(i) = ((S(i-1) + S(i)) + S(i+1)) / 27 keeps the
//             reference's association;
//   7-point:  (((((W + S) + D) + C) + U) + N) + E, then / 7
//             (ref/stencil.c:82-101), W and E from the column's registers.
// The FP64 path is therefore bit-identical to the reference.  New values
// go to the interior in place (only this workgroup reads its interior, and
// each plane is staged before any output of the plane is written) and to
// the z = 1 / z = nz copies.  When asked, the workgroup also reduces the
// values it stored into a per-block FP64 partial sum (the checksum of
// ref/check_sum.c:36-65 fused into the stencil); a second small kernel
// reduces the partials of each variable in a fixed order.
#include <stdlib.h>
#include <string.h>

#include "amr_internal.h"

namespace {

constexpr int kThreads = 256;
constexpr int kCols = 4;       // columns per thread: ny*nz <= 1024
constexpr int kPer = 8;        // plane elements per thread: gy*gz <= 2048
constexpr int kDepth = 4;      // planes of loads in flight per workgroup

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

template <typename T, int ST, int NCOL, int NPER, int VPW>
__global__ __launch_bounds__(kThreads) void k_stencil(T *__restrict__ pool,
                                                       const int32_t *__restrict__ sorted,
                                                       Lay L, int var0, int var_end,
                                                       double *__restrict__ partials,
                                                       int part_stride)
{
   typedef typename Acc<T>::type A;
   constexpr int D = kDepth;           // planes of global loads in flight
   extern __shared__ __align__(16) unsigned char smem[];
   const int plane = L.gy * L.gz, ncol = L.ny * L.nz, nplanes = L.nx + 2;
   A *ring = (A *) smem;               // per variable: 3 staged planes
   __shared__ double red[kThreads / 64];

   const int in = blockIdx.x, v0 = var0 + blockIdx.y * VPW;
   const int nv = min(VPW, var_end - v0);
   T *rec = pool + (int64_t) v0 * L.var_stride + (int64_t) sorted[in] * L.rec;
   const int64_t vs = L.var_stride;
   const int tid = threadIdx.x;

   // Record offsets of this thread's plane elements: plane 0, plane nx+1,
   // and interior planes p, where the offset is affine in p.
   int off0[NPER], offN[NPER], off1[NPER], stp[NPER];
#pragma unroll
   for (int q = 0; q < NPER; q++) {
      const int e = min(tid + q * kThreads, plane - 1);
      const int jj = e / L.gz, kk = e - jj * L.gz;
      off0[q] = home(L, 0, jj, kk);
      offN[q] = home(L, L.nx + 1, jj, kk);
      off1[q] = home(L, 1, jj, kk);
      stp[q] = L.nx > 1 ? home(L, 2, jj, kk) - off1[q] : 0;
   }
   T pre[D][VPW][NPER];
#define PREFETCH(slot, p)                                                  \
   _Pragma("unroll") for (int q = 0; q < NPER; q++) {                      \
      if (tid + q * kThreads < plane) {                                    \
         const int o = (p) == 0 ? off0[q] : (p) == L.nx + 1 ? offN[q]       \
                                         : off1[q] + ((p) - 1) * stp[q];    \
         _Pragma("unroll") for (int u = 0; u < VPW; u++)                   \
            if (u < nv) pre[slot][u][q] = rec[u * vs + o];                 \
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
               const int e = tid + q * kThreads;
               if (e < plane && u < nv)
                  slot[e] = to_acc(pre[d][u][q]);
            }
         }
         if (p + D < nplanes) { PREFETCH(d, p + D); }
         __syncthreads();
#pragma unroll
         for (int u = 0; u < VPW; u++) {
            if (u >= nv)
               continue;
            const A *slot = ring + (u * 3 + p % 3) * plane;
            const A *pc = ring + (u * 3 + (p + 2) % 3) * plane;   // plane p - 1
            T *ru = rec + u * vs;
#pragma unroll
            for (int q = 0; q < NCOL; q++) {
               const int col = tid + q * kThreads;
               if (col >= ncol)
                  continue;
               const int j = col / L.nz + 1, k = col - (j - 1) * L.nz + 1, c = j * L.gz + k;
               A w = 0;
               bool out_now = p >= 2;
               if (ST == 27) {
                  const A s_new = sum9(slot, c, L.gz);
                  if (out_now) {
                     w = (s_prev[u][q] + s_cur[u][q]) + s_new;
                     w = divide<A, 27>(w);
                  }
                  s_prev[u][q] = s_cur[u][q];
                  s_cur[u][q] = s_new;
               } else {
                  if (out_now) {
                     const A cc = pc[c];
                     w = c_prev[u][q] + pc[c - L.gz];
                     w = w + pc[c - 1];
                     w = w + cc;
                     w = w + pc[c + 1];
                     w = w + pc[c + L.gz];
                     w = w + slot[c];
                     w = divide<A, 7>(w);
                     c_prev[u][q] = cc;
                  } else if (p == 0) {
                     c_prev[u][q] = slot[c];
                  }
               }
               if (out_now) {
                  const T out = from_acc<T>(w);
                  const int i = p - 1;
                  ru[((i - 1) * L.ny + (j - 1)) * L.nz + (k - 1)] = out;
                  if (k == 1 || k == L.nz)
                     ru[zcopy(L, i, j, k)] = out;
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
         double s = block_reduce(bsum[u], red);
         if (threadIdx.x == 0 && u < nv)
            partials[(int64_t) (v0 + u) * part_stride + in] = s;
      }
   }
}

// Per-block interior sums of variables [var0, var0 + gridDim.y).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_block_sums(const T *__restrict__ pool,
                                                          const int32_t *__restrict__ sorted,
                                                          Lay L, int var0,
                                                          double *__restrict__ partials,
                                                          int part_stride)
{
   __shared__ double red[kThreads / 64];
   const int in = blockIdx.x, var = var0 + blockIdx.y;
   const T *rec = pool + (int64_t) var * L.var_stride + (int64_t) sorted[in] * L.rec;
   const int ncell = L.nx * L.ny * L.nz;
   double bsum = 0.0;
   for (int t = threadIdx.x; t < ncell; t += kThreads)
      bsum += (double) to_acc(rec[t]);
   double s = block_reduce(bsum, red);
   if (threadIdx.x == 0)
      partials[(int64_t) var * part_stride + in] = s;
}

// Sum of each variable's per-block partials in a fixed order: thread t
// takes blocks t, t + 256, ... in sorted order, then a fixed tree.
__global__ __launch_bounds__(kThreads) void k_reduce_sums(const double *__restrict__ partials,
                                                           int part_stride, int n_active,
                                                           int var0, double *__restrict__ out)
{
   __shared__ double red[kThreads / 64];
   const int var = var0 + blockIdx.x;
   const double *p = partials + (int64_t) var * part_stride;
   double s = 0.0;
   for (int t = threadIdx.x; t < n_active; t += kThreads)
      s += p[t];
   double tot = block_reduce(s, red);
   if (threadIdx.x == 0)
      out[blockIdx.x] = tot;
}

// Canonical [var][gx][gy][gz] FP64 array <-> record (with z-plane copies).
template <typename T>
__global__ void k_convert_in(T *__restrict__ pool, const double *__restrict__ stage, Lay L,
                             int slot, int cells)
{
   const int var = blockIdx.y;
   T *rec = pool + (int64_t) var * L.var_stride + (int64_t) slot * L.rec;
   const double *src = stage + (int64_t) var * cells;
   for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < cells; t += gridDim.x * blockDim.x) {
      const int i = t / (L.gy * L.gz), r = t - i * L.gy * L.gz, j = r / L.gz, k = r - j * L.gz;
      store_cell(rec, L, i, j, k, from_acc<T>((typename Acc<T>::type) src[t]));
   }
}

template <typename T>
__global__ void k_convert_out(const T *__restrict__ pool, double *__restrict__ stage, Lay L,
                              int slot, int cells)
{
   const int var = blockIdx.y;
   const T *rec = pool + (int64_t) var * L.var_stride + (int64_t) slot * L.rec;
   double *dst = stage + (int64_t) var * cells;
   for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < cells; t += gridDim.x * blockDim.x) {
      const int i = t / (L.gy * L.gz), r = t - i * L.gy * L.gz, j = r / L.gz, k = r - j * L.gz;
      dst[t] = (double) to_acc(rec[home(L, i, j, k)]);
   }
}

// Variables per workgroup: two for the narrow storage types (the per-cell
// index work and barriers are shared), one for FP64.
template <typename T> struct Vpw { static constexpr int v = 2; };
template <> struct Vpw<double> { static constexpr int v = 1; };

template <typename T, int ST>
int launch_st(amr_dev *d, int start, int num, double *parts)
{
   const Lay &L = d->g;
   constexpr int V = Vpw<T>::v;
   const int ncol = L.ny * L.nz, plane = L.gy * L.gz;
   dim3 grid(d->n_active, (num + V - 1) / V);
   size_t lds = sizeof(typename Acc<T>::type) * 3 * V * (size_t) plane;
#define GO(NC, NP)                                                                          \
   hipLaunchKernelGGL((k_stencil<T, ST, NC, NP, V>), grid, dim3(kThreads), lds, d->stream,  \
                      (T *) d->pool, d->d_sorted, L, start, start + num, parts,             \
                      d->c.max_blocks)
   if (ncol <= kThreads && plane <= 2 * kThreads)
      GO(1, 2);
   else if (ncol <= 2 * kThreads && plane <= 4 * kThreads)
      GO(2, 4);
   else if (ncol <= kCols * kThreads && plane <= kPer * kThreads)
      GO(kCols, kPer);
   else
      return amr_set_error(d, AMR_ERR_ARG, "block face %dx%d too large for the stencil kernel",
                           L.ny, L.nz);
#undef GO
   HIP_TRY(d, hipGetLastError());
   return AMR_OK;
}

template <typename T>
int launch_stencil(amr_dev *d, int start, int num, bool want_partials)
{
   if (!d->n_active || !num)
      return AMR_OK;
   double *parts = want_partials ? d->d_partials : nullptr;
   return d->g.stencil == 7 ? launch_st<T, 7>(d, start, num, parts)
                            : launch_st<T, 27>(d, start, num, parts);
}

template <typename T>
int launch_block_sums(amr_dev *d, int start, int num)
{
   if (!d->n_active || !num)
      return AMR_OK;
   hipLaunchKernelGGL(k_block_sums<T>, dim3(d->n_active, num), dim3(kThreads), 0, d->stream,
                      (const T *) d->pool, d->d_sorted, d->g, start, d->d_partials,
                      d->c.max_blocks);
   HIP_TRY(d, hipGetLastError());
   return AMR_OK;
}

template <typename T>
int launch_convert(amr_dev *d, int slot, bool in)
{
   int cells = d->g.gx * d->g.gy * d->g.gz;
   dim3 grid((cells + 255) / 256, d->c.num_vars);
   if (in)
      hipLaunchKernelGGL(k_convert_in<T>, grid, dim3(256), 0, d->stream, (T *) d->pool,
                         (const double *) d->d_stage, d->g, slot, cells);
   else
      hipLaunchKernelGGL(k_convert_out<T>, grid, dim3(256), 0, d->stream,
                         (const T *) d->pool, d->d_stage, d->g, slot, cells);
   HIP_TRY(d, hipGetLastError());
   return AMR_OK;
}

}  // namespace

int amr_launch_stencil(amr_dev *d, int start, int num, bool want_partials)
{
   switch (d->c.precision) {
   case AMR_PREC_F64: return launch_stencil<double>(d, start, num, want_partials);
   case AMR_PREC_F32: return launch_stencil<float>(d, start, num, want_partials);
   default: return launch_stencil<uint16_t>(d, start, num, want_partials);
   }
}

int amr_launch_block_sums(amr_dev *d, int start, int num)
{
   switch (d->c.precision) {
   case AMR_PREC_F64: return launch_block_sums<double>(d, start, num);
   case AMR_PREC_F32: return launch_block_sums<float>(d, start, num);
   default: return launch_block_sums<uint16_t>(d, start, num);
   }
}

int amr_launch_reduce_sums(amr_dev *d, int start, int num, double *dst_row)
{
   if (!num)
      return AMR_OK;
   if (!d->n_active) {
      HIP_TRY(d, hipMemsetAsync(dst_row, 0, sizeof(double) * num, d->stream));
      return AMR_OK;
   }
   hipLaunchKernelGGL(k_reduce_sums, dim3(num), dim3(kThreads), 0, d->stream, d->d_partials,
                      d->c.max_blocks, d->n_active, start, dst_row);
   HIP_TRY(d, hipGetLastError());
   return AMR_OK;
}

int amr_launch_convert_in(amr_dev *d, int slot)
{
   switch (d->c.precision) {
   case AMR_PREC_F64: return launch_convert<double>(d, slot, true);
   case AMR_PREC_F32: return launch_convert<float>(d, slot, true);
   default: return launch_convert<uint16_t>(d, slot, true);
   }
}

int amr_launch_convert_out(amr_dev *d, int slot)
{
   switch (d->c.precision) {
   case AMR_PREC_F64: return launch_convert<double>(d, slot, false);
   case AMR_PREC_F32: return launch_convert<float>(d, slot, false);
   default: return launch_convert<uint16_t>(d, slot, false);
   }
}
