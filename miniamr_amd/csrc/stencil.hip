// stencil.hip -- 7/27-point averaging stencil, checksum reductions and
// conversions between canonical cell arrays and records.
//
// Stencil: one workgroup per (active block, variable), one thread per
// (j, k) column (up to kCols columns per thread).  The workgroup walks the
// x planes 0..nx+1 once: each canonical plane is gathered from the record
// (interior rows, face planes, edge lines) into registers one step ahead,
// parked in a 3-deep LDS ring, and every thread derives from it what its
// columns need:
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

template <typename T, int ST, int NCOL, int NPER>
__global__ __launch_bounds__(kThreads) void k_stencil(T *__restrict__ pool,
                                                       const int32_t *__restrict__ sorted,
                                                       Lay L, int var0,
                                                       double *__restrict__ partials,
                                                       int part_stride)
{
   typedef typename Acc<T>::type A;
   extern __shared__ __align__(16) unsigned char smem[];
   const int plane = L.gy * L.gz, ncol = L.ny * L.nz;
   A *ring = (A *) smem;
   __shared__ double red[kThreads / 64];

   const int in = blockIdx.x, var = var0 + blockIdx.y;
   T *rec = pool + (int64_t) var * L.var_stride + (int64_t) sorted[in] * L.rec;
   const int tid = threadIdx.x;

   // record offsets of this thread's plane elements, per plane kind
   // (offsets inside a plane depend on p only through the plane's base)
   T pre[NPER];
   auto prefetch = [&](int p) {
#pragma unroll
      for (int q = 0; q < NPER; q++) {
         const int e = tid + q * kThreads;
         if (e < plane) {
            const int j = e / L.gz, k = e - j * L.gz;
            pre[q] = rec[home(L, p, j, k)];
         }
      }
   };
   auto park = [&](int p) {
#pragma unroll
      for (int q = 0; q < NPER; q++) {
         const int e = tid + q * kThreads;
         if (e < plane)
            ring[(p % 3) * plane + e] = to_acc(pre[q]);
      }
   };

   A s_prev[NCOL], s_cur[NCOL], c_prev[NCOL];
   double bsum = 0.0;
   prefetch(0);
   park(0);
   prefetch(1);
   __syncthreads();
#pragma unroll
   for (int q = 0; q < NCOL; q++) {
      const int col = tid + q * kThreads;
      if (col < ncol) {
         const int j = col / L.nz + 1, k = col - (j - 1) * L.nz + 1, c = j * L.gz + k;
         if (ST == 27)
            s_prev[q] = sum9(ring, c, L.gz);
         else
            c_prev[q] = ring[c];
      }
   }
   park(1);
   if (2 <= L.nx + 1)
      prefetch(2);
   __syncthreads();
   if (ST == 27) {
#pragma unroll
      for (int q = 0; q < NCOL; q++) {
         const int col = tid + q * kThreads;
         if (col < ncol) {
            const int j = col / L.nz + 1, k = col - (j - 1) * L.nz + 1, c = j * L.gz + k;
            s_cur[q] = sum9(ring + plane, c, L.gz);
         }
      }
   }
   for (int i = 1; i <= L.nx; i++) {
      // plane i+1 into the ring slot of plane i-2 (last read two steps ago)
      park(i + 1);
      if (i + 2 <= L.nx + 1)
         prefetch(i + 2);
      __syncthreads();
      const A *pc = ring + (i % 3) * plane;
      const A *pn = ring + ((i + 1) % 3) * plane;
#pragma unroll
      for (int q = 0; q < NCOL; q++) {
         const int col = tid + q * kThreads;
         if (col >= ncol)
            continue;
         const int j = col / L.nz + 1, k = col - (j - 1) * L.nz + 1, c = j * L.gz + k;
         A w;
         if (ST == 27) {
            const A s_next = sum9(pn, c, L.gz);
            w = (s_prev[q] + s_cur[q]) + s_next;
            w = w / A(27);
            s_prev[q] = s_cur[q];
            s_cur[q] = s_next;
         } else {
            const A cc = pc[c];
            w = c_prev[q] + pc[c - L.gz];
            w = w + pc[c - 1];
            w = w + cc;
            w = w + pc[c + 1];
            w = w + pc[c + L.gz];
            w = w + pn[c];
            w = w / A(7);
            c_prev[q] = cc;
         }
         const T out = from_acc<T>(w);
         rec[((i - 1) * L.ny + (j - 1)) * L.nz + (k - 1)] = out;
         if (k == 1 || k == L.nz)
            rec[zcopy(L, i, j, k)] = out;
         if (partials)
            bsum += (double) to_acc(out);
      }
   }
   if (partials) {
      double s = block_reduce(bsum, red);
      if (threadIdx.x == 0)
         partials[(int64_t) var * part_stride + in] = s;
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

template <typename T, int ST>
int launch_st(amr_dev *d, dim3 grid, double *parts, int start)
{
   const Lay &L = d->g;
   const int ncol = L.ny * L.nz, plane = L.gy * L.gz;
   size_t lds = sizeof(typename Acc<T>::type) * 3 * (size_t) plane;
#define GO(NC, NP)                                                                        \
   hipLaunchKernelGGL((k_stencil<T, ST, NC, NP>), grid, dim3(kThreads), lds, d->stream,   \
                      (T *) d->pool, d->d_sorted, L, start, parts, d->c.max_blocks)
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
   dim3 grid(d->n_active, num);
   double *parts = want_partials ? d->d_partials : nullptr;
   return d->g.stencil == 7 ? launch_st<T, 7>(d, grid, parts, start)
                            : launch_st<T, 27>(d, grid, parts, start);
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
