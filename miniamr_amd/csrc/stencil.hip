// stencil.hip -- 7/27-point averaging stencil, checksum reductions and
// host<->device conversions.
//
// Stencil: one workgroup per (active block, variable).  The block's whole
// (nx+2)(ny+2)(nz+2) slab is staged in LDS with coalesced loads along z,
// every interior cell is computed from LDS and written back in place (the
// workgroup is the only reader of its slab, so in-place is safe once the
// slab is staged).  Sums follow the reference's association exactly
// (ref/stencil.c:82-101 and 102-144):
//   7-point:  (((((W + S) + D) + C) + U) + N) + E, then / 7
//   27-point: each i-plane's 9 terms summed left to right over (j, k),
//             then (back + mid + front) / 27
// so the FP64 path is bit-identical to the reference.  When asked, the
// workgroup also reduces the values it stored into a per-block FP64
// partial sum (the checksum of ref/check_sum.c:36-65 fused into the
// stencil epilogue); a second small kernel reduces the partials of each
// variable in a fixed order.
#include <stdlib.h>
#include <string.h>

#include "amr_internal.h"

namespace {

constexpr int kThreads = 256;

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

template <typename T, int ST>
__global__ __launch_bounds__(kThreads) void k_stencil(T *__restrict__ pool,
                                                       const int32_t *__restrict__ sorted,
                                                       Geo g, int var0,
                                                       double *__restrict__ partials,
                                                       int part_stride)
{
   typedef typename Acc<T>::type A;
   extern __shared__ __align__(16) unsigned char smem[];
   A *tile = (A *) smem;
   __shared__ double red[kThreads / 64];

   const int in = blockIdx.x, var = var0 + blockIdx.y;
   T *base = pool + (int64_t) var * g.var_stride + (int64_t) sorted[in] * g.slab;
   const int cells = g.gx * g.gy * g.gz;
   for (int t = threadIdx.x; t < cells; t += kThreads)
      tile[t] = to_acc(base[t]);
   __syncthreads();

   const int sx = g.gy * g.gz, sy = g.gz;
   const int nyz = g.ny * g.nz, ncell = g.nx * nyz;
   double bsum = 0.0;
   for (int t = threadIdx.x; t < ncell; t += kThreads) {
      const int i = t / nyz + 1, r = t - (i - 1) * nyz;
      const int j = r / g.nz + 1, k = r - (j - 1) * g.nz + 1;
      const int c = i * sx + j * sy + k;
      A w;
      if (ST == 7) {
         w = tile[c - sx] + tile[c - sy];
         w = w + tile[c - 1];
         w = w + tile[c];
         w = w + tile[c + 1];
         w = w + tile[c + sy];
         w = w + tile[c + sx];
         w = w / A(7);
      } else {
         A s[3];
#pragma unroll
         for (int p = 0; p < 3; p++) {
            const int cc = c + (p - 1) * sx;
            A a = tile[cc - sy - 1] + tile[cc - sy];
            a = a + tile[cc - sy + 1];
            a = a + tile[cc - 1];
            a = a + tile[cc];
            a = a + tile[cc + 1];
            a = a + tile[cc + sy - 1];
            a = a + tile[cc + sy];
            a = a + tile[cc + sy + 1];
            s[p] = a;
         }
         w = (s[0] + s[1]) + s[2];
         w = w / A(27);
      }
      const T out = from_acc<T>(w);
      base[c] = out;
      if (partials)
         bsum += (double) to_acc(out);
   }
   if (partials) {
      double s = block_reduce(bsum, red);
      if (threadIdx.x == 0)
         partials[(int64_t) var * part_stride + in] = s;
   }
}

// Plane-streaming stencil: one workgroup per (block, variable), one thread
// per (j, k) column of the block.  The workgroup walks the x planes
// 0..nx+1 once: each plane is prefetched into registers one step ahead,
// parked in a 3-deep LDS ring, and every thread derives from it what its
// column needs:
//   27-point: the plane's 9-term sum S(p) at (j, k) in the reference's
//             (j, k) order -- exactly the reference's sb/sm/sf for the
//             outputs at x = p+1, p, p-1 -- so each output costs one 9-term
//             sum, and w(i) = ((S(i-1) + S(i)) + S(i+1)) / 27 keeps the
//             reference's association;
//   7-point:  the four in-plane neighbours of plane i plus the column
//             values of planes i-1 and i+1.
// Output rows are written whole (the z-ghosts of the row are written back
// unchanged from the staged copy) so every written cache line is full.
template <typename T, int ST>
__global__ __launch_bounds__(kThreads) void k_stencil_stream(T *__restrict__ pool,
                                                              const int32_t *__restrict__ sorted,
                                                              Geo g, int var0,
                                                              double *__restrict__ partials,
                                                              int part_stride)
{
   typedef typename Acc<T>::type A;
   extern __shared__ __align__(16) unsigned char smem[];
   const int plane = g.gy * g.gz;
   A *ring = (A *) smem;                  // 3 planes
   __shared__ double red[kThreads / 64];

   const int in = blockIdx.x, var = var0 + blockIdx.y;
   T *base = pool + (int64_t) var * g.var_stride + (int64_t) sorted[in] * g.slab;
   const int tid = threadIdx.x;
   const bool active = tid < g.ny * g.nz;
   const int j = tid / g.nz + 1, k = tid - (j - 1) * g.nz + 1;
   const int c = j * g.gz + k;            // column offset within a plane

   // each thread stages at most kMaxPer elements of a plane
   constexpr int kMaxPer = 4;
   T pre[kMaxPer];
#define PREFETCH(p)                                                        \
   _Pragma("unroll") for (int q = 0; q < kMaxPer; q++) {                   \
      const int e = tid + q * kThreads;                                    \
      if (e < plane) pre[q] = base[(int64_t) (p) * plane + e];             \
   }
#define PARK(p)                                                            \
   _Pragma("unroll") for (int q = 0; q < kMaxPer; q++) {                   \
      const int e = tid + q * kThreads;                                    \
      if (e < plane) ring[((p) % 3) * plane + e] = to_acc(pre[q]);         \
   }

   double bsum = 0.0;
   PREFETCH(0);
   PARK(0);
   PREFETCH(1);
   __syncthreads();
   A s_prev = 0, s_cur = 0, c_prev = 0;
   if (active) {
      const A *pl = ring;
      if (ST == 27) {
         A a = pl[c - g.gz - 1] + pl[c - g.gz];
         a = a + pl[c - g.gz + 1];
         a = a + pl[c - 1];
         a = a + pl[c];
         a = a + pl[c + 1];
         a = a + pl[c + g.gz - 1];
         a = a + pl[c + g.gz];
         a = a + pl[c + g.gz + 1];
         s_prev = a;
      } else {
         c_prev = pl[c];
      }
   }
   PARK(1);
   if (2 <= g.nx + 1) { PREFETCH(2); }
   __syncthreads();
   if (active && ST == 27) {
      const A *pl = ring + plane;
      A a = pl[c - g.gz - 1] + pl[c - g.gz];
      a = a + pl[c - g.gz + 1];
      a = a + pl[c - 1];
      a = a + pl[c];
      a = a + pl[c + 1];
      a = a + pl[c + g.gz - 1];
      a = a + pl[c + g.gz];
      a = a + pl[c + g.gz + 1];
      s_cur = a;
   }
   for (int i = 1; i <= g.nx; i++) {
      // plane i+1 into the ring slot of plane i-2 (last read two steps ago)
      PARK(i + 1);
      if (i + 2 <= g.nx + 1) { PREFETCH(i + 2); }
      __syncthreads();
      if (active) {
         const A *pn = ring + ((i + 1) % 3) * plane;
         A w;
         if (ST == 27) {
            A a = pn[c - g.gz - 1] + pn[c - g.gz];
            a = a + pn[c - g.gz + 1];
            a = a + pn[c - 1];
            a = a + pn[c];
            a = a + pn[c + 1];
            a = a + pn[c + g.gz - 1];
            a = a + pn[c + g.gz];
            a = a + pn[c + g.gz + 1];
            w = (s_prev + s_cur) + a;
            w = w / A(27);
            s_prev = s_cur;
            s_cur = a;
         } else {
            const A *pc = ring + (i % 3) * plane;
            const A cc = pc[c];
            w = c_prev + pc[c - g.gz];
            w = w + pc[c - 1];
            w = w + cc;
            w = w + pc[c + 1];
            w = w + pc[c + g.gz];
            w = w + pn[c];
            w = w / A(7);
            c_prev = cc;
         }
         const T out = from_acc<T>(w);
         T *row = base + (int64_t) i * plane + j * g.gz;
         row[k] = out;
         // rewrite the row's z-ghosts unchanged so the row's lines are whole
         const A *pc = ring + (i % 3) * plane;
         if (k == 1)
            row[0] = from_acc<T>(pc[c - 1]);
         if (k == g.nz)
            row[g.nz + 1] = from_acc<T>(pc[c + 1]);
         if (partials)
            bsum += (double) to_acc(out);
      }
   }
#undef PREFETCH
#undef PARK
   if (partials) {
      double s = block_reduce(bsum, red);
      if (threadIdx.x == 0)
         partials[(int64_t) var * part_stride + in] = s;
   }
}

// Per-block interior sums of variables [var0, var0 + gridDim.y) without
// touching the cells (stand-alone check_sum).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_block_sums(const T *__restrict__ pool,
                                                          const int32_t *__restrict__ sorted,
                                                          Geo g, int var0,
                                                          double *__restrict__ partials,
                                                          int part_stride)
{
   __shared__ double red[kThreads / 64];
   const int in = blockIdx.x, var = var0 + blockIdx.y;
   const T *base = pool + (int64_t) var * g.var_stride + (int64_t) sorted[in] * g.slab;
   const int nyz = g.ny * g.nz, ncell = g.nx * nyz;
   double bsum = 0.0;
   for (int t = threadIdx.x; t < ncell; t += kThreads) {
      const int i = t / nyz + 1, r = t - (i - 1) * nyz;
      const int j = r / g.nz + 1, k = r - (j - 1) * g.nz + 1;
      bsum += (double) to_acc(base[((int64_t) i * g.gy + j) * g.gz + k]);
   }
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

template <typename T>
__global__ void k_convert_in(T *__restrict__ pool, const double *__restrict__ stage, Geo g,
                             int slot, int cells)
{
   const int var = blockIdx.y;
   T *dst = pool + (int64_t) var * g.var_stride + (int64_t) slot * g.slab;
   const double *src = stage + (int64_t) var * cells;
   for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < cells; t += gridDim.x * blockDim.x)
      dst[t] = from_acc<T>((typename Acc<T>::type) src[t]);
}

template <typename T>
__global__ void k_convert_out(const T *__restrict__ pool, double *__restrict__ stage, Geo g,
                              int slot, int cells)
{
   const int var = blockIdx.y;
   const T *src = pool + (int64_t) var * g.var_stride + (int64_t) slot * g.slab;
   double *dst = stage + (int64_t) var * cells;
   for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < cells; t += gridDim.x * blockDim.x)
      dst[t] = (double) to_acc(src[t]);
}

// The plane-streaming kernel needs one thread per (j, k) column and at most
// 4 staged elements per thread; larger block faces use the whole-slab kernel.
static bool use_stream(const Geo &g)
{
   static int forced = -1;
   if (forced < 0) {
      const char *e = getenv("MINIAMR_STENCIL");
      forced = (e && !strcmp(e, "slab")) ? 0 : 1;
   }
   return forced && g.ny * g.nz <= kThreads && g.gy * g.gz <= 4 * kThreads;
}

template <typename T>
int launch_stencil(amr_dev *d, int start, int num, bool want_partials)
{
   typedef typename Acc<T>::type A;
   if (!d->n_active || !num)
      return AMR_OK;
   dim3 grid(d->n_active, num);
   double *parts = want_partials ? d->d_partials : nullptr;
   if (use_stream(d->g)) {
      size_t lds = sizeof(A) * 3 * (size_t) d->g.gy * d->g.gz;
      if (d->g.stencil == 7)
         hipLaunchKernelGGL((k_stencil_stream<T, 7>), grid, dim3(kThreads), lds, d->stream,
                            (T *) d->pool, d->d_sorted, d->g, start, parts, d->c.max_blocks);
      else
         hipLaunchKernelGGL((k_stencil_stream<T, 27>), grid, dim3(kThreads), lds, d->stream,
                            (T *) d->pool, d->d_sorted, d->g, start, parts, d->c.max_blocks);
      HIP_TRY(d, hipGetLastError());
      return AMR_OK;
   }
   size_t lds = sizeof(A) * (size_t) d->g.gx * d->g.gy * d->g.gz;
   static bool attr_set[2][2] = {{false, false}, {false, false}};
   const int si = d->g.stencil == 7 ? 0 : 1;
   const void *fn = d->g.stencil == 7 ? (const void *) k_stencil<T, 7>
                                      : (const void *) k_stencil<T, 27>;
   if (lds > 64 * 1024) {
      if (lds > 160 * 1024)
         return amr_set_error(d, AMR_ERR_ARG, "block of %dx%dx%d does not fit in LDS",
                              d->g.nx, d->g.ny, d->g.nz);
      const int pi = sizeof(A) == 8 ? 0 : 1;
      if (!attr_set[pi][si]) {
         HIP_TRY(d, hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int) lds));
         attr_set[pi][si] = true;
      }
   }
   if (d->g.stencil == 7)
      hipLaunchKernelGGL((k_stencil<T, 7>), grid, dim3(kThreads), lds, d->stream,
                         (T *) d->pool, d->d_sorted, d->g, start, parts, d->c.max_blocks);
   else
      hipLaunchKernelGGL((k_stencil<T, 27>), grid, dim3(kThreads), lds, d->stream,
                         (T *) d->pool, d->d_sorted, d->g, start, parts, d->c.max_blocks);
   HIP_TRY(d, hipGetLastError());
   return AMR_OK;
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
