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

template <typename T>
int launch_stencil(amr_dev *d, int start, int num, bool want_partials)
{
   typedef typename Acc<T>::type A;
   if (!d->n_active || !num)
      return AMR_OK;
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
   dim3 grid(d->n_active, num);
   double *parts = want_partials ? d->d_partials : nullptr;
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
