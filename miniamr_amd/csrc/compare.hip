// compare.hip -- amr_dev_compare_cells: the largest difference between the
// interior cells of two contexts holding the same mesh, per variable, in
// any pair of storage precisions.  The precision form's cell error against
// FP64 after each checksum interval (no counterpart in the reference,
// whose cells are FP64 only).
#include <algorithm>

#include "amr_internal.h"

namespace {

__device__ __forceinline__ double as_f64(double v) { return v; }
__device__ __forceinline__ double as_f64(float v) { return v; }
__device__ __forceinline__ double as_f64(uint16_t v) { return (double) to_acc(v); }

// One workgroup per (active block, variable); interiors are contiguous
// [nx][ny][nz] runs at each record's current interior copy.  Differences
// are non-negative, so their IEEE bit patterns order as unsigned integers
// and the maximum is an integer atomic.
template <typename Ta, typename Tb>
__global__ __launch_bounds__(256) void k_compare(const Ta *__restrict__ pa, Lay La,
                                                 const Tb *__restrict__ pb, Lay Lb,
                                                 const int32_t *__restrict__ sorted, int start,
                                                 unsigned long long *__restrict__ out)
{
   const int slot = sorted[blockIdx.x], var = start + blockIdx.y;
   const Ta *a = pa + (int64_t) var * La.var_stride + (int64_t) slot * La.rec + La.oI;
   const Tb *b = pb + (int64_t) var * Lb.var_stride + (int64_t) slot * Lb.rec + Lb.oI;
   const int n = La.nx * La.ny * La.nz;
   double m = 0.0;
   for (int t = threadIdx.x; t < n; t += blockDim.x) {
      const double d = fabs(as_f64(a[t]) - as_f64(b[t]));
      m = d > m ? d : m;                 // a NaN difference leaves m (caught below)
      if (d != d)
         m = __builtin_inf();
   }
   for (int o = 32; o > 0; o >>= 1)
      m = fmax(m, __shfl_xor(m, o));
   __shared__ double w[4];
   if ((threadIdx.x & 63) == 0)
      w[threadIdx.x >> 6] = m;
   __syncthreads();
   if (threadIdx.x == 0) {
      m = fmax(fmax(w[0], w[1]), fmax(w[2], w[3]));
      atomicMax(&out[blockIdx.y], (unsigned long long) __double_as_longlong(m));
   }
}

__global__ void k_zero(unsigned long long *p, int n)
{
   if ((int) threadIdx.x < n)
      p[threadIdx.x] = 0ull;
}

template <typename Ta, typename Tb>
void launch(amr_dev *a, const Lay &La, amr_dev *b, const Lay &Lb, int start, int num)
{
   hipLaunchKernelGGL((k_compare<Ta, Tb>), dim3(a->n_active, num), dim3(256), 0, a->stream,
                      (const Ta *) a->pool, La, (const Tb *) b->pool, Lb, a->d_sorted, start,
                      a->d_cmp);
}

template <typename Ta>
void launch_b(amr_dev *a, const Lay &La, amr_dev *b, const Lay &Lb, int start, int num)
{
   switch (b->c.precision) {
   case AMR_PREC_F64: launch<Ta, double>(a, La, b, Lb, start, num); break;
   case AMR_PREC_F32: launch<Ta, float>(a, La, b, Lb, start, num); break;
   default: launch<Ta, uint16_t>(a, La, b, Lb, start, num); break;
   }
}

}  // namespace

extern "C" int amr_dev_compare_cells(amr_dev *a, amr_dev *b, int start, int num,
                                     double *max_abs)
{
   if (!a || !b || !max_abs || start < 0 || num < 1 || start + num > a->c.num_vars)
      return amr_set_error(a, AMR_ERR_ARG, "compare_cells: bad arguments");
   if (a->c.nx != b->c.nx || a->c.ny != b->c.ny || a->c.nz != b->c.nz ||
       a->c.num_vars != b->c.num_vars || a->c.max_blocks != b->c.max_blocks ||
       a->c.device != b->c.device)
      return amr_set_error(a, AMR_ERR_ARG, "compare_cells: the contexts differ in block size, "
                           "variables, slots or device");
   if (a->h_sorted != b->h_sorted)
      return amr_set_error(a, AMR_ERR_STATE, "compare_cells: the contexts hold different "
                           "meshes");
   Lay La, Lb;
   if (int rc = amr_lay(a, start, num, &La))
      return rc;
   if (int rc = amr_lay(b, start, num, &Lb))
      return amr_set_error(a, rc, "compare_cells: %s", amr_dev_last_error(b));
   if (!a->d_cmp)
      HIP_TRY(a, hipMalloc(&a->d_cmp, sizeof(unsigned long long) * std::max(a->c.num_vars, 1)));
   HIP_TRY(a, hipStreamSynchronize(b->stream));   // b's cells are final
   if (num > 1024)
      return amr_set_error(a, AMR_ERR_ARG, "compare_cells: at most 1024 variables per call");
   hipLaunchKernelGGL(k_zero, dim3(1), dim3(1024), 0, a->stream, a->d_cmp, num);
   if (a->n_active > 0) {
      switch (a->c.precision) {
      case AMR_PREC_F64: launch_b<double>(a, La, b, Lb, start, num); break;
      case AMR_PREC_F32: launch_b<float>(a, La, b, Lb, start, num); break;
      default: launch_b<uint16_t>(a, La, b, Lb, start, num); break;
      }
      HIP_TRY(a, hipGetLastError());
   }
   HIP_TRY(a, hipMemcpyAsync(a->h_pinned, a->d_cmp, sizeof(double) * num, hipMemcpyDeviceToHost,
                             a->stream));
   HIP_TRY(a, hipStreamSynchronize(a->stream));
   for (int v = 0; v < num; v++)
      max_abs[v] = a->h_pinned[v];
   return AMR_OK;
}
