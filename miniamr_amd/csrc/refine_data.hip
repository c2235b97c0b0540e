// refine_data.hip -- cell data of block split and consolidate.
//
// split (ref/block.c:161-173): every child cell of octant o is the parent
// cell it lies in divided by 8.  consolidate (ref/block.c:411-431): every
// cell of the merged block's octant o is the sum of the 2x2x2 child cells
// it covers, added in the reference's order (x fastest, then y, then z).
// Only interior cells (and their z-plane copies) are written; a new block's
// ghost cells keep what the slot held before, as in the reference.
//
// A batch is applied "as if one job after another": a slot freed by an
// earlier job (a split parent, a consolidated child) can be the target of
// a later job.  Each launch pair therefore first snapshots what the batch
// reads (the parent octants, or the merged sums) into scratch and only then
// writes the targets, so reads never see the batch's own writes.
#include "amr_internal.h"

namespace {

// Snapshot of parent octants: scratch[job][var][hx][hy][hz].
template <typename T>
__global__ void k_split_snap(const T *__restrict__ pool, const amr_split_job *__restrict__ jobs,
                             T *__restrict__ scratch, Lay g, int nvars)
{
   const amr_split_job j = jobs[blockIdx.x];
   const int var = blockIdx.y;
   const int hx = g.nx / 2, hy = g.ny / 2, hz = g.nz / 2, hc = hx * hy * hz;
   const int i1 = (j.octant & 1) * hx, j1 = ((j.octant >> 1) & 1) * hy, k1 = (j.octant >> 2) * hz;
   const T *p = pool + (int64_t) var * g.var_stride + (int64_t) j.parent * g.rec;
   T *s = scratch + ((int64_t) blockIdx.x * nvars + var) * hc;
   for (int t = threadIdx.x; t < hc; t += blockDim.x) {
      const int i = t / (hy * hz), r = t - i * hy * hz, jj = r / hz, k = r - jj * hz;
      s[t] = p[g.oI + ((i + i1) * g.ny + (jj + j1)) * g.nz + (k + k1)];
   }
}

template <typename T>
__global__ void k_split_write(T *__restrict__ pool, const amr_split_job *__restrict__ jobs,
                              const T *__restrict__ scratch, Lay g, int nvars)
{
   typedef typename Acc<T>::type A;
   const amr_split_job j = jobs[blockIdx.x];
   const int var = blockIdx.y;
   const int hy = g.ny / 2, hz = g.nz / 2, hc = (g.nx / 2) * hy * hz;
   T *c = pool + (int64_t) var * g.var_stride + (int64_t) j.child * g.rec;
   const T *s = scratch + ((int64_t) blockIdx.x * nvars + var) * hc;
   const int nyz = g.ny * g.nz;
   // Each thread owns (j, k) columns and walks the x planes, so the index
   // arithmetic (divisions by run-time sizes) is done once per column, not
   // once per cell: per cell the kernel was instruction-bound (~2.7 TB/s of
   // writes).  A plane's cells are contiguous in I, so each pass writes a
   // whole plane; the cells at z = 1 / z = nz also go to their z-plane copy
   // (the stores of store_cell, spelled out).
   for (int r = threadIdx.x; r < nyz; r += blockDim.x) {
      const int jj = r / g.nz, k = r - jj * g.nz;
      const int so = (jj >> 1) * hz + (k >> 1);
      const int zs = k == g.nz - 1 ? 1 : 0;
      const bool zc = k == 0 || k == g.nz - 1;
      for (int i = 0; i < g.nx; i++) {
         const A v = to_acc(s[(i >> 1) * hy * hz + so]) / A(8);
         const T x = from_acc<T>(v);
         c[g.oI + (i * g.ny + jj) * g.nz + k] = x;
         if (zc)
            c[g.oZI + (zs * g.nx + i) * g.ny + jj] = x;
      }
   }
}

// Merged sums: scratch[job][var][hx][hy][hz].
template <typename T>
__global__ void k_merge_sum(const T *__restrict__ pool, const amr_merge_job *__restrict__ jobs,
                            T *__restrict__ scratch, Lay g, int nvars)
{
   typedef typename Acc<T>::type A;
   const amr_merge_job j = jobs[blockIdx.x];
   const int var = blockIdx.y;
   const int hy = g.ny / 2, hz = g.nz / 2, hc = (g.nx / 2) * hy * hz;
   const T *c = pool + (int64_t) var * g.var_stride + (int64_t) j.child * g.rec + g.oI;
   T *s = scratch + ((int64_t) blockIdx.x * nvars + var) * hc;
   const int sx = g.ny * g.nz, sy = g.nz;
   for (int t = threadIdx.x; t < hc; t += blockDim.x) {
      const int i = t / (hy * hz), r = t - i * hy * hz, jj = r / hz, k = r - jj * hz;
      const int o = (2 * i) * sx + (2 * jj) * sy + (2 * k);
      A v = to_acc(c[o]);
      v = v + to_acc(c[o + sx]);
      v = v + to_acc(c[o + sy]);
      v = v + to_acc(c[o + sx + sy]);
      v = v + to_acc(c[o + 1]);
      v = v + to_acc(c[o + sx + 1]);
      v = v + to_acc(c[o + sy + 1]);
      v = v + to_acc(c[o + sx + sy + 1]);
      s[t] = from_acc<T>(v);
   }
}

template <typename T>
__global__ void k_merge_write(T *__restrict__ pool, const amr_merge_job *__restrict__ jobs,
                              const T *__restrict__ scratch, Lay g, int nvars)
{
   const amr_merge_job j = jobs[blockIdx.x];
   const int var = blockIdx.y;
   const int hx = g.nx / 2, hy = g.ny / 2, hz = g.nz / 2, hc = hx * hy * hz;
   const int i1 = (j.octant & 1) * hx, j1 = ((j.octant >> 1) & 1) * hy, k1 = (j.octant >> 2) * hz;
   T *p = pool + (int64_t) var * g.var_stride + (int64_t) j.dst * g.rec;
   const T *s = scratch + ((int64_t) blockIdx.x * nvars + var) * hc;
   for (int t = threadIdx.x; t < hc; t += blockDim.x) {
      const int i = t / (hy * hz), r = t - i * hy * hz, jj = r / hz, k = r - jj * hz;
      store_cell(p, g, i + 1 + i1, jj + 1 + j1, k + 1 + k1, s[t]);
   }
}

template <typename T>
int split(amr_dev *d, const amr_split_job *jobs, int n, void *scratch)
{
   dim3 grid(n, d->c.num_vars);
   Lay L;
   if (int rc = amr_lay(d, 0, d->c.num_vars, &L))
      return rc;
   hipLaunchKernelGGL(k_split_snap<T>, grid, dim3(256), 0, d->stream, (const T *) d->pool, jobs,
                      (T *) scratch, L, d->c.num_vars);
   HIP_TRY(d, hipGetLastError());
   hipLaunchKernelGGL(k_split_write<T>, grid, dim3(256), 0, d->stream, (T *) d->pool, jobs,
                      (const T *) scratch, L, d->c.num_vars);
   HIP_TRY(d, hipGetLastError());
   return AMR_OK;
}

template <typename T>
int merge(amr_dev *d, const amr_merge_job *jobs, int n, void *scratch)
{
   dim3 grid(n, d->c.num_vars);
   Lay L;
   if (int rc = amr_lay(d, 0, d->c.num_vars, &L))
      return rc;
   hipLaunchKernelGGL(k_merge_sum<T>, grid, dim3(256), 0, d->stream, (const T *) d->pool, jobs,
                      (T *) scratch, L, d->c.num_vars);
   HIP_TRY(d, hipGetLastError());
   hipLaunchKernelGGL(k_merge_write<T>, grid, dim3(256), 0, d->stream, (T *) d->pool, jobs,
                      (const T *) scratch, L, d->c.num_vars);
   HIP_TRY(d, hipGetLastError());
   return AMR_OK;
}

}  // namespace

int amr_launch_split(amr_dev *d, const amr_split_job *jobs, int n, void *scratch)
{
   switch (d->c.precision) {
   case AMR_PREC_F64: return split<double>(d, jobs, n, scratch);
   case AMR_PREC_F32: return split<float>(d, jobs, n, scratch);
   default: return split<uint16_t>(d, jobs, n, scratch);
   }
}

int amr_launch_merge(amr_dev *d, const amr_merge_job *jobs, int n, void *scratch)
{
   switch (d->c.precision) {
   case AMR_PREC_F64: return merge<double>(d, jobs, n, scratch);
   case AMR_PREC_F32: return merge<float>(d, jobs, n, scratch);
   default: return merge<uint16_t>(d, jobs, n, scratch);
   }
}
