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

#include "kern_stencil.h"
#include "kern_stencil_vec.h"

namespace {

using namespace amrk;

template <typename T, int ST, int NCOL, int NPER, int VPW, int NZ, int D = kDepth>
__global__ __launch_bounds__(kThreads) void k_stencil(T *__restrict__ pool,
                                                       const int32_t *__restrict__ sorted,
                                                       Lay L, int var0, int var_end,
                                                       double *__restrict__ partials,
                                                       int part_stride, int remap)
{
   typedef typename Acc<T>::type A;
   extern __shared__ __align__(16) unsigned char smem[];
   __shared__ double red[kThreads / 64];
   int in = blockIdx.x, vg = blockIdx.y;
   if (remap) {
      const int q = xcd_remap(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
      in = q % gridDim.x;
      vg = q / gridDim.x;
   }
   const int v0 = var0 + vg * VPW;
   const int nv = min(VPW, var_end - v0);
   T *rec = pool + (int64_t) v0 * L.var_stride + (int64_t) sorted[in] * L.rec;
   stencil_block<T, ST, NCOL, NPER, VPW, kThreads, NZ, D>(rec, L, nv, (A *) smem, red,
                                         partials ? partials + (int64_t) v0 * part_stride + in
                                                  : nullptr,
                                         part_stride);
}

// The stencil of a fused stage: interiors read from the copy L describes,
// written to the copy Lo describes, same-level faces read in place through
// redir (six slots per active block, sorted order), ghost edge lines and
// corners where edges (kEdgeSrcs entries per slot) says.
template <typename T, int ST, int NC, int NP, int V, int NZ>
__global__ __launch_bounds__(kThreads) void k_stencil_fused(T *__restrict__ pool,
                                                             const int32_t *__restrict__ sorted,
                                                             const int32_t *__restrict__ redir,
                                                             const EdgeSrc *__restrict__ edges,
                                                             Lay L, Lay Lo, int var0, int var_end,
                                                             double *__restrict__ partials,
                                                             int part_stride, int remap)
{
   typedef typename Acc<T>::type A;
   extern __shared__ __align__(16) unsigned char smem[];
   __shared__ double red[kThreads / 64];
   int in = blockIdx.x, vg = blockIdx.y;
   if (remap) {
      const int q = xcd_remap(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
      in = q % gridDim.x;
      vg = q / gridDim.x;
   }
   const int v0 = var0 + vg * V;
   const int nv = min(V, var_end - v0);
   const int slot = sorted[in];
   T *rec = pool + (int64_t) v0 * L.var_stride + (int64_t) slot * L.rec;
   stencil_block<T, ST, NC, NP, V, kThreads, NZ, kDepth, true>(
      rec, L, nv, (A *) smem, red, partials ? partials + (int64_t) v0 * part_stride + in : nullptr,
      part_stride, &Lo, slot, redir + 6 * (int64_t) in, edges + (int64_t) kEdgeSrcs * slot);
}

// The 27-point stencil of blocks with 16 x 16 faces, C columns per thread
// (kern_stencil_vec.h), plain (interiors read and written in the copy L
// describes) or fused (as k_stencil_fused).
template <typename T, bool FUSED, int D>
__global__ __launch_bounds__(kThreads) void k_stencil_v16(T *__restrict__ pool,
                                                           const int32_t *__restrict__ sorted,
                                                           const int32_t *__restrict__ redir,
                                                           const EdgeSrc *__restrict__ edges,
                                                           Lay L, Lay Lo, int var0, int var_end,
                                                           double *__restrict__ partials,
                                                           int part_stride)
{
   typedef typename Acc<T>::type A;
   constexpr int V = VecVars<T>::v;
   extern __shared__ __align__(16) unsigned char smem[];
   __shared__ double red[kThreads / 64];
   const int in = blockIdx.x;
   const int v0 = var0 + blockIdx.y * V;
   const int nv = min(V, var_end - v0);
   const int slot = sorted[in];
   T *rec = pool + (int64_t) v0 * L.var_stride + (int64_t) slot * L.rec;
   stencil27_v16<T, D, FUSED>(
      rec, L, nv, (A *) smem, red, partials ? partials + (int64_t) v0 * part_stride + in : nullptr,
      part_stride, FUSED ? &Lo : nullptr, slot, FUSED ? redir + 6 * (int64_t) in : nullptr,
      FUSED ? edges + (int64_t) kEdgeSrcs * slot : nullptr);
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
   const T *rec = pool + (int64_t) var * L.var_stride + (int64_t) sorted[in] * L.rec + L.oI;
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

// Variables per workgroup: the per-cell index work, plane barriers and
// ring bookkeeping are shared by several variables.  Measured on the 16^3
// workload, fused stages, ms per launch for 1 / 2 / 4 variables: FP64
// 3.240 / 3.093 / 3.170, FP32 2.537 / 1.825 / 1.805, BF16 2.166 / 1.535 /
// 1.443; so two for FP64, four for FP32 and BF16.  MINIAMR_STENCIL_VPW=
// 1|2|4 overrides the choice.
template <typename T> struct Vpw { static constexpr int v = sizeof(T) == 8 ? 2 : 4; };

template <typename T>
static int stencil_vpw()
{
   static int v = -1;
   if (v < 0) {
      const char *e = getenv("MINIAMR_STENCIL_VPW");
      v = e ? atoi(e) : Vpw<T>::v;
      if (v != 1 && v != 2 && v != 4)
         v = Vpw<T>::v;
   }
   return v;
}

// MINIAMR_XCD_REMAP=1 gives each XCD a contiguous share of the stencil's
// workgroups (xcd_remap).  Off by default: on the 16^3 workload it made the
// fused FP64 stencil 1.7% slower (2.943 against 2.893 ms per launch) -- the
// blocks next to each other in sorted order are not the ones whose faces
// a workgroup reads, and the round-robin spread over XCDs balances the
// memory channels.
static int xcd_remap_on()
{
   static int v = -1;
   if (v < 0) {
      const char *e = getenv("MINIAMR_XCD_REMAP");
      v = e ? atoi(e) != 0 : 0;
   }
   return v;
}

static int depth6()
{
   static int v = -1;
   if (v < 0) {
      const char *e = getenv("MINIAMR_STENCIL_DEPTH");
      v = e && atoi(e) == 6;
   }
   return v;
}

// The vector form (k_stencil_v16) serves 27-point stencils of 16 x 16
// faces whose interiors and x faces start on 16-byte boundaries, unless
// MINIAMR_STENCIL_VEC=0.  Plain and fused stages then both use it, so their
// checksum partials are summed in the same order.
static bool stencil_vec_on()
{
   static int v = -1;
   if (v < 0) {
      const char *e = getenv("MINIAMR_STENCIL_VEC");
      v = e ? atoi(e) != 0 : 1;
   }
   return v != 0;
}

template <typename T>
static bool use_vec(const amr_dev *d, int st)
{
   const int C = VecCols<T>::c;
   const Lay &g = d->g;
   return st == 27 && g.ny == 16 && g.nz == 16 && stencil_vec_on() && d->oI_par[0] % C == 0 &&
          d->oI_par[1] % C == 0 && g.oF[0] % C == 0 && g.oF[1] % C == 0 && g.rec % C == 0;
}

template <typename T, bool FUSED>
int launch_v16(amr_dev *d, const Lay &L, const Lay &Lo, int start, int num, double *parts)
{
   constexpr int V = VecVars<T>::v;
   dim3 grid(d->n_active, (num + V - 1) / V);
   const size_t lds = sizeof(typename Acc<T>::type) * 3 * V * 18 * 18;
   hipLaunchKernelGGL((k_stencil_v16<T, FUSED, VecDepth<T>::d>), grid, dim3(kThreads), lds,
                      d->stream, (T *) d->pool, d->d_sorted, d->d_redir, d->d_edges, L, Lo, start,
                      start + num, parts, d->c.max_blocks);
   HIP_TRY(d, hipGetLastError());
   return AMR_OK;
}

template <typename T, int ST, int V>
int launch_st_v(amr_dev *d, int start, int num, double *parts)
{
   Lay L;
   if (int rc = amr_lay(d, start, num, &L))
      return rc;
   const int ncol = L.ny * L.nz, plane = L.gy * L.gz;
   dim3 grid(d->n_active, (num + V - 1) / V);
   size_t lds = sizeof(typename Acc<T>::type) * 3 * V * (size_t) plane;
#define GO(NC, NP, NZ)                                                                      \
   hipLaunchKernelGGL((k_stencil<T, ST, NC, NP, V, NZ>), grid, dim3(kThreads), lds, d->stream, \
                      (T *) d->pool, d->d_sorted, L, start, start + num, parts,             \
                      d->c.max_blocks, xcd_remap_on())
   // common face widths get a compile-time nz (constant LDS offsets)
   if (ncol <= kThreads && plane <= 2 * kThreads) {
      if (L.nz == 16 && depth6())
         hipLaunchKernelGGL((k_stencil<T, ST, 1, 2, V, 16, 6>), grid, dim3(kThreads), lds,
                            d->stream, (T *) d->pool, d->d_sorted, L, start, start + num, parts,
                            d->c.max_blocks, xcd_remap_on());
      else if (L.nz == 16) GO(1, 2, 16);
      else if (L.nz == 12) GO(1, 2, 12);
      else if (L.nz == 8) GO(1, 2, 8);
      else GO(1, 2, 0);
   } else if (ncol <= 2 * kThreads && plane <= 4 * kThreads)
      GO(2, 4, 0);
   else if (ncol <= kCols * kThreads && plane <= kPer * kThreads)
      GO(kCols, kPer, 0);
   else
      return amr_set_error(d, AMR_ERR_ARG, "block face %dx%d too large for the stencil kernel",
                           L.ny, L.nz);
#undef GO
   HIP_TRY(d, hipGetLastError());
   return AMR_OK;
}

template <typename T, int ST>
int launch_st(amr_dev *d, int start, int num, double *parts)
{
   if (use_vec<T>(d, ST)) {
      Lay L;
      if (int rc = amr_lay(d, start, num, &L))
         return rc;
      return launch_v16<T, false>(d, L, L, start, num, parts);
   }
   if (stencil_vpw<T>() == 4)
      return launch_st_v<T, ST, 4>(d, start, num, parts);
   if (stencil_vpw<T>() == 1)
      return launch_st_v<T, ST, 1>(d, start, num, parts);
   return launch_st_v<T, ST, 2>(d, start, num, parts);
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

template <typename T, int ST, int V>
int launch_fused_v(amr_dev *d, int start, int num, double *parts)
{
   Lay L;
   if (int rc = amr_lay(d, start, num, &L))
      return rc;
   const Lay Lo = amr_lay_other(d, L);
   const int ncol = L.ny * L.nz, plane = L.gy * L.gz;
   dim3 grid(d->n_active, (num + V - 1) / V);
   size_t lds = sizeof(typename Acc<T>::type) * 3 * V * (size_t) plane;
#define GO(NC, NP, NZ)                                                                        \
   hipLaunchKernelGGL((k_stencil_fused<T, ST, NC, NP, V, NZ>), grid, dim3(kThreads), lds,        \
                      d->stream, (T *) d->pool, d->d_sorted, d->d_redir, d->d_edges, L, Lo,     \
                      start,                                                                    \
                      start + num, parts, d->c.max_blocks, xcd_remap_on())
   if (ncol <= kThreads && plane <= 2 * kThreads) {
      if (L.nz == 16) GO(1, 2, 16);
      else if (L.nz == 12) GO(1, 2, 12);
      else if (L.nz == 8) GO(1, 2, 8);
      else GO(1, 2, 0);
   } else if (ncol <= 2 * kThreads && plane <= 4 * kThreads)
      GO(2, 4, 0);
   else if (ncol <= kCols * kThreads && plane <= kPer * kThreads)
      GO(kCols, kPer, 0);
   else
      return amr_set_error(d, AMR_ERR_ARG, "block face %dx%d too large for the stencil kernel",
                           L.ny, L.nz);
#undef GO
   HIP_TRY(d, hipGetLastError());
   // the variables' cells now live in the other copy
   for (int v = start; v < start + num; v++)
      d->par[v] ^= 1;
   return AMR_OK;
}

// variables per workgroup as for the plain stencil (MINIAMR_STENCIL_VPW)
template <typename T, int ST>
int launch_fused(amr_dev *d, int start, int num, double *parts)
{
   if (use_vec<T>(d, ST)) {
      Lay L;
      if (int rc = amr_lay(d, start, num, &L))
         return rc;
      if (int rc = launch_v16<T, true>(d, L, amr_lay_other(d, L), start, num, parts))
         return rc;
      for (int v = start; v < start + num; v++)
         d->par[v] ^= 1;   // the variables' cells now live in the other copy
      return AMR_OK;
   }
   if (stencil_vpw<T>() == 4)
      return launch_fused_v<T, ST, 4>(d, start, num, parts);
   if (stencil_vpw<T>() == 1)
      return launch_fused_v<T, ST, 1>(d, start, num, parts);
   return launch_fused_v<T, ST, 2>(d, start, num, parts);
}

template <typename T>
int launch_block_sums(amr_dev *d, int start, int num)
{
   if (!d->n_active || !num)
      return AMR_OK;
   Lay L;
   if (int rc = amr_lay(d, start, num, &L))
      return rc;
   hipLaunchKernelGGL(k_block_sums<T>, dim3(d->n_active, num), dim3(kThreads), 0, d->stream,
                      (const T *) d->pool, d->d_sorted, L, start, d->d_partials,
                      d->c.max_blocks);
   HIP_TRY(d, hipGetLastError());
   return AMR_OK;
}

template <typename T>
int launch_convert(amr_dev *d, int slot, bool in)
{
   int cells = d->g.gx * d->g.gy * d->g.gz;
   dim3 grid((cells + 255) / 256, d->c.num_vars);
   Lay L;
   if (int rc = amr_lay(d, 0, d->c.num_vars, &L))
      return rc;
   if (in)
      hipLaunchKernelGGL(k_convert_in<T>, grid, dim3(256), 0, d->stream, (T *) d->pool,
                         (const double *) d->d_stage, L, slot, cells);
   else
      hipLaunchKernelGGL(k_convert_out<T>, grid, dim3(256), 0, d->stream,
                         (const T *) d->pool, d->d_stage, L, slot, cells);
   HIP_TRY(d, hipGetLastError());
   return AMR_OK;
}

}  // namespace

int amr_launch_stencil_fused(amr_dev *d, int start, int num, bool want_partials)
{
   if (!d->n_active || !num) {
      for (int v = start; v < start + num; v++)
         d->par[v] ^= 1;   // nothing to compute: the stage still flips the copy
      return AMR_OK;
   }
   double *parts = want_partials ? d->d_partials : nullptr;
   const bool s7 = d->g.stencil == 7;
   switch (d->c.precision) {
   case AMR_PREC_F64: return s7 ? launch_fused<double, 7>(d, start, num, parts)
                                : launch_fused<double, 27>(d, start, num, parts);
   case AMR_PREC_F32: return s7 ? launch_fused<float, 7>(d, start, num, parts)
                                : launch_fused<float, 27>(d, start, num, parts);
   default: return s7 ? launch_fused<uint16_t, 7>(d, start, num, parts)
                      : launch_fused<uint16_t, 27>(d, start, num, parts);
   }
}

// --stencil 0 (ref/stencil.c:43-73): variables in order, since each may read
// the variables before it after their update of this stage; variable 0 and
// variables from 4*mat on take the 7-point average, the others the
// variable-work updates.  Checksum partials come from a block-sum pass.
template <typename T>
int launch_stencil0(amr_dev *d, int start, int num, bool want_partials, int calc_stage)
{
   if (!d->n_active || !num)
      return AMR_OK;
   for (int v = start; v < start + num; v++) {
      int rc = (v == 0 || v >= 4 * d->mat) ? launch_st_v<T, 7, 1>(d, v, 1, nullptr)
                                           : amr_launch_mat_stencil(d, v, calc_stage);
      if (rc)
         return rc;
   }
   return want_partials ? launch_block_sums<T>(d, start, num) : AMR_OK;
}

int amr_launch_stencil(amr_dev *d, int start, int num, bool want_partials, int calc_stage)
{
   if (d->g.stencil == 0) {
      if (!d->mat)
         return amr_set_error(d, AMR_ERR_STATE, "stencil 0: amr_dev_set_materials not called");
      switch (d->c.precision) {
      case AMR_PREC_F64: return launch_stencil0<double>(d, start, num, want_partials, calc_stage);
      case AMR_PREC_F32: return launch_stencil0<float>(d, start, num, want_partials, calc_stage);
      default: return launch_stencil0<uint16_t>(d, start, num, want_partials, calc_stage);
      }
   }
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
