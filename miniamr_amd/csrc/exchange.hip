// exchange.hip -- ghost-plane fills of one exchange direction.
//
// One launch per direction phase; one workgroup per (face job, group of
// up to 8 variables).
// Within a phase every fill is a gather from cells no fill of that phase
// writes (face-interior planes, and for the 27-point stencil the ghosts of
// directions done earlier), and no two fills write the same cell, so the
// jobs of a phase run fully in parallel; the phases are ordered by the
// stream.  Arithmetic per kind, reference code and operation order:
//   SAME      copy                                   ref/comm.c:1473-1534
//   PROLONG   fine ghost = coarse / 4                ref/comm.c:1618-1625
//   RESTRICT  coarse ghost = f00 + f01 + f10 + f11   ref/comm.c:1626-1630
//             (slower in-plane index first, as the reference sums them)
//   BC        ghost plane = adjacent interior plane  ref/comm.c:1911-1965
// In-plane extents (code 0): x faces always interior; for the 27-point
// stencil y faces extend along x and z faces along x and y so that edges
// and corners fill transitively (ref/comm.c:1496-1533), and boundary
// planes cover the whole ghost plane (ref/comm.c:1941-1964).
//
// Cells are addressed canonically (i, j, k) and mapped to the record
// layout (amr_internal.h); the z phase reads source planes z = 1 / z = nz
// from their contiguous copies.  A workgroup fills one face job for up to
// kVars variables, issuing all of their loads before the stores.
#include <stdio.h>
#include <stdlib.h>

#include "kern_exchange.h"

namespace {

using namespace amrk;

template <typename T, int KV, int NT>
__global__ __launch_bounds__(NT) void k_exchange(T *__restrict__ pool,
                                                 const amr_face_job *__restrict__ jobs,
                                                 Lay L, int var0, int var_end)
{
   const amr_face_job jb = jobs[blockIdx.x];
   const int v0 = var0 + blockIdx.y * KV;
   exchange_job<T, KV>(pool, jb, L, v0, min(KV, var_end - v0), threadIdx.x, NT);
}

// The fused stage's jobs (FusedJob), same launch shape.
template <typename T, int KV, int NT>
__global__ __launch_bounds__(NT) void k_exchange_fused(T *__restrict__ pool,
                                                       const FusedJob *__restrict__ jobs,
                                                       const EdgeSrc *__restrict__ edges, Lay L,
                                                       int var0, int var_end)
{
   const FusedJob fj = jobs[blockIdx.x];
   const int v0 = var0 + blockIdx.y * KV;
   exchange_job_fused<T, KV>(pool, fj, edges, L, v0, min(KV, var_end - v0), threadIdx.x, NT);
}

// Variables and threads per workgroup.  Default 8 x 256 (measured best for
// FP64 and FP32 on the 16^3 workload: 8 x 128 was 3% slower in FP32, 16
// variables 1-3% slower).  Fused stages, whose jobs are mostly coarse/fine
// faces: 8 variables in FP64 (16 were 6% slower, 4 11% slower), 20 in
// FP32 (exchange 5% shorter than with 16; 8 and 4 slower) and 40 in BF16
// (exchange 17% shorter than with 8; 16 and 20 in between): the narrower
// the type, the more variables a workgroup's loads need to fill the
// memory pipeline.  MINIAMR_EXCHANGE ("vars,threads" with vars
// 4/8/16/20/40 and threads 128/256) overrides it (fused stages take the
// variable count, at 256 threads; plain stages 16 for 20 and 40).
struct ExShape {
   int kv = kVars, nt = 256;
   bool chosen = false;   // set by MINIAMR_EXCHANGE
};

ExShape ex_shape()
{
   static ExShape sh;
   static bool read = false;
   if (!read) {
      read = true;
      const char *e = getenv("MINIAMR_EXCHANGE");
      int kv = 0, nt = 0;
      if (e && sscanf(e, "%d,%d", &kv, &nt) == 2 && (kv == 4 || kv == 8 || kv == 16 || kv == 20 || kv == 40) &&
          (nt == 128 || nt == 256)) {
         sh.kv = kv;
         sh.nt = nt;
         sh.chosen = true;
      }
   }
   return sh;
}

template <typename T, int KV, int NT>
int launch_kv(amr_dev *d, int off, int n, int start, int num)
{
   dim3 grid(n, (num + KV - 1) / KV);
   Lay L;
   if (int rc = amr_lay(d, start, num, &L))
      return rc;
   hipLaunchKernelGGL((k_exchange<T, KV, NT>), grid, dim3(NT), 0, d->stream, (T *) d->pool,
                      d->d_jobs + off, L, start, start + num);
   HIP_TRY(d, hipGetLastError());
   return AMR_OK;
}

template <typename T>
int launch(amr_dev *d, int dir, int start, int num)
{
   int off = dir == 0 ? 0 : dir == 1 ? d->count[0] : d->count[0] + d->count[1];
   int n = d->count[dir];
   if (!n)
      return AMR_OK;
   const ExShape sh = ex_shape();
   if (sh.nt == 256)
      return sh.kv >= 16 ? launch_kv<T, 16, 256>(d, off, n, start, num)
           : sh.kv == 4 ? launch_kv<T, 4, 256>(d, off, n, start, num)
                        : launch_kv<T, 8, 256>(d, off, n, start, num);
   return sh.kv >= 16 ? launch_kv<T, 16, 128>(d, off, n, start, num)
        : sh.kv == 4 ? launch_kv<T, 4, 128>(d, off, n, start, num)
                     : launch_kv<T, 8, 128>(d, off, n, start, num);
}

template <typename T>
int launch_fused(amr_dev *d, int dir, int start, int num)
{
   const int off = dir == 0 ? 0 : dir == 1 ? d->fcount[0] : d->fcount[0] + d->fcount[1];
   const int n = d->fcount[dir];
   if (!n || !num)
      return AMR_OK;
   Lay L;
   if (int rc = amr_lay(d, start, num, &L))
      return rc;
   // variables per item by storage type (see ExShape), unless
   // MINIAMR_EXCHANGE chooses
   const ExShape sh = ex_shape();
   const int kv = sh.chosen ? sh.kv : sizeof(T) == 4 ? 20 : sizeof(T) == 2 ? 40 : kVars;
   if (kv == 20) {
      dim3 grid(n, (num + 19) / 20);
      hipLaunchKernelGGL((k_exchange_fused<T, 20, 256>), grid, dim3(256), 0, d->stream,
                         (T *) d->pool, d->d_fjobs + off, d->d_edges, L, start,
                         start + num);
   } else if (kv == 40) {
      dim3 grid(n, (num + 39) / 40);
      hipLaunchKernelGGL((k_exchange_fused<T, 40, 256>), grid, dim3(256), 0, d->stream,
                         (T *) d->pool, d->d_fjobs + off, d->d_edges, L, start,
                         start + num);
   } else if (kv == 16) {
      dim3 grid(n, (num + 15) / 16);
      hipLaunchKernelGGL((k_exchange_fused<T, 16, 256>), grid, dim3(256), 0, d->stream,
                         (T *) d->pool, d->d_fjobs + off, d->d_edges, L, start,
                         start + num);
   } else if (kv == 4) {
      dim3 grid(n, (num + 3) / 4);
      hipLaunchKernelGGL((k_exchange_fused<T, 4, 256>), grid, dim3(256), 0, d->stream,
                         (T *) d->pool, d->d_fjobs + off, d->d_edges, L, start,
                         start + num);
   } else {
      dim3 grid(n, (num + kVars - 1) / kVars);
      hipLaunchKernelGGL((k_exchange_fused<T, kVars, 256>), grid, dim3(256), 0, d->stream,
                         (T *) d->pool, d->d_fjobs + off, d->d_edges, L, start,
                         start + num);
   }
   HIP_TRY(d, hipGetLastError());
   return AMR_OK;
}

}  // namespace

int amr_launch_exchange_fused_dir(amr_dev *d, int dir, int start, int num)
{
   switch (d->c.precision) {
   case AMR_PREC_F64: return launch_fused<double>(d, dir, start, num);
   case AMR_PREC_F32: return launch_fused<float>(d, dir, start, num);
   default: return launch_fused<uint16_t>(d, dir, start, num);
   }
}

namespace {

template <typename T>
int launch_jobs(amr_dev *d, const amr_face_job *jobs, int n, const Lay &L, int start, int num)
{
   dim3 grid(n, (num + kVars - 1) / kVars);
   hipLaunchKernelGGL((k_exchange<T, kVars, 256>), grid, dim3(256), 0, d->stream, (T *) d->pool,
                      jobs, L, start, start + num);
   HIP_TRY(d, hipGetLastError());
   return AMR_OK;
}

}  // namespace

int amr_launch_exchange_jobs(amr_dev *d, const amr_face_job *jobs, int n, const Lay &L,
                             int start, int num)
{
   if (!n || !num)
      return AMR_OK;
   switch (d->c.precision) {
   case AMR_PREC_F64: return launch_jobs<double>(d, jobs, n, L, start, num);
   case AMR_PREC_F32: return launch_jobs<float>(d, jobs, n, L, start, num);
   default: return launch_jobs<uint16_t>(d, jobs, n, L, start, num);
   }
}

int amr_launch_exchange_dir(amr_dev *d, int dir, int start, int num)
{
   if (!d->d_jobs && (d->count[0] + d->count[1] + d->count[2]))
      return amr_set_error(d, AMR_ERR_STATE, "exchange before set_mesh");
   switch (d->c.precision) {
   case AMR_PREC_F64: return launch<double>(d, dir, start, num);
   case AMR_PREC_F32: return launch<float>(d, dir, start, num);
   default: return launch<uint16_t>(d, dir, start, num);
   }
}
