// exchange.hip -- ghost-plane fills of one exchange direction.
//
// One launch per direction phase; one workgroup per (face job, group of
// up to kVars variables).
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
#include "kern_exchange.h"

namespace {

using namespace amrk;

template <typename T>
__global__ __launch_bounds__(128) void k_exchange(T *__restrict__ pool,
                                                  const amr_face_job *__restrict__ jobs,
                                                  Lay L, int var0, int var_end)
{
   const amr_face_job jb = jobs[blockIdx.x];
   const int v0 = var0 + blockIdx.y * kVars;
   exchange_job<T>(pool, jb, L, v0, min(kVars, var_end - v0), threadIdx.x, blockDim.x);
}

template <typename T>
int launch(amr_dev *d, int dir, int start, int num)
{
   int off = dir == 0 ? 0 : dir == 1 ? d->count[0] : d->count[0] + d->count[1];
   int n = d->count[dir];
   if (!n)
      return AMR_OK;
   dim3 grid(n, (num + kVars - 1) / kVars);
   hipLaunchKernelGGL(k_exchange<T>, grid, dim3(128), 0, d->stream, (T *) d->pool,
                      d->d_jobs + off, d->g, start, start + num);
   HIP_TRY(d, hipGetLastError());
   return AMR_OK;
}

}  // namespace

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
