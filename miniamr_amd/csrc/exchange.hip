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
#include "amr_internal.h"

namespace {

// canonical coordinates of in-plane point (a, b) on plane p of direction dir
__device__ __forceinline__ void canon(int dir, int p, int a, int b, int &i, int &j, int &k)
{
   if (dir == 0) { i = p; j = a; k = b; }
   else if (dir == 1) { i = a; j = p; k = b; }
   else { i = a; j = b; k = p; }
}

constexpr int kVars = 8;   // variables per workgroup, loads issued together

template <typename T>
__global__ __launch_bounds__(128) void k_exchange(T *__restrict__ pool,
                                                  const amr_face_job *__restrict__ jobs,
                                                  Lay L, int var0, int var_end)
{
   typedef typename Acc<T>::type A;
   const amr_face_job jb = jobs[blockIdx.x];
   const int v0 = var0 + blockIdx.y * kVars;
   const int nv = min(kVars, var_end - v0);
   const int dir = jb.face >> 1, hi = jb.face & 1;
   const int n = dir == 0 ? L.nx : dir == 1 ? L.ny : L.nz;
   const int na = dir == 0 ? L.ny : L.nx;   // slower in-plane axis
   const int nb = dir == 2 ? L.ny : L.nz;   // faster in-plane axis
   const int ghost = hi ? n + 1 : 0;
   const int plane = hi ? 1 : n;            // neighbour plane next to the shared face
   T *dst = pool + (int64_t) v0 * L.var_stride + (int64_t) jb.dst * L.rec;
   const T *src = pool + (int64_t) v0 * L.var_stride + (int64_t) jb.src * L.rec;
   const int64_t vs = L.var_stride;

   if (jb.kind == AMR_FACE_SAME || jb.kind == AMR_FACE_BC) {
      int ext_a = 0, ext_b = 0;
      if (L.stencil == 27) {
         if (jb.kind == AMR_FACE_BC) ext_a = ext_b = 1;
         else if (dir == 1) ext_a = 1;
         else if (dir == 2) ext_a = ext_b = 1;
      }
      const int a0 = 1 - ext_a, b0 = 1 - ext_b;
      const int wa = na + 2 * ext_a, wb = nb + 2 * ext_b;
      // BC reads the block's own plane next to the face
      const int sp = jb.kind == AMR_FACE_BC ? (hi ? n : 1) : plane;
      for (int t = threadIdx.x; t < wa * wb; t += blockDim.x) {
         const int a = a0 + t / wb, b = b0 + t % wb;
         int i, j, k, si, sj, sk;
         canon(dir, sp, a, b, si, sj, sk);
         canon(dir, ghost, a, b, i, j, k);
         const int so = dir == 2 ? zcopy(L, si, sj, sk) : home(L, si, sj, sk);
         const int dh = home(L, i, j, k);
         const int dc = (k == 1 || k == L.nz) ? zcopy(L, i, j, k) : -1;
         T v[kVars];
#pragma unroll
         for (int q = 0; q < kVars; q++)
            if (q < nv) v[q] = src[q * vs + so];
#pragma unroll
         for (int q = 0; q < kVars; q++)
            if (q < nv) {
               dst[q * vs + dh] = v[q];
               if (dc >= 0) dst[q * vs + dc] = v[q];
            }
      }
   } else if (jb.kind == AMR_FACE_PROLONG) {
      // fine cell (a, b) of the whole face <- coarse cell of the quarter
      const int ao = jb.jq * (na / 2), bo = jb.iq * (nb / 2);
      for (int t = threadIdx.x; t < na * nb; t += blockDim.x) {
         const int a = 1 + t / nb, b = 1 + t % nb;
         const int ca = ((a + 1) >> 1) + ao, cb = ((b + 1) >> 1) + bo;
         int i, j, k, si, sj, sk;
         canon(dir, plane, ca, cb, si, sj, sk);
         canon(dir, ghost, a, b, i, j, k);
         const int so = dir == 2 ? zcopy(L, si, sj, sk) : home(L, si, sj, sk);
         const int dh = home(L, i, j, k);
         const int dc = (k == 1 || k == L.nz) ? zcopy(L, i, j, k) : -1;
         T v[kVars];
#pragma unroll
         for (int q = 0; q < kVars; q++)
            if (q < nv) v[q] = src[q * vs + so];
#pragma unroll
         for (int q = 0; q < kVars; q++)
            if (q < nv) {
               const T o = from_acc<T>(to_acc(v[q]) / A(4));
               dst[q * vs + dh] = o;
               if (dc >= 0) dst[q * vs + dc] = o;
            }
      }
   } else {  // RESTRICT: coarse quarter cell <- sum of 4 fine cells
      const int ha = na / 2, hb = nb / 2;
      const int ao = jb.jq * ha, bo = jb.iq * hb;
      for (int t = threadIdx.x; t < ha * hb; t += blockDim.x) {
         const int a = 1 + t / hb, b = 1 + t % hb;
         int so[4];
#pragma unroll
         for (int r = 0; r < 4; r++) {
            int si, sj, sk;
            canon(dir, plane, 2 * a - 1 + (r >> 1), 2 * b - 1 + (r & 1), si, sj, sk);
            so[r] = dir == 2 ? zcopy(L, si, sj, sk) : home(L, si, sj, sk);
         }
         int i, j, k;
         canon(dir, ghost, a + ao, b + bo, i, j, k);
         const int dh = home(L, i, j, k);
         const int dc = (k == 1 || k == L.nz) ? zcopy(L, i, j, k) : -1;
         T v[kVars][4];
#pragma unroll
         for (int q = 0; q < kVars; q++)
            if (q < nv) {
#pragma unroll
               for (int r = 0; r < 4; r++)
                  v[q][r] = src[q * vs + so[r]];
            }
#pragma unroll
         for (int q = 0; q < kVars; q++)
            if (q < nv) {
               A s = to_acc(v[q][0]);
               s = s + to_acc(v[q][1]);
               s = s + to_acc(v[q][2]);
               s = s + to_acc(v[q][3]);
               const T o = from_acc<T>(s);
               dst[q * vs + dh] = o;
               if (dc >= 0) dst[q * vs + dc] = o;
            }
      }
   }
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
