// exchange.hip -- ghost-plane fills of one exchange direction.
//
// One launch per direction phase; one workgroup per (face job, variable).
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
#include "amr_internal.h"

namespace {

struct FaceFrame {
   int64_t sp, sa, sb;   // strides of the normal axis and the two in-plane axes
   int n, na, nb;        // sizes along those axes (interior)
};

__device__ __forceinline__ FaceFrame frame(const Geo &g, int dir)
{
   FaceFrame f;
   const int64_t sx = (int64_t) g.gy * g.gz, sy = g.gz, sz = 1;
   if (dir == 0) {        // normal x, in-plane (y, z)
      f.sp = sx; f.sa = sy; f.sb = sz; f.n = g.nx; f.na = g.ny; f.nb = g.nz;
   } else if (dir == 1) { // normal y, in-plane (x, z)
      f.sp = sy; f.sa = sx; f.sb = sz; f.n = g.ny; f.na = g.nx; f.nb = g.nz;
   } else {               // normal z, in-plane (x, y)
      f.sp = sz; f.sa = sx; f.sb = sy; f.n = g.nz; f.na = g.nx; f.nb = g.ny;
   }
   return f;
}

template <typename T>
__global__ __launch_bounds__(128) void k_exchange(T *__restrict__ pool,
                                                  const amr_face_job *__restrict__ jobs,
                                                  Geo g, int var0)
{
   typedef typename Acc<T>::type A;
   const amr_face_job j = jobs[blockIdx.x];
   const int var = var0 + blockIdx.y;
   const int dir = j.face >> 1, hi = j.face & 1;
   const FaceFrame f = frame(g, dir);
   const int ghost = hi ? f.n + 1 : 0;
   const int plane = hi ? 1 : f.n;      // neighbour plane next to the shared face
   T *dst = pool + (int64_t) var * g.var_stride + (int64_t) j.dst * g.slab + ghost * f.sp;
   const T *src = pool + (int64_t) var * g.var_stride + (int64_t) j.src * g.slab;

   if (j.kind == AMR_FACE_SAME || j.kind == AMR_FACE_BC) {
      int ext_a = 0, ext_b = 0;
      if (g.stencil == 27) {
         if (j.kind == AMR_FACE_BC) ext_a = ext_b = 1;
         else if (dir == 1) ext_a = 1;
         else if (dir == 2) ext_a = ext_b = 1;
      }
      const int a0 = 1 - ext_a, b0 = 1 - ext_b;
      const int wa = f.na + 2 * ext_a, wb = f.nb + 2 * ext_b;
      // BC reads the block's own plane next to the face
      const int sp = j.kind == AMR_FACE_BC ? (hi ? f.n : 1) : plane;
      const T *s = src + sp * f.sp;
      for (int t = threadIdx.x; t < wa * wb; t += blockDim.x) {
         const int a = a0 + t / wb, b = b0 + t % wb;
         const int64_t o = a * f.sa + b * f.sb;
         dst[o] = s[o];
      }
   } else if (j.kind == AMR_FACE_PROLONG) {
      // fine cell (a, b) of the whole face <- coarse cell of the quarter
      const int ao = j.jq * (f.na / 2), bo = j.iq * (f.nb / 2);
      const T *s = src + plane * f.sp;
      for (int t = threadIdx.x; t < f.na * f.nb; t += blockDim.x) {
         const int a = 1 + t / f.nb, b = 1 + t % f.nb;
         const int ca = ((a + 1) >> 1) + ao, cb = ((b + 1) >> 1) + bo;
         A v = to_acc(s[ca * f.sa + cb * f.sb]) / A(4);
         dst[a * f.sa + b * f.sb] = from_acc<T>(v);
      }
   } else {  // RESTRICT: coarse quarter cell <- sum of 4 fine cells
      const int ha = f.na / 2, hb = f.nb / 2;
      const int ao = j.jq * ha, bo = j.iq * hb;
      const T *s = src + plane * f.sp;
      for (int t = threadIdx.x; t < ha * hb; t += blockDim.x) {
         const int a = 1 + t / hb, b = 1 + t % hb;
         const int64_t o00 = (2 * a - 1) * f.sa + (2 * b - 1) * f.sb;
         A v = to_acc(s[o00]);
         v = v + to_acc(s[o00 + f.sb]);
         v = v + to_acc(s[o00 + f.sa]);
         v = v + to_acc(s[o00 + f.sa + f.sb]);
         dst[(a + ao) * f.sa + (b + bo) * f.sb] = from_acc<T>(v);
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
   dim3 grid(n, num);
   hipLaunchKernelGGL(k_exchange<T>, grid, dim3(128), 0, d->stream, (T *) d->pool,
                      d->d_jobs + off, d->g, start);
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
