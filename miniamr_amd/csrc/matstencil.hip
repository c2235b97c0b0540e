// matstencil.hip -- the variable-work stencils of --stencil 0
// (ref/stencil.c:43-73 stencil_driver, :147-981).
//
// With --stencil 0 the variables form mat = num_vars/4 "materials" in four
// classes: class c holds variables [c*mat, (c+1)*mat).  Variable 0 and the
// variables from 4*mat on take the plain 7-point average (k_stencil).  Every
// other variable takes, by calculation stage % 6, one of six updates coupled
// to other variables, followed by the clamp of stencil_check:
//   0    a pointwise update from variables 0, 1 and class partners;
//   1-3  a sweep along x, y or z that updates in place: each cell reads its
//        predecessor's new value, so one thread owns a line and walks it in
//        order (variable 1 and class 0 are pointwise here);
//   4    a 7-point product stencil with the class partners;
//   5    a 27-point sum of the class partners.
// 4 and 5 read the old values of the whole block (a Jacobi update): they
// write a work buffer that a second kernel commits (clamping as it goes).
// Every expression keeps the reference's evaluation order (left to right,
// no contraction), so the FP64 path is bit-identical; the clamp's
// data-dependent operation counts are summed on the device for the report.
//
// One workgroup per active block; cells are addressed through home(), so
// face, edge and z-copy parts of the record stay consistent.
#include "kern_exchange.h"
#include "kern_stencil.h"

namespace {

using namespace amrk;

// Class partners (offsets in units of mat) of the 7-point update -- x, y, z
// neighbours -- and of the 27-point update -- neighbours with 1, 2 and 3
// non-zero offsets (ref/stencil.c:661-845).
__constant__ int kPartner[4][3] = {{1, 2, 3}, {1, 2, -1}, {1, -2, -1}, {-3, -2, -1}};

struct MatArgs {
   int var, mat;
   double a1;
   const double *a0;          // mat values (device)
   unsigned long long *counts;  // [0] clamps that divide, [1] that scale
};

// One block's cells of every variable.
template <typename T>
struct BlockView {
   T *pool;
   int64_t vs, base;   // variable stride, slot offset
   typedef typename Acc<T>::type A;
   __device__ A get(const Lay &L, int u, int i, int j, int k) const
   {
      return to_acc(pool[u * vs + base + home(L, i, j, k)]);
   }
   // store a new value of variable u; returns the value as stored
   __device__ A put(const Lay &L, int u, int i, int j, int k, A v) const
   {
      const T t = from_acc<T>(v);
      T *rec = pool + u * vs + base;
      rec[home(L, i, j, k)] = t;
      if (k == 1 || k == L.nz)
         rec[zcopy(L, i, j, k)] = t;
      return to_acc(t);
   }
};

template <typename T>
__device__ __forceinline__ BlockView<T> view(T *pool, const int32_t *sorted, const Lay &L)
{
   BlockView<T> b;
   b.pool = pool;
   b.vs = L.var_stride;
   b.base = (int64_t) sorted[blockIdx.x] * L.rec;
   return b;
}

// stencil_check (ref/stencil.c:966-981): |v|, then v / (a1 + a0[0] + v) for
// v >= 1 or v * (10 - a1) for v < 0.1.
template <typename A>
__device__ __forceinline__ A clamp(A v, const MatArgs &m, unsigned &ndiv, unsigned &nmul)
{
   v = fabs(v);
   if (v >= A(1.0)) {
      v = v / ((A(m.a1) + A(m.a0[0])) + v);
      ndiv++;
   } else if (v < A(0.1)) {
      v = v * (A(10.0) - A(m.a1));
      nmul++;
   }
   return v;
}

__device__ __forceinline__ void add_counts(const MatArgs &m, unsigned ndiv, unsigned nmul)
{
   for (int o = 32; o > 0; o >>= 1) {
      ndiv += __shfl_down(ndiv, o, 64);
      nmul += __shfl_down(nmul, o, 64);
   }
   if ((threadIdx.x & 63) == 0 && (ndiv | nmul)) {
      atomicAdd(&m.counts[0], (unsigned long long) ndiv);
      atomicAdd(&m.counts[1], (unsigned long long) nmul);
   }
}

// Pointwise updates: stage 0 (stencil_0, ref/stencil.c:147-225) and, for
// stages 1-3, variable 1 and class 0 (ref/stencil.c:233-262 and twins).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_mat_point(T *pool, const int32_t *sorted, Lay L,
                                                        MatArgs m, int op)
{
   typedef typename Acc<T>::type A;
   const BlockView<T> b = view(pool, sorted, L);
   const int var = m.var, mat = m.mat, ncell = L.nx * L.ny * L.nz;
   const A a1 = A(m.a1), one = A(1.0);
   unsigned ndiv = 0, nmul = 0;
   for (int t = threadIdx.x; t < ncell; t += kThreads) {
      const int i = 1 + t / (L.ny * L.nz), j = 1 + (t / L.nz) % L.ny, k = 1 + t % L.nz;
      const A v = b.get(L, var, i, j, k), p0 = b.get(L, 0, i, j, k), p1 = b.get(L, 1, i, j, k);
      A r;
      if (op == 0) {
         if (var == 1) {
            r = p1;
            for (int u = mat; u < 2 * mat; u++)
               r = r + b.get(L, u, i, j, k) * p0;
         } else if (var < mat) {
            r = v + v * ((p0 + p1) - a1 * v);
         } else if (var < 2 * mat) {
            r = v * (((p0 + v) + a1 * b.get(L, var + mat, i, j, k)) +
                     (one - a1) * b.get(L, var + 2 * mat, i, j, k)) / p1;
         } else if (var < 3 * mat) {
            r = v + b.get(L, var - mat, i, j, k) *
                       ((a1 * p0 + A(m.a0[var - 2 * mat]) * v) +
                        (one - a1) * b.get(L, var + mat, i, j, k)) / p1;
         } else {
            const A a0 = A(m.a0[var - 3 * mat]), q = b.get(L, var - 2 * mat, i, j, k);
            r = v + q * (((a1 * p0 + a0 * v) + (one - a0) * b.get(L, var - mat, i, j, k)) +
                         (one - a1) * q) / (p1 * p1);
         }
      } else if (var == 1) {
         r = p1;
         for (int u = 2; u < mat + 2; u++)
            r = r + b.get(L, u, i, j, k) * p0;
         r = r / (a1 + r);
      } else {   // class 0 of a sweep stage
         r = v + v * ((p0 + p1) - a1 * v) / (A(m.a0[var]) + p1);
      }
      b.put(L, var, i, j, k, clamp(r, m, ndiv, nmul));
   }
   add_counts(m, ndiv, nmul);
}

// In-place sweeps of classes 1-3 along axis ax (stages 1-3: stencil_x,
// stencil_y, stencil_z, ref/stencil.c:263-660).  Thread = line; the
// predecessor's value is the new one, the successor's the old one.  The
// gate variable g (the variable itself for class 1, var - mat for class 2,
// var - 2*mat for class 3) picks the weights.  After the line is done the
// thread clamps it (only its own line's cells are read by the sweep).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_mat_sweep(T *pool, const int32_t *sorted, Lay L,
                                                        MatArgs m, int ax)
{
   typedef typename Acc<T>::type A;
   const BlockView<T> b = view(pool, sorted, L);
   const int var = m.var, mat = m.mat, cls = var / mat;
   const int n = ax == 0 ? L.nx : ax == 1 ? L.ny : L.nz;
   const int n1 = ax == 0 ? L.ny : L.nx;   // slower in-plane axis
   const int n2 = ax == 2 ? L.ny : L.nz;   // faster in-plane axis
   const A a1 = A(m.a1), a0 = A(m.a0[var - cls * mat]);
   const int g = cls == 1 ? var : var - (cls - 1) * mat;
   unsigned ndiv = 0, nmul = 0;
   for (int t = threadIdx.x; t < n1 * n2; t += kThreads) {
      const int c1 = 1 + t / n2, c2 = 1 + t % n2;
      auto at = [&](int u, int p) {
         int i, j, k;
         canon(ax, p, c1, c2, i, j, k);
         return b.get(L, u, i, j, k);
      };
      A prev = at(var, 0);
      for (int p = 1; p <= n; p++) {
         const A v = at(var, p), vn = at(var, p + 1), q1 = at(1, p);
         const A gv = cls == 1 ? v : at(g, p);
         const A gm = cls == 1 ? prev : at(g, p - 1);
         const A gp = cls == 1 ? vn : at(g, p + 1);
         const A tmp1 = fabs(gv - gm), tmp2 = fabs(gv - gp);
         const A w = tmp1 > tmp2 ? tmp1 - tmp2 : tmp2 - tmp1;
         const A num = ((tmp1 * prev) + (w * (v + q1))) + (tmp2 * vn);
         A den;
         if (cls == 1) {
            den = ((((a1 + a0) + prev) + v) + vn) + at(0, p);
            den = den + q1;
         } else {
            den = ((a1 + a0) + at(var - mat, p)) + at(cls == 2 ? var + mat : var - 2 * mat, p);
            den = ((den + prev) + v) + vn;
         }
         int i, j, k;
         canon(ax, p, c1, c2, i, j, k);
         prev = b.put(L, var, i, j, k, num / den);
      }
      for (int p = 1; p <= n; p++) {
         int i, j, k;
         canon(ax, p, c1, c2, i, j, k);
         b.put(L, var, i, j, k, clamp(b.get(L, var, i, j, k), m, ndiv, nmul));
      }
   }
   add_counts(m, ndiv, nmul);
}

// Jacobi updates, stages 4 and 5 (stencil_7 / stencil_27, ref/stencil.c:661-958):
// new values into work[in][cell] (accumulator type), old values everywhere.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_mat_jacobi(T *pool, const int32_t *sorted, Lay L,
                                                         MatArgs m, int op,
                                                         typename Acc<T>::type *work)
{
   typedef typename Acc<T>::type A;
   const BlockView<T> b = view(pool, sorted, L);
   const int var = m.var, mat = m.mat, cls = var / mat, ncell = L.nx * L.ny * L.nz;
   const int qx = var + kPartner[cls][0] * mat, qy = var + kPartner[cls][1] * mat,
             qz = var + kPartner[cls][2] * mat;
   const A a1 = A(m.a1);
   A *out = work + (int64_t) blockIdx.x * ncell;
   for (int t = threadIdx.x; t < ncell; t += kThreads) {
      const int i = 1 + t / (L.ny * L.nz), j = 1 + (t / L.nz) % L.ny, k = 1 + t % L.nz;
      const A v = b.get(L, var, i, j, k);
      A r;
      if (op == 4) {
         A s = b.get(L, var, i - 1, j, k) * b.get(L, qx, i - 1, j, k);
         s = s + b.get(L, var, i, j - 1, k) * b.get(L, qy, i, j - 1, k);
         s = s + b.get(L, var, i, j, k - 1) * b.get(L, qz, i, j, k - 1);
         s = s + v * v;
         s = s + b.get(L, var, i, j, k + 1) * b.get(L, qz, i, j, k + 1);
         s = s + b.get(L, var, i, j + 1, k) * b.get(L, qy, i, j + 1, k);
         s = s + b.get(L, var, i + 1, j, k) * b.get(L, qx, i + 1, j, k);
         r = s / A(7.0) * (a1 + v);
      } else {
         const int q[4] = {var, qx, qy, qz};   // by number of non-zero offsets
         A s = A(0);
         for (int d = 0; d < 27; d++) {
            const int di = d / 9 - 1, dj = (d / 3) % 3 - 1, dk = d % 3 - 1;
            const A e = b.get(L, q[(di != 0) + (dj != 0) + (dk != 0)], i + di, j + dj, k + dk);
            s = d ? s + e : e;
         }
         r = s / (a1 + A(27.0));
      }
      out[t] = r;
   }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_mat_commit(T *pool, const int32_t *sorted, Lay L,
                                                         MatArgs m,
                                                         const typename Acc<T>::type *work)
{
   const BlockView<T> b = view(pool, sorted, L);
   const int ncell = L.nx * L.ny * L.nz;
   const typename Acc<T>::type *in = work + (int64_t) blockIdx.x * ncell;
   unsigned ndiv = 0, nmul = 0;
   for (int t = threadIdx.x; t < ncell; t += kThreads) {
      const int i = 1 + t / (L.ny * L.nz), j = 1 + (t / L.nz) % L.ny, k = 1 + t % L.nz;
      b.put(L, m.var, i, j, k, clamp(in[t], m, ndiv, nmul));
   }
   add_counts(m, ndiv, nmul);
}

template <typename T>
int launch_mat(amr_dev *d, int var, int calc_stage)
{
   typedef typename Acc<T>::type A;
   Lay L;
   if (int rc = amr_lay(d, 0, d->c.num_vars, &L))
      return rc;
   MatArgs m;
   m.var = var;
   m.mat = d->mat;
   m.a1 = d->a1;
   m.a0 = d->d_a0;
   m.counts = d->d_check_counts;
   const dim3 grid(d->n_active), blk(kThreads);
   const int op = calc_stage % 6, cls = var / d->mat;
   if (op == 0 || (op <= 3 && (var == 1 || cls == 0))) {
      hipLaunchKernelGGL(k_mat_point<T>, grid, blk, 0, d->stream, (T *) d->pool, d->d_sorted, L,
                         m, op);
   } else if (op <= 3) {
      hipLaunchKernelGGL(k_mat_sweep<T>, grid, blk, 0, d->stream, (T *) d->pool, d->d_sorted, L,
                         m, op - 1);
   } else {
      const size_t need = (size_t) d->n_active * d->g.nx * d->g.ny * d->g.nz * sizeof(A);
      if (need > d->matwork_bytes) {
         HIP_TRY(d, hipStreamSynchronize(d->stream));
         hipFree(d->d_matwork);
         d->d_matwork = nullptr;
         d->matwork_bytes = 0;
         HIP_TRY(d, hipMalloc(&d->d_matwork, need));
         d->matwork_bytes = need;
      }
      hipLaunchKernelGGL(k_mat_jacobi<T>, grid, blk, 0, d->stream, (T *) d->pool, d->d_sorted, L,
                         m, op, (A *) d->d_matwork);
      hipLaunchKernelGGL(k_mat_commit<T>, grid, blk, 0, d->stream, (T *) d->pool, d->d_sorted, L,
                         m, (const A *) d->d_matwork);
   }
   HIP_TRY(d, hipGetLastError());
   return AMR_OK;
}

}  // namespace

// One variable's update of a --stencil 0 stage other than the plain 7-point
// ones (variable 0 and variables >= 4*mat are left to the caller).
int amr_launch_mat_stencil(amr_dev *d, int var, int calc_stage)
{
   if (!d->n_active)
      return AMR_OK;
   if (!d->d_a0)
      return amr_set_error(d, AMR_ERR_STATE, "stencil 0: amr_dev_set_materials not called");
   switch (d->c.precision) {
   case AMR_PREC_F64: return launch_mat<double>(d, var, calc_stage);
   case AMR_PREC_F32: return launch_mat<float>(d, var, calc_stage);
   default: return launch_mat<uint16_t>(d, var, calc_stage);
   }
}
