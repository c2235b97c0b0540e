// batch.hip -- many independent instances stepped together
// (include/amr_batch.h).
//
// Work lists.  Every member is an ordinary context (its own pool, layout,
// sorted slot list, face jobs, partials and checksum table).  For each
// kernel of a stage the batch keeps a flat list of work items over all
// active members, stored implicitly: members are ordered by block shape
// (largest first) and a prefix table pre[m] gives the first item of member
// m, so item t belongs to the member m with pre[m] <= t < pre[m+1] (binary
// search on a wave-uniform value) and its local index addresses the
// member's own data:
//   exchange, direction d: (var group g, face job j), local = g*count[d] + j
//   stencil:               (var group g, sorted block in), local = g*n_active + in
//   checksum reduce:       variable v, local = v
// Once per rebuild the lists are expanded into flat item records (the job
// or block, the member, the first variable), so a stage's kernels load one
// record per item.  Exchange items are computed by a wavefront or a
// workgroup (64/128/256 threads), stencil items by a wavefront (small
// blocks, whose planes would leave most of a workgroup idle) or a 256-thread
// workgroup, one stencil kernel per block-size class; with a grid of `wgs`
// workgroups the kernels are persistent and stride through the list.  One
// launch per stage phase (per class for the stencil) serves every member.
//
// Arithmetic.  Items run exactly the bodies of the single-instance kernels
// (kern_stencil.h, kern_exchange.h), so cells are bit-identical to running
// each member alone; only the association of the per-block checksum
// partials and of their sum differs (FP64, a few ulp).
//
// Graphs.  amr_batch_timestep captures a timestep's stages (3 exchange
// launches, a stencil launch and, on checksum stages, a reduce launch per
// stage) into one hipGraph, reused until the work lists or launch shapes
// change.  No memset or copy node is recorded.
//
// Autotuner.  amr_batch_autotune snapshots the members' pools, times
// candidate shapes (variables per stencil item; persistent grid sizes) on
// the live meshes with HIP events, keeps the fastest, and restores the
// pools, so tuning never changes results.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <vector>

#include "amr_batch.h"
#include "kern_exchange.h"
#include "kern_stencil.h"

namespace {

using namespace amrk;

constexpr int kWavesPerWg = 4;   // workgroups of 256 threads, one item per wave

// One member as the kernels see it.
struct BMember {
   void *pool;
   const amr_face_job *jobs;
   const int32_t *sorted;
   double *partials;
   double *sums;
   Lay L;                   // the interior copy holding the cells
   Lay Lo;                  // the other copy (fused stages write it)
   const FusedJob *fjobs;   // the member's fused plan (amr_dev_stage_fused)
   const int32_t *redir;
   const EdgeSrc *edges;
   int n_active, num_vars, part_stride, cls;
   int count[3], joff[3];
   int fcount[3], fjoff[3];
};

// Shape classes of the wave-per-item stencil: columns (ny*nz) per lane and
// plane elements ((ny+2)*(nz+2)) per lane over 64 lanes, each with a
// compile-time nz for the block size the class is meant for:
//   0/1: ny*nz <= 64,  plane <= 128   (nz == 8 / other; e.g. 8^3)
//   2/3: ny*nz <= 192, plane <= 256   (nz == 12 / other; e.g. 10^3, 12^3)
//   4/5: ny*nz <= 256, plane <= 384   (nz == 16 / other; e.g. 14^3, 16^3)
int shape_class(const Lay &L)
{
   const int ncol = L.ny * L.nz, plane = L.gy * L.gz;
   if (ncol <= 64 && plane <= 128) return L.nz == 8 ? 0 : 1;
   if (ncol <= 192 && plane <= 256) return L.nz == 12 ? 2 : 3;
   if (ncol <= 256 && plane <= 384) return L.nz == 16 ? 4 : 5;
   return -1;
}

__device__ __forceinline__ int find_member(const int *__restrict__ pre, int M, int t)
{
   int lo = 0, hi = M;   // pre[lo] <= t < pre[hi]
   while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (pre[mid] <= t) lo = mid;
      else hi = mid;
   }
   return lo;
}

__device__ __forceinline__ int first_item()
{
   return __builtin_amdgcn_readfirstlane((int) (blockIdx.x * kWavesPerWg + (threadIdx.x >> 6)));
}

// Materialised work items.  The prefix tables define the lists; once per
// rebuild a small kernel expands them into flat records, so that a stage's
// kernels reach their data with one record load per item instead of a
// binary search and a chain of dependent descriptor loads.
struct BExItem {          // exchange: one face job, a group of variables
   amr_face_job jb;
   int32_t m, v0;
};
struct BStItem {          // stencil: one block, a group of variables
   void *rec;             // the block's record of the first variable
   double *part;          // its checksum partial of the first variable
   int32_t m, nv;         // member, variables in the group
   int32_t in, slot;      // the block's place in the sorted list, its slot
};
struct BFxItem {          // fused-stage exchange: one fused job, a group of variables
   FusedJob fj;
   int32_t m, v0;
};

__global__ void k_mat_exchange(const BMember *__restrict__ mem, const int *__restrict__ pre,
                               int M, int dir, int kv, BExItem *__restrict__ out)
{
   const int total = pre[M];
   for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
      const int m = find_member(pre, M, t);
      const BMember &b = mem[m];
      const int local = t - pre[m], cnt = b.count[dir];
      const int g = local / cnt, j = local - g * cnt;
      BExItem it;
      it.jb = b.jobs[b.joff[dir] + j];
      it.m = m;
      it.v0 = g * kv;
      out[t] = it;
   }
}

__global__ void k_mat_stencil(const BMember *__restrict__ mem, const int *__restrict__ pre,
                              int M, int vpw, int elem_bytes, BStItem *__restrict__ out)
{
   const int total = pre[M];
   for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
      const int m = find_member(pre, M, t);
      const BMember &b = mem[m];
      const int local = t - pre[m], na = b.n_active;
      const int g = local / na, in = local - g * na;
      const int v0 = g * vpw;
      BStItem it;
      it.rec = (char *) b.pool +
               ((int64_t) v0 * b.L.var_stride + (int64_t) b.sorted[in] * b.L.rec) * elem_bytes;
      it.part = b.partials + (int64_t) v0 * b.part_stride + in;
      it.m = m;
      it.nv = min(vpw, b.num_vars - v0);
      it.in = in;
      it.slot = b.sorted[in];
      out[t] = it;
   }
}

__global__ void k_mat_fexchange(const BMember *__restrict__ mem, const int *__restrict__ pre,
                                int M, int dir, int kv, BFxItem *__restrict__ out)
{
   const int total = pre[M];
   for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
      const int m = find_member(pre, M, t);
      const BMember &b = mem[m];
      const int local = t - pre[m], cnt = b.fcount[dir];
      const int g = local / cnt, j = local - g * cnt;
      BFxItem it;
      it.fj = b.fjobs[b.fjoff[dir] + j];
      it.m = m;
      it.v0 = g * kv;
      out[t] = it;
   }
}

// Exchange items computed by NT threads each: a wavefront (NT = 64, four
// items per workgroup) or a whole workgroup of NT threads.
template <typename T, int KV, int NT>
__global__ __launch_bounds__(NT == 64 ? 256 : NT) void kb_exchange(
   const BMember *__restrict__ mem, const BExItem *__restrict__ items, int total)
{
   const int tid = NT == 64 ? (int) (threadIdx.x & 63) : (int) threadIdx.x;
   const int per = NT == 64 ? kWavesPerWg : 1;
   const int first = NT == 64 ? first_item() : (int) blockIdx.x;
   for (int t = first; t < total; t += gridDim.x * per) {
      const BExItem it = items[t];
      const BMember &b = mem[it.m];
      exchange_job<T, KV>((T *) b.pool, it.jb, b.L, it.v0, min(KV, b.num_vars - it.v0), tid, NT);
   }
}

// The same over the members' fused plans (a fused stage): `flip` says
// which interior copy holds the cells (0: BMember::L, 1: BMember::Lo).
template <typename T, int KV, int NT>
__global__ __launch_bounds__(NT == 64 ? 256 : NT) void kb_exchange_fused(
   const BMember *__restrict__ mem, const BFxItem *__restrict__ items, int total, int flip)
{
   const int tid = NT == 64 ? (int) (threadIdx.x & 63) : (int) threadIdx.x;
   const int per = NT == 64 ? kWavesPerWg : 1;
   const int first = NT == 64 ? first_item() : (int) blockIdx.x;
   for (int t = first; t < total; t += gridDim.x * per) {
      const BFxItem it = items[t];
      const BMember &b = mem[it.m];
      const Lay L = flip ? b.Lo : b.L;
      exchange_job_fused<T, KV>((T *) b.pool, it.fj, b.edges, L, it.v0,
                                min(KV, b.num_vars - it.v0), tid, NT);
   }
}

// One stencil kernel per block-size class (the register needs of the
// classes differ 2.5x, and a kernel holding every class's code would run
// all of them at the largest one's occupancy).  Each walks the class's
// items over all members: in wavefront mode one (block, variable group)
// item per wave, in workgroup mode one per 256-thread workgroup (the
// single-instance kernel's shape, for large blocks).
template <int C> struct Cls;
template <> struct Cls<0> { enum { NC = 1, NP = 2, NZ = 8, D = 8 }; };
template <> struct Cls<1> { enum { NC = 1, NP = 2, NZ = 0, D = 8 }; };
template <> struct Cls<2> { enum { NC = 3, NP = 4, NZ = 12, D = 4 }; };
template <> struct Cls<3> { enum { NC = 3, NP = 4, NZ = 0, D = 4 }; };
template <> struct Cls<4> { enum { NC = 4, NP = 6, NZ = 16, D = 4 }; };
template <> struct Cls<5> { enum { NC = 4, NP = 6, NZ = 0, D = 4 }; };

// An item's layout is its member's (a scalar load: the member index is
// uniform across the item's threads).
// FUSED: the stencil of a fused stage, reading the copy `flip` names and
// same-level faces in place, writing the other copy (as k_stencil_fused).
template <typename T, int ST, int VPW, int C, bool WGM, bool FUSED = false>
__global__ __launch_bounds__(256) void kb_stencil(const BMember *__restrict__ mem,
                                                  const BStItem *__restrict__ items, int total,
                                                  int want_partials, int ring_elems, int flip)
{
   typedef typename Acc<T>::type A;
   typedef Cls<C> K;
   extern __shared__ __align__(16) unsigned char smem[];
   if constexpr (WGM) {
      __shared__ double red[kThreads / 64];
      for (int t = blockIdx.x; t < total; t += gridDim.x) {
         const BStItem it = items[t];
         const BMember &b = mem[it.m];
         if constexpr (FUSED) {
            const Lay Lr = flip ? b.Lo : b.L, Lw = flip ? b.L : b.Lo;
            stencil_block<T, ST, 1, 2, VPW, kThreads, K::NZ, kDepth, true>(
               (T *) it.rec, Lr, it.nv, (A *) smem, red, want_partials ? it.part : nullptr,
               b.part_stride, &Lw, it.slot, b.redir + 6 * (int64_t) it.in,
               b.edges + (int64_t) kEdgeSrcs * it.slot);
         } else {
            stencil_block<T, ST, 1, 2, VPW, kThreads, K::NZ>((T *) it.rec, b.L, it.nv, (A *) smem,
                                                             red, want_partials ? it.part : nullptr,
                                                             b.part_stride);
         }
      }
   } else {
      A *ring = (A *) smem + (threadIdx.x >> 6) * ring_elems;
      const int stride = gridDim.x * kWavesPerWg;
      for (int t = first_item(); t < total; t += stride) {
         const BStItem it = items[t];
         const BMember &b = mem[it.m];
         if constexpr (FUSED) {
            const Lay Lr = flip ? b.Lo : b.L, Lw = flip ? b.L : b.Lo;
            stencil_block<T, ST, K::NC, K::NP, VPW, 64, K::NZ, K::D, true>(
               (T *) it.rec, Lr, it.nv, ring, nullptr, want_partials ? it.part : nullptr,
               b.part_stride, &Lw, it.slot, b.redir + 6 * (int64_t) it.in,
               b.edges + (int64_t) kEdgeSrcs * it.slot);
         } else {
            stencil_block<T, ST, K::NC, K::NP, VPW, 64, K::NZ, K::D>(
               (T *) it.rec, b.L, it.nv, ring, nullptr, want_partials ? it.part : nullptr,
               b.part_stride);
         }
      }
   }
}

// Checksum of variable v of a member: the per-block partials summed in
// sorted-block order per lane (blocks lane, lane+64, ...), then a fixed
// shuffle tree.
__global__ __launch_bounds__(256) void kb_reduce(const BMember *__restrict__ mem,
                                                 const int *__restrict__ pre, int M, int row)
{
   const int lane = threadIdx.x & 63;
   const int total = pre[M], stride = gridDim.x * kWavesPerWg;
   for (int t = first_item(); t < total; t += stride) {
      const int m = find_member(pre, M, t);
      const BMember &b = mem[m];
      const int v = t - pre[m];
      const double *p = b.partials + (int64_t) v * b.part_stride;
      double s = 0.0;
      for (int i = lane; i < b.n_active; i += 64)
         s += p[i];
      s = wave_reduce(s);
      if (lane == 0)
         b.sums[(int64_t) row * b.num_vars + v] = s;
   }
}

}  // namespace

struct amr_batch {
   int device = 0, precision = 0, stencil = 7, cus = 256;
   hipStream_t stream = nullptr;
   std::vector<amr_dev *> members;
   std::vector<char> active;
   std::vector<int> seen;            // plan_version of each member at the last build

   // work lists
   BMember *d_mem = nullptr;
   int *d_pre = nullptr;             // ex[3][M+1], st[6][M+1] (per class), rd[M+1]
   int mem_cap = 0;
   int M = 0, plane_max = 0;
   int tot_ex[3] = {0, 0, 0}, tot_st[6] = {0, 0, 0, 0, 0, 0}, tot_rd = 0;
   int tot_fx[3] = {0, 0, 0};
   bool fusable = false;             // every member has a fused plan (two interiors)
   int plane_max_c[6] = {0, 0, 0, 0, 0, 0};
   bool table_ok = false;
   int table_vpw = 0, table_wg = -1, table_exv = 0;
   long epoch = 0;
   std::vector<int> member_rows;     // sum_rows of each table member

   // launch shape
   int vpw = 1, st_wgs = 0, ex_wgs = 0;
   int wg_mask = 0;                  // bit c: size class c in workgroup mode
   int ex_vars = kVars;              // variables per exchange item (4, 8 or 16)
   int ex_threads = 128;             // threads per exchange item (64, 128 or 256)

   // materialised items: exchange per direction, stencil per class
   BExItem *d_exitems = nullptr;
   BFxItem *d_fxitems = nullptr;     // fused-stage exchange items per direction
   size_t fx_cap = 0, fx_off[3] = {0, 0, 0};
   BStItem *d_stitems = nullptr;
   size_t ex_cap = 0, st_cap = 0;
   size_t ex_off[3] = {0, 0, 0}, st_off[6] = {0, 0, 0, 0, 0, 0};

   // graph
   hipGraphExec_t gexec = nullptr;
   long gkey[11] = {0};

   // timing
   bool timing = false;
   struct Span { hipEvent_t a, b; int cat; };
   std::vector<Span> spans;
   std::vector<hipEvent_t> ev_free;

   char err[512] = {0};
};

namespace {

int berr(amr_batch *b, int rc, const char *fmt, ...)
{
   if (b) {
      va_list ap;
      va_start(ap, fmt);
      vsnprintf(b->err, sizeof(b->err), fmt, ap);
      va_end(ap);
   }
   return rc;
}

#define BHIP(b, call)                                                                   \
   do {                                                                                 \
      hipError_t e_ = (call);                                                           \
      if (e_ != hipSuccess)                                                             \
         return berr((b), e_ == hipErrorOutOfMemory ? AMR_ERR_NOMEM : AMR_ERR_HIP,      \
                     "%s: %s", #call, hipGetErrorString(e_));                           \
   } while (0)

hipEvent_t bevent(amr_batch *b)
{
   if (!b->ev_free.empty()) {
      hipEvent_t e = b->ev_free.back();
      b->ev_free.pop_back();
      return e;
   }
   hipEvent_t e = nullptr;
   return hipEventCreate(&e) == hipSuccess ? e : nullptr;
}

void span_begin(amr_batch *b, int cat)
{
   if (!b->timing)
      return;
   amr_batch::Span s{bevent(b), nullptr, cat};
   if (s.a)
      hipEventRecord(s.a, b->stream);
   b->spans.push_back(s);
}

void span_end(amr_batch *b)
{
   if (!b->timing || b->spans.empty())
      return;
   amr_batch::Span &s = b->spans.back();
   s.b = bevent(b);
   if (s.b)
      hipEventRecord(s.b, b->stream);
}

int ceil_div(int a, int b) { return (a + b - 1) / b; }

// Rebuild the member table and prefix lists when a member's plan changed,
// a member joined or left, or the stencil's variables per item changed.
int ensure_table(amr_batch *b)
{
   bool need = !b->table_ok || b->table_vpw != b->vpw || b->table_wg != b->wg_mask ||
               b->table_exv != b->ex_vars;
   for (size_t i = 0; i < b->members.size() && !need; i++)
      if (b->seen[i] != b->members[i]->plan_version)
         need = true;
   if (!need)
      return AMR_OK;

   std::vector<int> order;
   for (size_t i = 0; i < b->members.size(); i++)
      if (b->active[i])
         order.push_back((int) i);
   std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
      return shape_class(b->members[x]->g) > shape_class(b->members[y]->g);
   });
   const int M = (int) order.size();
   const int P = M + 1;
   std::vector<BMember> hm(M);
   std::vector<int> pre(13 * P, 0);
   int *ex[3] = {&pre[0], &pre[P], &pre[2 * P]};
   int *st[6];
   for (int c = 0; c < 6; c++)
      st[c] = &pre[(3 + c) * P];
   int *rd = &pre[9 * P];
   int *fx[3] = {&pre[10 * P], &pre[11 * P], &pre[12 * P]};
   bool fusable = M > 0;
   b->plane_max = 0;
   for (int c = 0; c < 6; c++)
      b->plane_max_c[c] = 0;
   b->member_rows.assign(M, 0);
   for (int q = 0; q < M; q++) {
      amr_dev *d = b->members[order[q]];
      BMember &m = hm[q];
      m.pool = d->pool;
      m.jobs = d->d_jobs;
      m.sorted = d->d_sorted;
      m.partials = d->d_partials;
      m.sums = d->d_sums;
      if (int rc = amr_lay(d, 0, d->c.num_vars, &m.L))
         return berr(b, rc, "member %d: %s", q, d->err);
      m.Lo = amr_lay_other(d, m.L);
      m.fjobs = d->d_fjobs;
      m.redir = d->d_redir;
      m.edges = d->d_edges;
      fusable = fusable && d->two_interiors && !d->whole_ghosts && d->g.stencil != 0 &&
                !(d->tape && d->tape->on) && (d->d_fjobs || !(d->fcount[0] + d->fcount[1] +
                                                              d->fcount[2])) &&
                d->d_redir && d->d_edges;
      m.n_active = d->n_active;
      m.num_vars = d->c.num_vars;
      m.part_stride = d->c.max_blocks;
      m.cls = shape_class(d->g);
      int off = 0;
      for (int dir = 0; dir < 3; dir++) {
         m.count[dir] = d->count[dir];
         m.joff[dir] = off;
         off += d->count[dir];
         ex[dir][q + 1] = ex[dir][q] + d->count[dir] * ceil_div(m.num_vars, b->ex_vars);
      }
      for (int dir = 0, foff = 0; dir < 3; dir++) {
         m.fcount[dir] = d->fcount[dir];
         m.fjoff[dir] = foff;
         foff += d->fcount[dir];
         fx[dir][q + 1] = fx[dir][q] + d->fcount[dir] * ceil_div(m.num_vars, b->ex_vars);
      }
      const int items = d->n_active * ceil_div(m.num_vars, b->vpw);
      for (int c = 0; c < 6; c++)
         st[c][q + 1] = st[c][q] + (c == m.cls ? items : 0);
      rd[q + 1] = rd[q] + m.num_vars;
      const int pl = d->g.gy * d->g.gz;
      b->plane_max = std::max(b->plane_max, pl);
      b->plane_max_c[m.cls] = std::max(b->plane_max_c[m.cls], pl);
      b->member_rows[q] = d->d_sums ? d->sum_rows : 0;
   }
   BHIP(b, hipStreamSynchronize(b->stream));   // queued kernels may read the table
   if (M > b->mem_cap) {
      const int cap = std::max(M, 2 * b->mem_cap);
      hipFree(b->d_mem);
      hipFree(b->d_pre);
      b->d_mem = nullptr;
      b->d_pre = nullptr;
      b->mem_cap = 0;
      BHIP(b, hipMalloc(&b->d_mem, sizeof(BMember) * cap));
      BHIP(b, hipMalloc(&b->d_pre, sizeof(int) * 13 * (cap + 1)));
      b->mem_cap = cap;
   }
   if (!b->d_pre)
      BHIP(b, hipMalloc(&b->d_pre, sizeof(int) * 13));
   if (M)
      BHIP(b, hipMemcpy(b->d_mem, hm.data(), sizeof(BMember) * M, hipMemcpyHostToDevice));
   BHIP(b, hipMemcpy(b->d_pre, pre.data(), sizeof(int) * pre.size(), hipMemcpyHostToDevice));
   b->M = M;
   for (int dir = 0; dir < 3; dir++)
      b->tot_ex[dir] = ex[dir][M];
   for (int c = 0; c < 6; c++)
      b->tot_st[c] = st[c][M];
   b->tot_rd = rd[M];
   b->fusable = fusable;
   for (int dir = 0; dir < 3; dir++)
      b->tot_fx[dir] = fusable ? fx[dir][M] : 0;
   // expand the lists into item records
   size_t nex = 0, nst = 0;
   for (int dir = 0; dir < 3; dir++) {
      b->ex_off[dir] = nex;
      nex += b->tot_ex[dir];
   }
   for (int c = 0; c < 6; c++) {
      b->st_off[c] = nst;
      nst += b->tot_st[c];
   }
   if (nex > b->ex_cap) {
      hipFree(b->d_exitems);
      b->d_exitems = nullptr;
      b->ex_cap = 0;
      BHIP(b, hipMalloc(&b->d_exitems, sizeof(BExItem) * (nex + nex / 4)));
      b->ex_cap = nex + nex / 4;
   }
   size_t nfx = 0;
   for (int dir = 0; dir < 3; dir++) {
      b->fx_off[dir] = nfx;
      nfx += b->tot_fx[dir];
   }
   if (nfx > b->fx_cap) {
      hipFree(b->d_fxitems);
      b->d_fxitems = nullptr;
      b->fx_cap = 0;
      BHIP(b, hipMalloc(&b->d_fxitems, sizeof(BFxItem) * (nfx + nfx / 4)));
      b->fx_cap = nfx + nfx / 4;
   }
   if (nst > b->st_cap) {
      hipFree(b->d_stitems);
      b->d_stitems = nullptr;
      b->st_cap = 0;
      BHIP(b, hipMalloc(&b->d_stitems, sizeof(BStItem) * (nst + nst / 4)));
      b->st_cap = nst + nst / 4;
   }
   for (int dir = 0; dir < 3 && M; dir++)
      if (b->tot_ex[dir]) {
         hipLaunchKernelGGL(k_mat_exchange, dim3(std::min(4096, ceil_div(b->tot_ex[dir], 256))),
                            dim3(256), 0, b->stream, b->d_mem, b->d_pre + dir * P, M, dir,
                            b->ex_vars, b->d_exitems + b->ex_off[dir]);
         BHIP(b, hipGetLastError());
      }
   for (int dir = 0; dir < 3 && M; dir++)
      if (b->tot_fx[dir]) {
         hipLaunchKernelGGL(k_mat_fexchange, dim3(std::min(4096, ceil_div(b->tot_fx[dir], 256))),
                            dim3(256), 0, b->stream, b->d_mem, b->d_pre + (10 + dir) * P, M, dir,
                            b->ex_vars, b->d_fxitems + b->fx_off[dir]);
         BHIP(b, hipGetLastError());
      }
   for (int c = 0; c < 6 && M; c++)
      if (b->tot_st[c]) {
         hipLaunchKernelGGL(k_mat_stencil, dim3(std::min(4096, ceil_div(b->tot_st[c], 256))),
                            dim3(256), 0, b->stream, b->d_mem, b->d_pre + (3 + c) * P, M, b->vpw,
                            b->precision == AMR_PREC_F64 ? 8 : b->precision == AMR_PREC_F32 ? 4 : 2,
                            b->d_stitems + b->st_off[c]);
         BHIP(b, hipGetLastError());
      }
   BHIP(b, hipStreamSynchronize(b->stream));
   for (size_t i = 0; i < b->members.size(); i++)
      b->seen[i] = b->members[i]->plan_version;
   b->table_ok = true;
   b->table_vpw = b->vpw;
   b->table_wg = b->wg_mask;
   b->table_exv = b->ex_vars;
   b->epoch++;
   return AMR_OK;
}

int grid_of(int wgs, int total)
{
   const int need = ceil_div(total, kWavesPerWg);
   return wgs > 0 ? std::min(wgs, need) : need;
}

template <typename T, int ST, int VPW, int C, bool FUSED>
int launch_class(amr_batch *b, int want_partials, int flip)
{
   typedef typename Acc<T>::type A;
   const int total = b->tot_st[C];
   if (!total)
      return AMR_OK;
   const BStItem *items = b->d_stitems + b->st_off[C];
   if ((b->wg_mask >> (C / 2)) & 1) {
      const size_t lds = sizeof(A) * 3 * VPW * (size_t) b->plane_max_c[C];
      const int grid = b->st_wgs > 0 ? std::min(b->st_wgs, total) : total;
      hipLaunchKernelGGL((kb_stencil<T, ST, VPW, C, true, FUSED>), dim3(grid), dim3(256), lds,
                         b->stream, b->d_mem, items, total, want_partials, 0, flip);
   } else {
      const int ring = 3 * VPW * b->plane_max_c[C];
      const size_t lds = sizeof(A) * (size_t) ring * kWavesPerWg;
      hipLaunchKernelGGL((kb_stencil<T, ST, VPW, C, false, FUSED>),
                         dim3(grid_of(b->st_wgs, total)), dim3(256), lds, b->stream, b->d_mem,
                         items, total, want_partials, ring, flip);
   }
   BHIP(b, hipGetLastError());
   return AMR_OK;
}

template <typename T, int ST, int VPW, bool FUSED>
int launch_stencil_t(amr_batch *b, int want_partials, int flip)
{
   int rc = launch_class<T, ST, VPW, 4, FUSED>(b, want_partials, flip);
   if (!rc) rc = launch_class<T, ST, VPW, 5, FUSED>(b, want_partials, flip);
   if (!rc) rc = launch_class<T, ST, VPW, 2, FUSED>(b, want_partials, flip);
   if (!rc) rc = launch_class<T, ST, VPW, 3, FUSED>(b, want_partials, flip);
   if (!rc) rc = launch_class<T, ST, VPW, 0, FUSED>(b, want_partials, flip);
   if (!rc) rc = launch_class<T, ST, VPW, 1, FUSED>(b, want_partials, flip);
   return rc;
}

template <typename T, bool FUSED>
int launch_stencil_st(amr_batch *b, int want_partials, int flip)
{
   if (b->stencil == 7)
      return b->vpw == 2 ? launch_stencil_t<T, 7, 2, FUSED>(b, want_partials, flip)
                         : launch_stencil_t<T, 7, 1, FUSED>(b, want_partials, flip);
   return b->vpw == 2 ? launch_stencil_t<T, 27, 2, FUSED>(b, want_partials, flip)
                      : launch_stencil_t<T, 27, 1, FUSED>(b, want_partials, flip);
}

// The exchange of a fused stage: every member's fused plan, the cells in
// the copy `flip` names.
template <typename T>
int enqueue_fused_exchange(amr_batch *b, int flip)
{
   for (int dir = 0; dir < 3; dir++) {
      const int total = b->tot_fx[dir];
      if (!total)
         continue;
      const BFxItem *items = b->d_fxitems + b->fx_off[dir];
      if (b->ex_threads == 64) {
         auto k = b->ex_vars == 16 ? kb_exchange_fused<T, 16, 64>
                : b->ex_vars == 4 ? kb_exchange_fused<T, 4, 64> : kb_exchange_fused<T, 8, 64>;
         hipLaunchKernelGGL(k, dim3(grid_of(b->ex_wgs, total)), dim3(256), 0, b->stream,
                            b->d_mem, items, total, flip);
      } else {
         const int grid = b->ex_wgs > 0 ? std::min(b->ex_wgs, total) : total;
         auto k = b->ex_threads == 256
                     ? (b->ex_vars == 16 ? kb_exchange_fused<T, 16, 256>
                        : b->ex_vars == 4 ? kb_exchange_fused<T, 4, 256>
                                          : kb_exchange_fused<T, 8, 256>)
                     : (b->ex_vars == 16 ? kb_exchange_fused<T, 16, 128>
                        : b->ex_vars == 4 ? kb_exchange_fused<T, 4, 128>
                                          : kb_exchange_fused<T, 8, 128>);
         hipLaunchKernelGGL(k, dim3(grid), dim3(b->ex_threads), 0, b->stream, b->d_mem, items,
                            total, flip);
      }
      BHIP(b, hipGetLastError());
   }
   return AMR_OK;
}

// One stage; fused (every member's fused plan, reading the copy `flip`
// names and writing the other) or plain.
template <typename T>
int enqueue_stage_t(amr_batch *b, int stage, int permute, int row, bool fused, int flip)
{
   static const int perm[6][3] = { {0, 1, 2}, {1, 2, 0}, {2, 0, 1},
                                   {0, 2, 1}, {1, 0, 2}, {2, 1, 0} };
   span_begin(b, 0);
   if (fused) {
      if (int rc = enqueue_fused_exchange<T>(b, flip))
         return rc;
   }
   for (int o = 0; o < 3 && !fused; o++) {
      const int dir = permute ? perm[stage % 6][o] : o;
      if (!b->tot_ex[dir])
         continue;
      const int total = b->tot_ex[dir];
      const BExItem *items = b->d_exitems + b->ex_off[dir];
      if (b->ex_threads == 64) {
         auto k = b->ex_vars == 16 ? kb_exchange<T, 16, 64>
                : b->ex_vars == 4 ? kb_exchange<T, 4, 64> : kb_exchange<T, 8, 64>;
         hipLaunchKernelGGL(k, dim3(grid_of(b->ex_wgs, total)), dim3(256), 0, b->stream,
                            b->d_mem, items, total);
      } else {
         const int grid = b->ex_wgs > 0 ? std::min(b->ex_wgs, total) : total;
         auto k = b->ex_threads == 256
                     ? (b->ex_vars == 16 ? kb_exchange<T, 16, 256>
                        : b->ex_vars == 4 ? kb_exchange<T, 4, 256> : kb_exchange<T, 8, 256>)
                     : (b->ex_vars == 16 ? kb_exchange<T, 16, 128>
                        : b->ex_vars == 4 ? kb_exchange<T, 4, 128> : kb_exchange<T, 8, 128>);
         hipLaunchKernelGGL(k, dim3(grid), dim3(b->ex_threads), 0, b->stream, b->d_mem, items,
                            total);
      }
      BHIP(b, hipGetLastError());
   }
   span_end(b);
   const int tot_st = b->tot_st[0] + b->tot_st[1] + b->tot_st[2] + b->tot_st[3] +
                      b->tot_st[4] + b->tot_st[5];
   if (tot_st) {
      span_begin(b, 1);
      const int wp = row >= 0;
      const int rc = fused ? launch_stencil_st<T, true>(b, wp, flip)
                           : launch_stencil_st<T, false>(b, wp, 0);
      span_end(b);
      if (rc)
         return rc;
   }
   if (row >= 0 && b->tot_rd) {
      span_begin(b, 2);
      hipLaunchKernelGGL(kb_reduce, dim3(grid_of(0, b->tot_rd)), dim3(256), 0, b->stream,
                         b->d_mem, b->d_pre + 9 * (b->M + 1), b->M, row);
      span_end(b);
      BHIP(b, hipGetLastError());
   }
   return AMR_OK;
}

int enqueue_stage(amr_batch *b, int stage, int permute, int row, bool fused = false,
                  int flip = 0)
{
   switch (b->precision) {
   case AMR_PREC_F64: return enqueue_stage_t<double>(b, stage, permute, row, fused, flip);
   case AMR_PREC_F32: return enqueue_stage_t<float>(b, stage, permute, row, fused, flip);
   default: return enqueue_stage_t<uint16_t>(b, stage, permute, row, fused, flip);
   }
}

// MINIAMR_BATCH_FUSED=0 keeps the batch's stages plain.
bool batch_fused_on()
{
   static int v = -1;
   if (v < 0) {
      const char *e = getenv("MINIAMR_BATCH_FUSED");
      v = e ? atoi(e) != 0 : 1;
   }
   return v != 0;
}

// A timestep of `nstages` stages runs fused when every member has a fused
// plan and the count is even (so each interior ends in the copy it started
// in, which the work lists and the captured graph assume).
bool fuse_timestep(const amr_batch *b, int nstages, int permute)
{
   return b->fusable && !permute && nstages % 2 == 0 && batch_fused_on();
}

// After fused stages: the members' ghost planes of in-place faces are
// stale (amr_require_fresh brings them up to date when a member call needs
// them).
void mark_stale(amr_batch *b, bool stale)
{
   for (size_t i = 0; i < b->members.size(); i++)
      if (b->active[i])
         std::fill(b->members[i]->stale.begin(), b->members[i]->stale.end(), stale ? 1 : 0);
}

int check_row(amr_batch *b, int row)
{
   if (row < 0)
      return AMR_OK;
   for (int q = 0; q < b->M; q++)
      if (row >= b->member_rows[q])
         return berr(b, AMR_ERR_ARG, "checksum row %d not reserved by every member", row);
   return AMR_OK;
}

void drop_graph(amr_batch *b)
{
   if (b->gexec) {
      hipGraphExecDestroy(b->gexec);
      b->gexec = nullptr;
   }
}

// Time `reps` stages (no checksums) of the current shape, after one warm-up.
int time_stages(amr_batch *b, int reps, double *ms)
{
   int rc = ensure_table(b);
   if (rc)
      return rc;
   const bool tm = b->timing;
   b->timing = false;
   hipEvent_t e0 = bevent(b), e1 = bevent(b);
   if (!e0 || !e1)
      return berr(b, AMR_ERR_HIP, "autotune: no events");
   // the stages a timestep would run: fused ones when the timesteps are
   // (warm-up plus reps stages, an even count, so each interior ends in
   // the copy it started in)
   const bool fused = fuse_timestep(b, reps + 1, 0);
   rc = enqueue_stage(b, 0, 0, -1, fused, 0);
   hipEventRecord(e0, b->stream);
   for (int r = 0; r < reps && !rc; r++)
      rc = enqueue_stage(b, r, 0, -1, fused, (r + 1) & 1);
   hipEventRecord(e1, b->stream);
   b->timing = tm;
   BHIP(b, hipEventSynchronize(e1));
   float t = 0.0f;
   BHIP(b, hipEventElapsedTime(&t, e0, e1));
   b->ev_free.push_back(e0);
   b->ev_free.push_back(e1);
   *ms = t / reps;
   return rc;
}

}  // namespace

extern "C" {

int amr_batch_create(int device, int precision, int stencil, amr_batch **out)
{
   if (!out)
      return AMR_ERR_ARG;
   amr_batch *b = new amr_batch();
   *out = b;
   b->device = device;
   b->precision = precision;
   b->stencil = stencil;
   if (precision < AMR_PREC_F64 || precision > AMR_PREC_BF16)
      return berr(b, AMR_ERR_ARG, "unknown precision %d", precision);
   if (stencil != 7 && stencil != 27)
      return berr(b, AMR_ERR_ARG,
                  "stencil must be 7 or 27 (--stencil 0 runs on the single-instance driver)");
   int ndev = 0;
   if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
      return berr(b, AMR_ERR_NODEV, "no HIP device available");
   if (device < 0 || device >= ndev)
      return berr(b, AMR_ERR_NODEV, "device %d out of range (%d devices)", device, ndev);
   BHIP(b, hipSetDevice(device));
   hipDeviceProp_t prop;
   BHIP(b, hipGetDeviceProperties(&prop, device));
   if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
      return berr(b, AMR_ERR_NODEV, "device %d is %s, this library is built for gfx950", device,
                  prop.gcnArchName);
   b->cus = prop.multiProcessorCount;
   b->vpw = precision == AMR_PREC_F64 ? 1 : 2;
   BHIP(b, hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
   return AMR_OK;
}

int amr_batch_destroy(amr_batch *b)
{
   if (!b)
      return AMR_OK;
   if (b->stream)
      hipStreamSynchronize(b->stream);
   drop_graph(b);
   for (amr_dev *d : b->members)
      amr_dev_destroy(d);
   for (auto &s : b->spans) {
      if (s.a) hipEventDestroy(s.a);
      if (s.b) hipEventDestroy(s.b);
   }
   for (auto e : b->ev_free)
      hipEventDestroy(e);
   hipFree(b->d_mem);
   hipFree(b->d_pre);
   hipFree(b->d_exitems);
   hipFree(b->d_fxitems);
   hipFree(b->d_stitems);
   if (b->stream)
      hipStreamDestroy(b->stream);
   delete b;
   return AMR_OK;
}

const char *amr_batch_last_error(const amr_batch *b)
{
   return b ? b->err : "no batch";
}

int amr_batch_add(amr_batch *b, const amr_dev_config *cfg, amr_dev **member)
{
   if (!b || !cfg || !member)
      return berr(b, AMR_ERR_ARG, "add: bad arguments");
   *member = nullptr;
   if (!b->stream)
      return berr(b, AMR_ERR_STATE, "add: batch not created");
   if (cfg->precision != b->precision || cfg->stencil != b->stencil || cfg->device != b->device)
      return berr(b, AMR_ERR_ARG, "add: member precision/stencil/device differ from the batch");
   amr_dev *d = nullptr;
   int rc = amr_dev_create(cfg, &d);
   if (rc) {
      berr(b, rc, "add: %s", d ? amr_dev_last_error(d) : "context creation failed");
      amr_dev_destroy(d);
      return rc;
   }
   if (shape_class(d->g) < 0) {
      amr_dev_destroy(d);
      return berr(b, AMR_ERR_ARG, "add: block face %dx%d larger than 16x16", cfg->ny, cfg->nz);
   }
   // members work on the batch's stream, so member calls and batch stages
   // are ordered without events
   hipStreamSynchronize(d->stream);
   hipStreamDestroy(d->stream);
   d->stream = b->stream;
   d->owns_stream = false;
   b->members.push_back(d);
   b->active.push_back(1);
   b->seen.push_back(-1);
   b->table_ok = false;
   *member = d;
   return AMR_OK;
}

int amr_batch_set_active(amr_batch *b, amr_dev *member, int on)
{
   if (!b)
      return AMR_ERR_ARG;
   for (size_t i = 0; i < b->members.size(); i++)
      if (b->members[i] == member) {
         if ((b->active[i] != 0) != (on != 0)) {
            b->active[i] = on != 0;
            b->table_ok = false;
         }
         return AMR_OK;
      }
   return berr(b, AMR_ERR_ARG, "set_active: not a member of this batch");
}

int amr_batch_set_shape(amr_batch *b, const amr_batch_shape *sh)
{
   if (!b || !sh || (sh->vpw != 1 && sh->vpw != 2) || sh->stencil_wgs < 0 ||
       sh->exchange_wgs < 0 || sh->wg_classes < 0 || sh->wg_classes > 7 ||
       (sh->exchange_vars != 4 && sh->exchange_vars != 8 && sh->exchange_vars != 16) ||
       (sh->exchange_threads != 64 && sh->exchange_threads != 128 &&
        sh->exchange_threads != 256))
      return berr(b, AMR_ERR_ARG, "set_shape: vpw 1 or 2, grid sizes >= 0, wg_classes in "
                                  "0..7, exchange_vars 4, 8 or 16, exchange_threads 64, 128 "
                                  "or 256");
   b->vpw = sh->vpw;
   b->st_wgs = sh->stencil_wgs;
   b->ex_wgs = sh->exchange_wgs;
   b->wg_mask = sh->wg_classes;
   b->ex_vars = sh->exchange_vars;
   b->ex_threads = sh->exchange_threads;
   return AMR_OK;
}

int amr_batch_get_shape(const amr_batch *b, amr_batch_shape *sh)
{
   if (!b || !sh)
      return AMR_ERR_ARG;
   sh->vpw = b->vpw;
   sh->stencil_wgs = b->st_wgs;
   sh->exchange_wgs = b->ex_wgs;
   sh->wg_classes = b->wg_mask;
   sh->exchange_vars = b->ex_vars;
   sh->exchange_threads = b->ex_threads;
   return AMR_OK;
}

int amr_batch_stage(amr_batch *b, int stage, int permute, int sum_row)
{
   if (!b || stage < 0)
      return berr(b, AMR_ERR_ARG, "stage: bad arguments");
   int rc = ensure_table(b);
   if (!rc)
      rc = check_row(b, sum_row);
   if (!rc)
      rc = enqueue_stage(b, stage, permute, sum_row);
   if (!rc)
      mark_stale(b, false);   // a plain stage writes every ghost plane
   return rc;
}

int amr_batch_timestep(amr_batch *b, int stage0, int nstages, int checksum_freq, int permute)
{
   if (!b || stage0 < 0 || nstages < 0 || checksum_freq < 0)
      return berr(b, AMR_ERR_ARG, "timestep: bad arguments");
   if (!nstages)
      return AMR_OK;
   int rc = ensure_table(b);
   if (rc)
      return rc;
   if (checksum_freq && (rc = check_row(b, (nstages - 1) / checksum_freq)))
      return rc;
   const bool fused = fuse_timestep(b, nstages, permute);
   const long key[11] = {b->epoch, permute ? stage0 % 6 : -1, nstages, checksum_freq,
                         b->vpw, b->st_wgs, b->ex_wgs, (long) b->wg_mask, (long) b->ex_vars,
                         (long) b->ex_threads, (long) fused};
   if (!b->gexec || memcmp(key, b->gkey, sizeof(key))) {
      drop_graph(b);
      const bool tm = b->timing;
      b->timing = false;   // no event records inside the graph
      BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      for (int s = 0; s < nstages && !rc; s++) {
         const int row = (checksum_freq && !(s % checksum_freq)) ? s / checksum_freq : -1;
         rc = enqueue_stage(b, stage0 + s, permute, row, fused, s & 1);
      }
      hipGraph_t graph = nullptr;
      hipError_t e = hipStreamEndCapture(b->stream, &graph);
      b->timing = tm;
      if (rc) {
         if (graph) hipGraphDestroy(graph);
         return rc;
      }
      if (e != hipSuccess)
         return berr(b, AMR_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
      e = hipGraphInstantiate(&b->gexec, graph, nullptr, nullptr, 0);
      hipGraphDestroy(graph);
      if (e != hipSuccess) {
         b->gexec = nullptr;
         return berr(b, AMR_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
      }
      memcpy(b->gkey, key, sizeof(key));
   }
   span_begin(b, 3);
   BHIP(b, hipGraphLaunch(b->gexec, b->stream));
   span_end(b);
   // fused: ghost planes of in-place faces stale; plain: all written
   mark_stale(b, fused);
   return AMR_OK;
}

int amr_batch_autotune(amr_batch *b, double *ms_out)
{
   if (!b)
      return AMR_ERR_ARG;
   int rc = ensure_table(b);
   if (rc)
      return rc;
   if (!(b->tot_st[0] + b->tot_st[1] + b->tot_st[2] + b->tot_st[3] + b->tot_st[4] +
         b->tot_st[5]))
      return AMR_OK;
   // snapshot the pools of the active members
   std::vector<std::pair<amr_dev *, size_t>> pools;
   size_t total = 0;
   for (size_t i = 0; i < b->members.size(); i++)
      if (b->active[i]) {
         amr_dev *d = b->members[i];
         const size_t bytes = (size_t) d->g.var_stride * d->c.num_vars * d->elem_bytes;
         pools.push_back({d, bytes});
         total += bytes;
      }
   char *backup = nullptr;
   BHIP(b, hipMalloc(&backup, total));
   size_t off = 0;
   for (auto &p : pools) {
      BHIP(b, hipMemcpyAsync(backup + off, p.first->pool, p.second, hipMemcpyDeviceToDevice,
                             b->stream));
      off += p.second;
   }
   // coordinate descent: workgroup/wavefront items per size class, then
   // variables per stencil item, then the persistent grids
   const int reps = 3;   // + 1 warm-up: an even number of stages per trial
   // the members' stale flags, restored with the pools
   std::vector<std::vector<uint8_t>> stale0;
   for (size_t i = 0; i < b->members.size(); i++)
      stale0.push_back(b->members[i]->stale);
   const int mult[5] = {0, 1, 2, 4, 8};
   double best = 1e30;
   auto trial = [&](int vpw, int st, int ex, int wgm, int exv, int ext, bool *better) {
      b->vpw = vpw;
      b->st_wgs = st;
      b->ex_wgs = ex;
      b->wg_mask = wgm;
      b->ex_vars = exv;
      b->ex_threads = ext;
      double t = 0.0;
      int r = time_stages(b, reps, &t);
      *better = !r && t < best;
      if (*better)
         best = t;
      return r;
   };
   int bv = b->vpw, bs = 0, be = 0, bm = 0, bx = kVars, bt = 128;
   bool better = false;
   for (int wgm = 0; wgm < 8 && !rc; wgm++)
      if (!(rc = trial(bv, bs, be, wgm, bx, bt, &better)) && better)
         bm = wgm;
   for (int v = 1; v <= 2 && !rc; v++)
      if (!(rc = trial(v, bs, be, bm, bx, bt, &better)) && better)
         bv = v;
   for (int si = 1; si < 5 && !rc; si++)
      if (!(rc = trial(bv, mult[si] * b->cus, be, bm, bx, bt, &better)) && better)
         bs = mult[si] * b->cus;
   for (int xt : {64, 256})
      if (!rc && !(rc = trial(bv, bs, be, bm, bx, xt, &better)) && better)
         bt = xt;
   for (int xv : {4, 16})
      if (!rc && !(rc = trial(bv, bs, be, bm, xv, bt, &better)) && better)
         bx = xv;
   for (int ei = 1; ei < 5 && !rc; ei++)
      if (!(rc = trial(bv, bs, mult[ei] * b->cus, bm, bx, bt, &better)) && better)
         be = mult[ei] * b->cus;
   b->vpw = bv;
   b->st_wgs = bs;
   b->ex_wgs = be;
   b->wg_mask = bm;
   b->ex_vars = bx;
   b->ex_threads = bt;
   off = 0;
   for (auto &p : pools) {
      hipMemcpyAsync(p.first->pool, backup + off, p.second, hipMemcpyDeviceToDevice, b->stream);
      off += p.second;
   }
   hipError_t e = hipStreamSynchronize(b->stream);
   hipFree(backup);
   for (size_t i = 0; i < b->members.size(); i++)
      b->members[i]->stale = stale0[i];
   drop_graph(b);
   if (rc)
      return rc;
   if (e != hipSuccess)
      return berr(b, AMR_ERR_HIP, "autotune restore: %s", hipGetErrorString(e));
   if (ms_out)
      *ms_out = best;
   return ensure_table(b);
}

int amr_batch_sync(amr_batch *b)
{
   if (!b)
      return AMR_ERR_ARG;
   BHIP(b, hipStreamSynchronize(b->stream));
   return AMR_OK;
}

int amr_batch_set_timing(amr_batch *b, int on)
{
   if (!b)
      return AMR_ERR_ARG;
   b->timing = on != 0;
   return AMR_OK;
}

int amr_batch_read_timing(amr_batch *b, double ms[4])
{
   if (!b || !ms)
      return AMR_ERR_ARG;
   for (int i = 0; i < 4; i++)
      ms[i] = 0.0;
   BHIP(b, hipStreamSynchronize(b->stream));
   for (auto &s : b->spans) {
      float t = 0.0f;
      if (s.a && s.b && hipEventElapsedTime(&t, s.a, s.b) == hipSuccess)
         ms[s.cat] += t;
      if (s.a) b->ev_free.push_back(s.a);
      if (s.b) b->ev_free.push_back(s.b);
   }
   b->spans.clear();
   return AMR_OK;
}

}  // extern "C"
