// context.hip -- context lifetime, host<->device copies, mesh plan upload,
// checksum table, fused stage entry point and timing for libminiamr_dev.so.
#include <stdarg.h>
#include <stdlib.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <thread>

#include "amr_adjoint.h"
#include "amr_internal.h"

int amr_set_error(amr_dev *d, int rc, const char *fmt, ...)
{
   if (d) {
      va_list ap;
      va_start(ap, fmt);
      vsnprintf(d->err, sizeof(d->err), fmt, ap);
      va_end(ap);
   }
   return rc;
}

int amr_hip_check(amr_dev *d, hipError_t e, const char *what)
{
   return amr_set_error(d, e == hipErrorOutOfMemory ? AMR_ERR_NOMEM : AMR_ERR_HIP,
                        "%s: %s", what, hipGetErrorString(e));
}

static hipEvent_t get_event(amr_dev *d)
{
   if (!d->ev_free.empty()) {
      hipEvent_t e = d->ev_free.back();
      d->ev_free.pop_back();
      return e;
   }
   hipEvent_t e = nullptr;
   if (hipEventCreate(&e) != hipSuccess)
      return nullptr;
   return e;
}

void amr_span_begin(amr_dev *d, int cat)
{
   d->span_skipped = !d->timing || (d->timing_stencil_only && cat != 1);
   if (d->span_skipped)
      return;
   amr_dev::Span s;
   s.a = get_event(d);
   s.b = nullptr;
   s.cat = cat;
   if (s.a)
      hipEventRecord(s.a, d->stream);
   d->spans.push_back(s);
}

void amr_span_end(amr_dev *d)
{
   if (d->span_skipped || d->spans.empty())
      return;
   amr_dev::Span &s = d->spans.back();
   s.b = get_event(d);
   if (s.b)
      hipEventRecord(s.b, d->stream);
}

extern "C" {

int amr_dev_create(const amr_dev_config *cfg, amr_dev **out)
{
   if (!out)
      return AMR_ERR_ARG;
   *out = nullptr;
   if (!cfg)
      return AMR_ERR_ARG;
   amr_dev *d = new amr_dev();
   *out = d;
   d->c = *cfg;
   if (cfg->nx < 2 || cfg->ny < 2 || cfg->nz < 2 || (cfg->nx | cfg->ny | cfg->nz) & 1)
      return amr_set_error(d, AMR_ERR_ARG, "block sizes must be even and >= 2");
   if (cfg->num_vars < 1 || cfg->max_blocks < 1)
      return amr_set_error(d, AMR_ERR_ARG, "num_vars and max_blocks must be positive");
   if (cfg->stencil != 0 && cfg->stencil != 7 && cfg->stencil != 27)
      return amr_set_error(d, AMR_ERR_ARG, "stencil must be 0, 7 or 27");
   if (cfg->precision < AMR_PREC_F64 || cfg->precision > AMR_PREC_BF16)
      return amr_set_error(d, AMR_ERR_ARG, "unknown precision %d", cfg->precision);

   int ndev = 0;
   if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
      return amr_set_error(d, AMR_ERR_NODEV, "no HIP device available");
   if (cfg->device < 0 || cfg->device >= ndev)
      return amr_set_error(d, AMR_ERR_NODEV, "device %d out of range (%d devices)",
                           cfg->device, ndev);
   HIP_TRY(d, hipSetDevice(cfg->device));
   hipDeviceProp_t prop;
   HIP_TRY(d, hipGetDeviceProperties(&prop, cfg->device));
   if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
      return amr_set_error(d, AMR_ERR_NODEV, "device %d is %s, this library is built for gfx950",
                           cfg->device, prop.gcnArchName);
   if (int rc = amr_noalloc_init(d))   // static scratch of amr_dev_split_noalloc
      return rc;

   d->elem_bytes = cfg->precision == AMR_PREC_F64 ? 8 : cfg->precision == AMR_PREC_F32 ? 4 : 2;
   const char *fz = getenv("MINIAMR_FUSED");
   d->fused_mode = fz ? atoi(fz) : 1;
   if (cfg->stencil == 0)
      d->fused_mode = 0;   // the variable-work stencils read other variables' faces
   d->two_interiors = d->fused_mode != 0;   // second I and ZI
   const RecordLayout rl = make_record_layout(cfg->nx, cfg->ny, cfg->nz, cfg->stencil,
                                              d->elem_bytes, d->two_interiors);
   Lay &g = d->g;
   g = rl.g;
   for (int p = 0; p < 2; p++) {
      d->oI_par[p] = rl.oI_par[p];
      d->oZI_par[p] = rl.oZI_par[p];
   }
   d->par.assign(cfg->num_vars, 0);
   d->stale.assign(cfg->num_vars, 0);
   int64_t cells = (int64_t) g.gx * g.gy * g.gz;
   g.var_stride = g.rec * cfg->max_blocks;
   size_t pool_bytes = (size_t) g.var_stride * cfg->num_vars * d->elem_bytes;

   HIP_TRY(d, hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking));
   HIP_TRY(d, hipMalloc(&d->pool, pool_bytes));
   HIP_TRY(d, hipMemsetAsync(d->pool, 0, pool_bytes, d->stream));
   HIP_TRY(d, hipMalloc(&d->d_sorted, sizeof(int32_t) * cfg->max_blocks));
   HIP_TRY(d, hipMalloc(&d->d_partials, sizeof(double) * (size_t) cfg->num_vars * cfg->max_blocks));
   HIP_TRY(d, hipMalloc(&d->d_stage, sizeof(double) * (size_t) cells * cfg->num_vars));
   d->h_pinned_bytes = sizeof(double) * std::max<size_t>(4096, (size_t) cfg->num_vars * 64);
   HIP_TRY(d, hipHostMalloc((void **) &d->h_pinned, d->h_pinned_bytes, hipHostMallocDefault));
   if (cfg->stencil == 0) {
      HIP_TRY(d, hipMalloc(&d->d_check_counts, 2 * sizeof(unsigned long long)));
      HIP_TRY(d, hipMemsetAsync(d->d_check_counts, 0, 2 * sizeof(unsigned long long), d->stream));
   }
   HIP_TRY(d, hipStreamSynchronize(d->stream));
   return AMR_OK;
}

int amr_dev_destroy(amr_dev *d)
{
   if (!d)
      return AMR_OK;
   if (d->stream)
      hipStreamSynchronize(d->stream);
   for (auto &s : d->spans) {
      if (s.a) hipEventDestroy(s.a);
      if (s.b) hipEventDestroy(s.b);
   }
   for (auto e : d->ev_free)
      hipEventDestroy(e);
   hipFree(d->pool);
   hipFree(d->d_sorted);
   hipFree(d->d_jobs);
   hipFree(d->d_partials);
   hipFree(d->d_sums);
   hipFree(d->d_stage);
   hipFree(d->d_scratch);
   hipFree(d->d_jobbuf);
   hipFree(d->d_codec);
   hipFree(d->d_msgs);
   hipFree(d->d_cmp);
   hipFree(d->d_fjobs);
   hipFree(d->d_redir);
   hipFree(d->d_edges);
   hipFree(d->d_mjobs);
   hipFree(d->d_a0);
   hipFree(d->d_check_counts);
   hipFree(d->d_matwork);
   hipFree(d->mk_blocks);
   hipFree(d->mk_parents);
   hipFree(d->mk_objects);
   hipFree(d->mk_levels);
   hipFree(d->mk_out);
   if (d->h_pinned)
      hipHostFree(d->h_pinned);
   if (d->h_plan)
      hipHostFree(d->h_plan);
   if (d->stream && d->owns_stream)
      hipStreamDestroy(d->stream);
   if (d->mark_stream && d->mark_stream != d->stream)
      hipStreamDestroy(d->mark_stream);
   if (d->h_jobstage)
      hipHostFree(d->h_jobstage);
   amr_tape_free(d);
   delete d;
   return AMR_OK;
}

const char *amr_dev_last_error(const amr_dev *d)
{
   return d ? d->err : "no context";
}

int amr_dev_upload_block(amr_dev *d, int slot, const double *cells)
{
   if (!d || !cells || slot < 0 || slot >= d->c.max_blocks)
      return amr_set_error(d, AMR_ERR_ARG, "upload: slot %d out of range", slot);
   size_t bytes = sizeof(double) * (size_t) d->g.gx * d->g.gy * d->g.gz * d->c.num_vars;
   HIP_TRY(d, hipMemcpyAsync(d->d_stage, cells, bytes, hipMemcpyHostToDevice, d->stream));
   int rc = amr_launch_convert_in(d, slot);
   if (rc)
      return rc;
   HIP_TRY(d, hipStreamSynchronize(d->stream));
   amr_tape_record(d, AMR_OP_UPLOAD, slot);
   return AMR_OK;
}

int amr_dev_download_block(amr_dev *d, int slot, double *cells)
{
   if (d && amr_require_fresh(d, 0, d->c.num_vars, "download_block"))
      return AMR_ERR_STATE;
   if (!d || !cells || slot < 0 || slot >= d->c.max_blocks)
      return amr_set_error(d, AMR_ERR_ARG, "download: slot %d out of range", slot);
   int rc = amr_launch_convert_out(d, slot);
   if (rc)
      return rc;
   size_t bytes = sizeof(double) * (size_t) d->g.gx * d->g.gy * d->g.gz * d->c.num_vars;
   HIP_TRY(d, hipMemcpyAsync(cells, d->d_stage, bytes, hipMemcpyDeviceToHost, d->stream));
   HIP_TRY(d, hipStreamSynchronize(d->stream));
   return AMR_OK;
}

}  // extern "C"

int amr_lay(amr_dev *d, int start, int num, Lay *out)
{
   *out = d->g;
   if (num <= 0)
      return AMR_OK;
   const int p = d->par[start];
   for (int v = start + 1; v < start + num; v++)
      if (d->par[v] != p)
         return amr_set_error(d, AMR_ERR_STATE,
                              "variables %d and %d hold their cells in different copies",
                              start, v);
   out->oI = d->oI_par[p];
   out->oZI = d->oZI_par[p];
   return AMR_OK;
}

Lay amr_lay_other(const amr_dev *d, const Lay &L)
{
   Lay o = L;
   const int p = L.oI == d->oI_par[1] && d->two_interiors ? 1 : 0;
   o.oI = d->oI_par[1 - p];
   o.oZI = d->oZI_par[1 - p];
   return o;
}

// MINIAMR_HOIST=0 keeps every fused job in its own direction's launch (the
// default moves the y and z coarse/fine jobs of blocks without a boundary
// face into the first launch, below).
static bool hoist_on()
{
   static int v = -1;
   if (v < 0) {
      const char *e = getenv("MINIAMR_HOIST");
      v = e ? atoi(e) != 0 : 1;
   }
   return v != 0;
}

// MINIAMR_FOLD_EDGES=0 keeps the 27-point edge extensions of in-place faces
// as exchange jobs (the default folds them into the reads, below).
static bool fold_edges_on()
{
   static int v = -1;
   if (v < 0) {
      const char *e = getenv("MINIAMR_FOLD_EDGES");
      v = e ? atoi(e) != 0 : 1;
   }
   return v != 0;
}

// Edge sources of a fused stage (EdgeSrc).  In the reference a ghost edge
// cell of a block gets its value from the last job that writes it in a
// stage (27-point stencil, --code 0; ref/comm.c:1597-1688): a cell with a z
// ghost coordinate from the block's z-face job, which copies the z-plane
// ring of its source; a cell with x and y ghost coordinates (z interior)
// from the y-face job, which copies the source's x-extended rows; each
// copies the source's cell at the same in-plane coordinates on the plane
// next to the shared face.  Only same-level jobs of blocks without a
// boundary face extend and are folded (a boundary block's jobs all run);
// a prolongation or restriction writes no edge cell, which then keeps what
// the slot held (stale).  Following these copies, and the in-place face
// reads of the plan, back to a cell that some job in the stage does write
// or that is an interior cell gives the value the stage's stencil sees.
namespace {

struct Cell { int slot, i, j, k; };

struct EdgeResolver {
   const Lay &g;
   const std::vector<int32_t> &same;
   const std::vector<uint8_t> &bc;
   Lay L0;   // g with the interior copies at offset 0 (ES_I / ES_ZI offsets)

   EdgeResolver(const Lay &g_, const std::vector<int32_t> &same_, const std::vector<uint8_t> &bc_)
      : g(g_), same(same_), bc(bc_), L0(g_)
   {
      L0.oI = 0;
      L0.oZI = 0;
   }

   // the cell whose value block b's cell (i, j, k) holds when the stencil runs
   Cell resolve(int b, int i, int j, int k) const
   {
      for (int hop = 0; hop < 8; hop++) {
         const bool gi = i == 0 || i == g.nx + 1, gj = j == 0 || j == g.ny + 1,
                    gk = k == 0 || k == g.nz + 1;
         const int ng = gi + gj + gk;
         if (!ng || bc[b])
            return {b, i, j, k};
         int f, src;
         if (ng == 1)           // face interior: read in place, or the job writes it
            f = gi ? (i != 0) : gj ? 2 + (j != 0) : 4 + (k != 0);
         else if (gk)           // z ring: the z-face job's extension
            f = 4 + (k != 0);
         else                   // x and y ghost, z interior: the y-face job's extension
            f = 2 + (j != 0);
         src = same[6 * (size_t) b + f];
         if (src < 0)
            return {b, i, j, k};
         b = src;
         if (f < 2) i = (f & 1) ? 1 : g.nx;
         else if (f < 4) j = (f & 1) ? 1 : g.ny;
         else k = (f & 1) ? 1 : g.nz;
      }
      return {-1, 0, 0, 0};
   }

   // a cell as an entry: axis is the direction of the line (-1 for a corner);
   // cells at z = 1 / z = nz of lines across z are read from the z-plane copies
   bool entry(const Cell &c, int axis, EdgeSrc &e) const
   {
      if (c.slot < 0)
         return false;
      const bool ghost = c.i == 0 || c.i == g.nx + 1 || c.j == 0 || c.j == g.ny + 1 ||
                         c.k == 0 || c.k == g.nz + 1;
      const bool zc = axis != 2 && (c.k == 1 || c.k == g.nz);
      e.slot = c.slot;
      e.stride = 0;
      if (!ghost) {
         e.kind = zc ? ES_ZI : ES_I;
         e.off = zc ? zcopy(L0, c.i, c.j, c.k) : home(L0, c.i, c.j, c.k);
      } else {
         e.kind = ES_ABS;
         e.off = zc ? zcopy(g, c.i, c.j, c.k) : home(g, c.i, c.j, c.k);
      }
      return true;
   }

   // entry of a line of n cells, cell(t) giving block b's cell t (resolved
   // when folding, else the block's own); false if the cells are not one
   // affine run of one block and kind.  Every cell of a line follows the
   // same copies (the jobs that write a line, and the in-place faces on the
   // way, are those of its whole face), so the first two cells give the
   // run and the last one checks it.
   template <typename F>
   bool line(int b, int axis, int n, F cell, bool fold, EdgeSrc &e) const
   {
      auto at = [&](int t, EdgeSrc &x) {
         const Cell c = cell(t);
         return entry(fold ? resolve(b, c.i, c.j, c.k) : c, axis, x);
      };
      if (!at(0, e))
         return false;
      if (n > 1) {
         EdgeSrc e1, el;
         if (!at(1, e1) || e1.slot != e.slot || e1.kind != e.kind)
            return false;
         e.stride = e1.off - e.off;
         if (n > 2 && (!at(n - 1, el) || el.slot != e.slot || el.kind != e.kind ||
                       el.off != e.off + (n - 1) * e.stride))
            return false;
      }
      return true;
   }

   // the kEdgeSrcs entries of block b
   bool block(int b, bool fold, EdgeSrc *out) const
   {
      const int X[2] = {0, g.nx + 1}, Y[2] = {0, g.ny + 1}, Z[2] = {0, g.nz + 1};
      bool ok = true;
      for (int xs = 0; xs < 2; xs++)
         for (int ys = 0; ys < 2; ys++)
            ok &= line(b, 2, g.nz, [&](int t) { return Cell{b, X[xs], Y[ys], t + 1}; }, fold,
                       out[xs * 2 + ys]);
      for (int xs = 0; xs < 2; xs++)
         for (int zs = 0; zs < 2; zs++)
            ok &= line(b, 1, g.ny, [&](int t) { return Cell{b, X[xs], t + 1, Z[zs]}; }, fold,
                       out[4 + xs * 2 + zs]);
      for (int ys = 0; ys < 2; ys++)
         for (int zs = 0; zs < 2; zs++)
            ok &= line(b, 0, g.nx, [&](int t) { return Cell{b, t + 1, Y[ys], Z[zs]}; }, fold,
                       out[8 + ys * 2 + zs]);
      for (int xs = 0; xs < 2; xs++)
         for (int ys = 0; ys < 2; ys++)
            for (int zs = 0; zs < 2; zs++)
               ok &= line(b, -1, 1, [&](int) { return Cell{b, X[xs], Y[ys], Z[zs]}; }, fold,
                          out[12 + xs * 4 + ys * 2 + zs]);
      return ok;
   }
};

}  // namespace

// The plan of amr_dev_stage_fused.  A block's face is read in place when
// its job is a same-level copy and the block has no boundary face (a
// boundary fill with the 27-point extension reads the block's own ghost
// planes as they were before the later phases of the stage, which an
// in-place read cannot reproduce).  Such a face's job is dropped (x, and
// y/z for the 7-point stencil) or reduced to its edge extension (y/z for
// the 27-point stencil); every job that reads a source's ghost row that is
// read in place is redirected to the block that row would be copied from.
// It is computed on the host only; set_mesh uploads it.
struct FusedPlan {
   std::vector<FusedJob> fj;
   int32_t fc[3] = {0, 0, 0};
   std::vector<EdgeSrc> edges;
   std::vector<int32_t> redir;
   std::vector<amr_face_job> mj;
   int32_t mc[3] = {0, 0, 0};
   bool fold = false;
};

static void compute_fused_plan(const amr_dev *d, const int32_t *sorted, int n_active,
                               const amr_face_job *jobs, const int32_t count[3], FusedPlan &P)
{
   const int M = d->c.max_blocks;
   std::vector<int32_t> same(6 * (size_t) M, -1);
   std::vector<uint8_t> bc(M, 0);
   const int nj = count[0] + count[1] + count[2];
   for (int t = 0; t < nj; t++) {
      const amr_face_job &j = jobs[t];
      if (j.kind == AMR_FACE_BC)
         bc[j.dst] = 1;
      else if (j.kind == AMR_FACE_SAME)
         same[6 * (size_t) j.dst + j.face] = j.src;
   }
   auto inplace = [&](int slot, int f) { return bc[slot] ? -1 : same[6 * (size_t) slot + f]; };
   // the in-place faces of each slot (those of boundary blocks are not)
   std::vector<int32_t> inp(6 * (size_t) M, -1);
   for (int in = 0; in < n_active; in++)
      for (int f = 0; f < 6; f++)
         inp[6 * (size_t) sorted[in] + f] = inplace(sorted[in], f);
   // edge sources of every active block: folded (27-point stencil) unless a
   // line is not one affine run, in which case the plan keeps the extensions
   std::vector<EdgeSrc> &edges = P.edges;
   edges.assign(kEdgeSrcs * (size_t) M, EdgeSrc{0, 0, 0, ES_ABS});
   bool fold = d->g.stencil == 27 && fold_edges_on();
   {
      // the blocks are independent: spread them over a few host threads
      // (the resolution takes ~2 ms for 5,500 blocks on one)
      const EdgeResolver R(d->g, inp, bc);
      auto resolve_range = [&](int i0, int i1, bool f) {
         bool ok = true;
         for (int in = i0; in < i1; in++)
            ok &= R.block(sorted[in], f, &edges[kEdgeSrcs * (size_t) sorted[in]]);
         return ok;
      };
      const int nt = n_active >= 2048 ? 8 : 1;
      auto run_all = [&](bool f) {
         if (nt == 1)
            return resolve_range(0, n_active, f);
         std::vector<std::thread> th;
         std::vector<char> ok(nt, 1);
         for (int t = 0; t < nt; t++)
            th.emplace_back([&, t] {
               ok[t] = resolve_range((int) ((int64_t) n_active * t / nt),
                                     (int) ((int64_t) n_active * (t + 1) / nt), f);
            });
         for (auto &x : th)
            x.join();
         return std::all_of(ok.begin(), ok.end(), [](char c) { return c != 0; });
      };
      if (fold)
         fold = run_all(true);
      if (!fold)
         run_all(false);
   }
   P.fold = fold;
   // Jobs by launch.  The y and z prolongations and restrictions of blocks
   // without a boundary face go with the x phase: they read only interior
   // cells, which no job writes, and write ghost cells that only later
   // phases' jobs (27-point rings) and the stencil read -- an x job reads
   // ghost cells only in a boundary fill of its own block.  The y and z
   // launches keep the jobs that read rings.
   std::vector<FusedJob> ph[3];
   const bool hoist = hoist_on();
   for (int t = 0, dir = 0; dir < 3; dir++)
      for (int e = 0; e < count[dir]; e++, t++) {
         const amr_face_job &j = jobs[t];
         FusedJob f;
         f.jb = j;
         f.mode = FJ_FULL;
         f.alt[0] = f.alt[1] = f.alt[2] = f.alt[3] = -1;
         f.pad = 0;
         if (j.kind == AMR_FACE_SAME && inplace(j.dst, j.face) >= 0) {
            if (dir == 0 || d->g.stencil == 7 || fold)
               continue;                    // nothing outside the face interior, or
                                            // the extension is read where it is
            f.mode = FJ_EXT_ONLY;
         }
         if (d->g.stencil == 27 && dir > 0 &&
             (j.kind == AMR_FACE_SAME || j.kind == AMR_FACE_BC)) {
            f.alt[0] = inplace(j.src, 0);
            f.alt[1] = inplace(j.src, 1);
            if (dir == 2) {
               f.alt[2] = inplace(j.src, 2);
               f.alt[3] = inplace(j.src, 3);
            }
         }
         const bool cf = j.kind == AMR_FACE_PROLONG || j.kind == AMR_FACE_RESTRICT;
         ph[hoist && dir > 0 && cf && !bc[j.dst] ? 0 : dir].push_back(f);
      }
   // the jobs that bring stale ghost planes up to date (amr_require_fresh)
   std::vector<amr_face_job> &mj = P.mj;
   int32_t *mc = P.mc;
   for (int t = 0, dir = 0; dir < 3; dir++)
      for (int e = 0; e < count[dir]; e++, t++)
         if (jobs[t].kind == AMR_FACE_SAME && inplace(jobs[t].dst, jobs[t].face) >= 0) {
            mj.push_back(jobs[t]);
            mc[dir]++;
         }
   std::vector<FusedJob> &fj = P.fj;
   fj.reserve(nj);
   for (int p = 0; p < 3; p++) {
      fj.insert(fj.end(), ph[p].begin(), ph[p].end());
      P.fc[p] = (int32_t) ph[p].size();
   }
   P.redir.assign(6 * (size_t) std::max(n_active, 1), -1);
   for (int in = 0; in < n_active; in++)
      for (int f = 0; f < 6; f++)
         P.redir[6 * (size_t) in + f] = inplace(sorted[in], f);
}

// A fused stage leaves unwritten the ghost cells its stencil reads where
// they are: the planes of the faces read in place and their 27-point edge
// extensions, which are exactly what the same-level jobs of blocks without
// a boundary face write (every other job runs in the fused stage).  The
// interiors those jobs read in the plain stage are the ones the fused stage
// read, which the record still holds in its other interior copy until the
// next stage writes it; a job's reads of other ghost cells find them as the
// plain stage's earlier phases left them (written by the fused stage, or by
// this rerun's earlier directions).  So rerunning those jobs, in direction
// order, on the other copy gives every ghost cell the plain stage's value.
int amr_require_fresh(amr_dev *d, int start, int num, const char *what)
{
   const int end = std::min(start + num, (int) d->stale.size());
   for (int a = start; a < end;) {
      if (!d->stale[a]) {
         a++;
         continue;
      }
      int b = a + 1;   // a run of stale variables whose cells are in one copy
      while (b < end && d->stale[b] && d->par[b] == d->par[a])
         b++;
      Lay L;
      if (int rc = amr_lay(d, a, b - a, &L))
         return rc;
      const Lay Lr = amr_lay_other(d, L);
      for (int dir = 0, off = 0; dir < 3; off += d->mcount[dir], dir++)
         if (int rc = amr_launch_exchange_jobs(d, d->d_mjobs + off, d->mcount[dir], Lr, a, b - a))
            return amr_set_error(d, rc, "%s: bringing ghost planes up to date: %s", what,
                                 d->err);
      for (int v = a; v < b; v++)
         d->stale[v] = 0;
      a = b;
   }
   return AMR_OK;
}

int amr_set_mesh_impl(amr_dev *d, const int32_t *sorted, int n_active,
                      const amr_face_job *jobs, const int32_t count[3])
{
   if (!d || n_active < 0 || n_active > d->c.max_blocks || (n_active && !sorted) || !count)
      return amr_set_error(d, AMR_ERR_ARG, "set_mesh: bad arguments");
   int nj = 0;
   for (int i = 0; i < 3; i++) {
      if (count[i] < 0)
         return amr_set_error(d, AMR_ERR_ARG, "set_mesh: negative job count");
      nj += count[i];
   }
   if (nj && !jobs)
      return amr_set_error(d, AMR_ERR_ARG, "set_mesh: jobs missing");
   if (int rc = amr_require_fresh(d, 0, d->c.num_vars, "set_mesh"))
      return rc;
   // Validate every job on the host: slots in range and faces/quarters legal,
   // so no kernel can address outside the pool.
   for (int i = 0; i < n_active; i++)
      if (sorted[i] < 0 || sorted[i] >= d->c.max_blocks)
         return amr_set_error(d, AMR_ERR_ARG, "set_mesh: slot %d out of range", sorted[i]);
   for (int t = 0; t < nj; t++) {
      const amr_face_job &j = jobs[t];
      int dir = t < count[0] ? 0 : t < count[0] + count[1] ? 1 : 2;
      if (j.kind < AMR_FACE_SAME || j.kind > AMR_FACE_BC_GHOSTS || j.face / 2 != dir ||
          j.face < 0 || j.face > 5 || j.dst < 0 || j.dst >= d->c.max_blocks ||
          j.src < 0 || j.src >= d->c.max_blocks || (j.iq & ~1) || (j.jq & ~1))
         return amr_set_error(d, AMR_ERR_ARG, "set_mesh: bad face job %d", t);
   }
   // The fused plan is host work; it runs before the wait below, so it
   // overlaps the device work still queued (split and consolidate data).
   FusedPlan fp;
   if (d->two_interiors)
      compute_fused_plan(d, sorted, n_active, jobs, count, fp);
   // Queued work may still read the old plan and the staging buffer may
   // still feed the last set_mesh's copies: wait, then replace.
   HIP_TRY(d, hipStreamSynchronize(d->stream));
   auto grow = [&](void **buf, int *cap, int need, size_t elem) -> int {
      if (need <= *cap)
         return AMR_OK;
      const int c = std::max(need, 2 * *cap);
      hipFree(*buf);
      *buf = nullptr;
      *cap = 0;
      HIP_TRY(d, hipMalloc(buf, elem * (size_t) c));
      *cap = c;
      return AMR_OK;
   };
   if (int rc = grow((void **) &d->d_jobs, &d->jobs_cap, nj, sizeof(amr_face_job)))
      return rc;
   if (d->two_interiors) {
      if (int rc = grow((void **) &d->d_fjobs, &d->fjobs_cap, (int) fp.fj.size(), sizeof(FusedJob)))
         return rc;
      if (int rc = grow((void **) &d->d_mjobs, &d->mjobs_cap, (int) fp.mj.size(),
                        sizeof(amr_face_job)))
         return rc;
      if (!d->d_edges)
         HIP_TRY(d, hipMalloc(&d->d_edges, sizeof(EdgeSrc) * fp.edges.size()));
      if (!d->d_redir)
         HIP_TRY(d, hipMalloc(&d->d_redir, sizeof(int32_t) * 6 * (size_t) d->c.max_blocks));
   }
   // One pinned staging buffer, then copies ordered on the stream before
   // any later launch (pageable hipMemcpy of the ~3 MB plan took 0.5-2 ms).
   struct Piece {
      void *dst;
      const void *src;
      size_t bytes;
   };
   std::vector<Piece> pieces;
   pieces.push_back({d->d_sorted, sorted, sizeof(int32_t) * (size_t) n_active});
   pieces.push_back({d->d_jobs, jobs, sizeof(amr_face_job) * (size_t) nj});
   if (d->two_interiors) {
      pieces.push_back({d->d_fjobs, fp.fj.data(), sizeof(FusedJob) * fp.fj.size()});
      pieces.push_back({d->d_edges, fp.edges.data(), sizeof(EdgeSrc) * fp.edges.size()});
      pieces.push_back({d->d_redir, fp.redir.data(), sizeof(int32_t) * fp.redir.size()});
      pieces.push_back({d->d_mjobs, fp.mj.data(), sizeof(amr_face_job) * fp.mj.size()});
   }
   size_t total = 0;
   for (const Piece &q : pieces)
      total += (q.bytes + 255) / 256 * 256;
   if (total > d->h_plan_bytes) {
      if (d->h_plan)
         hipHostFree(d->h_plan);
      d->h_plan = nullptr;
      d->h_plan_bytes = 0;
      HIP_TRY(d, hipHostMalloc(&d->h_plan, total, hipHostMallocDefault));
      d->h_plan_bytes = total;
   }
   size_t at = 0;
   for (const Piece &q : pieces) {
      if (!q.bytes)
         continue;
      char *h = (char *) d->h_plan + at;
      memcpy(h, q.src, q.bytes);
      HIP_TRY(d, hipMemcpyAsync(q.dst, h, q.bytes, hipMemcpyHostToDevice, d->stream));
      at += (q.bytes + 255) / 256 * 256;
   }
   d->h_sorted.assign(sorted, sorted + n_active);
   d->n_active = n_active;
   for (int i = 0; i < 3; i++)
      d->count[i] = count[i];
   d->whole_ghosts = false;
   for (int t = 0; t < nj; t++)
      d->whole_ghosts |= (jobs[t].kind & AMR_FACE_GHOSTS) != 0;
   d->plan_version++;
   if (d->two_interiors) {
      d->fold_edges = fp.fold;
      for (int i = 0; i < 3; i++) {
         d->fcount[i] = fp.fc[i];
         d->mcount[i] = fp.mc[i];
      }
   }
   return AMR_OK;
}

int amr_comm_impl(amr_dev *d, int start, int num, int stage, int permute)
{
   static const int perm[6][3] = { {0, 1, 2}, {1, 2, 0}, {2, 0, 1},
                                   {0, 2, 1}, {1, 0, 2}, {2, 1, 0} };
   if (!d || start < 0 || num < 0 || start + num > d->c.num_vars || stage < 0)
      return amr_set_error(d, AMR_ERR_ARG, "comm: variables [%d, %d) out of range", start,
                           start + num);
   if (!num)
      return AMR_OK;
   // one span per direction when timing per direction, else one in all
   if (!d->timing_dirs)
      amr_span_begin(d, 0);
   for (int o = 0; o < 3; o++) {
      int dir = permute ? perm[stage % 6][o] : o;
      if (d->timing_dirs)
         amr_span_begin(d, kSpanExchangeX + dir);
      int rc = amr_launch_exchange_dir(d, dir, start, num);
      if (d->timing_dirs)
         amr_span_end(d);
      if (rc)
         return rc;
   }
   if (!d->timing_dirs)
      amr_span_end(d);
   // every ghost cell of these variables is written again
   for (int v = start; v < start + num; v++)
      d->stale[v] = 0;
   return AMR_OK;
}

int amr_stencil_impl(amr_dev *d, int var, int calc_stage)
{
   if (!d || var < 0 || var >= d->c.num_vars || calc_stage < 0)
      return amr_set_error(d, AMR_ERR_ARG, "stencil: variable %d out of range", var);
   if (int rc = amr_require_fresh(d, var, 1, "stencil"))
      return rc;
   amr_span_begin(d, 1);
   int rc = amr_launch_stencil(d, var, 1, false, calc_stage);
   amr_span_end(d);
   return rc;
}

extern "C" {

int amr_dev_set_mesh(amr_dev *d, const int32_t *sorted, int n_active,
                     const amr_face_job *jobs, const int32_t count[3])
{
   int rc = amr_set_mesh_impl(d, sorted, n_active, jobs, count);
   if (rc || !d->tape || !d->tape->on || d->replaying)
      return rc;
   TapePlan pl;
   pl.sorted.assign(sorted, sorted + n_active);
   pl.jobs.assign(jobs, jobs + (count[0] + count[1] + count[2]));
   for (int i = 0; i < 3; i++)
      pl.count[i] = count[i];
   d->tape->plans.push_back(std::move(pl));
   d->tape->cur_plan = (int) d->tape->plans.size() - 1;
   amr_tape_record(d, AMR_OP_SET_MESH, d->tape->cur_plan);
   return AMR_OK;
}

int amr_dev_comm(amr_dev *d, int start, int num, int stage, int permute)
{
   int rc = amr_comm_impl(d, start, num, stage, permute);
   if (!rc)
      amr_tape_record(d, AMR_OP_COMM, start, num, stage, permute);
   return rc;
}

int amr_dev_stencil(amr_dev *d, int var)
{
   int rc = amr_stencil_impl(d, var, 0);
   if (!rc)
      amr_tape_record(d, AMR_OP_STENCIL, var);
   return rc;
}

int amr_dev_stencil_stage(amr_dev *d, int var, int calc_stage)
{
   int rc = amr_stencil_impl(d, var, calc_stage);
   if (!rc)
      amr_tape_record(d, AMR_OP_STENCIL, var);
   return rc;
}

int amr_dev_set_materials(amr_dev *d, double a1, const double *a0, int n)
{
   if (!d || d->g.stencil != 0)
      return amr_set_error(d, AMR_ERR_ARG, "set_materials: the context is not a --stencil 0 one");
   if (!a0 || n != d->c.num_vars / 4 || n < 1)
      return amr_set_error(d, AMR_ERR_ARG, "set_materials: %d coefficients, need num_vars/4 = %d",
                           n, d->c.num_vars / 4);
   HIP_TRY(d, hipStreamSynchronize(d->stream));
   if (!d->d_a0)
      HIP_TRY(d, hipMalloc(&d->d_a0, sizeof(double) * n));
   HIP_TRY(d, hipMemcpy(d->d_a0, a0, sizeof(double) * n, hipMemcpyHostToDevice));
   d->mat = n;
   d->a1 = a1;
   return AMR_OK;
}

int amr_dev_read_check_counts(amr_dev *d, int64_t counts[2])
{
   if (!d || !counts)
      return AMR_ERR_ARG;
   counts[0] = counts[1] = 0;
   if (!d->d_check_counts)
      return AMR_OK;
   unsigned long long h[2];
   HIP_TRY(d, hipStreamSynchronize(d->stream));
   HIP_TRY(d, hipMemcpy(h, d->d_check_counts, sizeof(h), hipMemcpyDeviceToHost));
   counts[0] = (int64_t) h[0];
   counts[1] = (int64_t) h[1];
   return AMR_OK;
}

int amr_dev_check_sum(amr_dev *d, int var, double *sum)
{
   if (!d || !sum || var < 0 || var >= d->c.num_vars)
      return amr_set_error(d, AMR_ERR_ARG, "check_sum: variable %d out of range", var);
   amr_span_begin(d, 2);
   int rc = amr_launch_block_sums(d, var, 1);
   if (!rc)
      rc = amr_launch_reduce_sums(d, var, 1, (double *) d->d_stage);
   amr_span_end(d);
   if (rc)
      return rc;
   HIP_TRY(d, hipMemcpyAsync(d->h_pinned, d->d_stage, sizeof(double), hipMemcpyDeviceToHost,
                             d->stream));
   HIP_TRY(d, hipStreamSynchronize(d->stream));
   *sum = d->h_pinned[0];
   return AMR_OK;
}

}  // extern "C"

// Snapshot scratch for split/merge: every job needs num_vars * (n/2)^3
// values; batches larger than the cap are processed in order in chunks
// (safe: a chunk can only overwrite slots freed by itself or earlier
// chunks, never a slot a later chunk still reads).
static const size_t kScratchCap = (size_t) 512 << 20;

static int ensure_scratch(amr_dev *d, size_t bytes)
{
   if (bytes <= d->scratch_bytes)
      return AMR_OK;
   HIP_TRY(d, hipStreamSynchronize(d->stream));
   hipFree(d->d_scratch);
   d->d_scratch = nullptr;
   d->scratch_bytes = 0;
   HIP_TRY(d, hipMalloc(&d->d_scratch, bytes));
   d->scratch_bytes = bytes;
   return AMR_OK;
}

static int ensure_jobbuf(amr_dev *d, size_t bytes)
{
   if (bytes <= d->jobbuf_bytes)
      return AMR_OK;
   HIP_TRY(d, hipStreamSynchronize(d->stream));
   hipFree(d->d_jobbuf);
   d->d_jobbuf = nullptr;
   d->jobbuf_bytes = 0;
   HIP_TRY(d, hipMalloc(&d->d_jobbuf, bytes));
   d->jobbuf_bytes = bytes;
   return AMR_OK;
}

template <typename J>
static int refine_data(amr_dev *d, const J *jobs, int n, bool split)
{
   if (!d || n < 0 || (n && !jobs))
      return amr_set_error(d, AMR_ERR_ARG, "split/consolidate: bad arguments");
   if (!n)
      return AMR_OK;
   for (int t = 0; t < n; t++) {
      int a = split ? ((const amr_split_job *) (const void *) &jobs[t])->parent
                    : ((const amr_merge_job *) (const void *) &jobs[t])->dst;
      int b = split ? ((const amr_split_job *) (const void *) &jobs[t])->child
                    : ((const amr_merge_job *) (const void *) &jobs[t])->child;
      int o = split ? ((const amr_split_job *) (const void *) &jobs[t])->octant
                    : ((const amr_merge_job *) (const void *) &jobs[t])->octant;
      if (a < 0 || a >= d->c.max_blocks || b < 0 || b >= d->c.max_blocks || o < 0 || o > 7)
         return amr_set_error(d, AMR_ERR_ARG, "split/consolidate: bad job %d", t);
   }
   size_t per_job = (size_t) d->c.num_vars * (d->g.nx / 2) * (d->g.ny / 2) * (d->g.nz / 2) *
                    d->elem_bytes;
   int chunk = (int) std::max<size_t>(1, std::min<size_t>(n, kScratchCap / per_job));
   // A context of its own sizes both buffers once for the largest batch of
   // jobs a refinement step can have (8 per slot): growing one waits for the
   // device, stages still queued included.  Batch members, which share one
   // stream and would each hold the scratch, size them as needed.
   const size_t most = 8 * (size_t) d->c.max_blocks;
   const size_t scratch_want =
      d->owns_stream ? per_job * std::max<size_t>(chunk, std::min(most, kScratchCap / per_job))
                     : per_job * chunk;
   int rc = ensure_scratch(d, scratch_want);
   if (!rc)
      rc = ensure_jobbuf(d, sizeof(J) * (d->owns_stream ? std::max<size_t>(n, most) : (size_t) n));
   if (rc)
      return rc;
   // The job list goes through pinned staging with a copy ordered on the
   // stream, so the host does not wait here for the stages still queued
   // (the refinement's host work then overlaps them).  A stream that has
   // drained frees the whole staging buffer; else the list is appended, and
   // only when it does not fit does the host wait.
   const size_t jb = (sizeof(J) * (size_t) n + 255) / 256 * 256;
   if (d->jobstage_used && hipStreamQuery(d->stream) == hipSuccess)
      d->jobstage_used = 0;
   if (d->jobstage_used + jb > d->jobstage_bytes) {
      HIP_TRY(d, hipStreamSynchronize(d->stream));
      d->jobstage_used = 0;
      if (jb > d->jobstage_bytes) {
         if (d->h_jobstage)
            hipHostFree(d->h_jobstage);
         d->h_jobstage = nullptr;
         d->jobstage_bytes = 0;
         const size_t cap = std::max<size_t>(jb, 4 << 20);
         HIP_TRY(d, hipHostMalloc(&d->h_jobstage, cap, hipHostMallocDefault));
         d->jobstage_bytes = cap;
      }
   }
   char *stage = (char *) d->h_jobstage + d->jobstage_used;
   memcpy(stage, jobs, sizeof(J) * n);
   d->jobstage_used += jb;
   HIP_TRY(d, hipMemcpyAsync(d->d_jobbuf, stage, sizeof(J) * n, hipMemcpyHostToDevice,
                             d->stream));
   amr_span_begin(d, 3);
   for (int t0 = 0; t0 < n && !rc; t0 += chunk) {
      int m = std::min(chunk, n - t0);
      const J *dj = (const J *) d->d_jobbuf + t0;
      rc = split ? amr_launch_split(d, (const amr_split_job *) dj, m, d->d_scratch)
                 : amr_launch_merge(d, (const amr_merge_job *) dj, m, d->d_scratch);
   }
   amr_span_end(d);
   return rc;
}

int amr_split_impl(amr_dev *d, const amr_split_job *jobs, int n)
{
   if (d && amr_require_fresh(d, 0, d->c.num_vars, "split"))
      return AMR_ERR_STATE;
   return refine_data(d, jobs, n, true);
}

int amr_merge_impl(amr_dev *d, const amr_merge_job *jobs, int n)
{
   if (d && amr_require_fresh(d, 0, d->c.num_vars, "consolidate"))
      return AMR_ERR_STATE;
   return refine_data(d, jobs, n, false);
}

extern "C" {

int amr_dev_split(amr_dev *d, const amr_split_job *jobs, int n)
{
   int rc = refine_data(d, jobs, n, true);
   if (!rc && n && d->tape && d->tape->on && !d->replaying) {
      const int first = (int) d->tape->splits.size();
      d->tape->splits.insert(d->tape->splits.end(), jobs, jobs + n);
      amr_tape_record(d, AMR_OP_SPLIT, first, n);
   }
   return rc;
}

int amr_dev_consolidate(amr_dev *d, const amr_merge_job *jobs, int n)
{
   int rc = refine_data(d, jobs, n, false);
   if (!rc && n && d->tape && d->tape->on && !d->replaying) {
      const int first = (int) d->tape->merges.size();
      d->tape->merges.insert(d->tape->merges.end(), jobs, jobs + n);
      amr_tape_record(d, AMR_OP_MERGE, first, n);
   }
   return rc;
}

int amr_dev_reserve_sums(amr_dev *d, int n_rows)
{
   if (!d || n_rows < 1)
      return amr_set_error(d, AMR_ERR_ARG, "reserve_sums: n_rows must be positive");
   if (n_rows <= d->sum_rows)
      return AMR_OK;
   HIP_TRY(d, hipStreamSynchronize(d->stream));
   hipFree(d->d_sums);
   d->d_sums = nullptr;
   d->sum_rows = 0;
   HIP_TRY(d, hipMalloc(&d->d_sums, sizeof(double) * (size_t) n_rows * d->c.num_vars));
   HIP_TRY(d, hipMemset(d->d_sums, 0, sizeof(double) * (size_t) n_rows * d->c.num_vars));
   d->sum_rows = n_rows;
   d->plan_version++;
   size_t need = sizeof(double) * (size_t) n_rows * d->c.num_vars;
   if (need > d->h_pinned_bytes) {
      hipHostFree(d->h_pinned);
      d->h_pinned = nullptr;
      HIP_TRY(d, hipHostMalloc((void **) &d->h_pinned, need, hipHostMallocDefault));
      d->h_pinned_bytes = need;
   }
   return AMR_OK;
}

int amr_dev_read_sums(amr_dev *d, int row0, int n_rows, double *out)
{
   if (!d || !out || row0 < 0 || n_rows < 0 || row0 + n_rows > d->sum_rows)
      return amr_set_error(d, AMR_ERR_ARG, "read_sums: rows [%d, %d) out of range", row0,
                           row0 + n_rows);
   size_t bytes = sizeof(double) * (size_t) n_rows * d->c.num_vars;
   HIP_TRY(d, hipMemcpyAsync(d->h_pinned, d->d_sums + (size_t) row0 * d->c.num_vars, bytes,
                             hipMemcpyDeviceToHost, d->stream));
   HIP_TRY(d, hipStreamSynchronize(d->stream));
   memcpy(out, d->h_pinned, bytes);
   return AMR_OK;
}

}  // extern "C"

int amr_stage_impl(amr_dev *d, int start, int num, int stage, int permute, int sum_row)
{
   if (!d || start < 0 || num < 0 || start + num > d->c.num_vars)
      return amr_set_error(d, AMR_ERR_ARG, "stage: variables out of range");
   if (sum_row >= d->sum_rows)
      return amr_set_error(d, AMR_ERR_ARG, "stage: checksum row %d not reserved", sum_row);
   int rc = amr_comm_impl(d, start, num, stage, permute);
   if (rc)
      return rc;
   amr_span_begin(d, 1);
   rc = amr_launch_stencil(d, start, num, sum_row >= 0, stage);
   amr_span_end(d);
   if (rc || sum_row < 0)
      return rc;
   amr_span_begin(d, 2);
   rc = amr_launch_reduce_sums(d, start, num,
                               d->d_sums + (size_t) sum_row * d->c.num_vars + start);
   amr_span_end(d);
   return rc;
}

extern "C" {

int amr_dev_stage(amr_dev *d, int start, int num, int stage, int permute, int sum_row)
{
   int rc = amr_stage_impl(d, start, num, stage, permute, sum_row);
   if (!rc)
      amr_tape_record(d, AMR_OP_STAGE, start, num, stage, permute);
   return rc;
}

int amr_dev_stage_fused(amr_dev *d, int start, int num, int stage, int permute, int sum_row)
{
   if (!d)
      return AMR_ERR_ARG;
   // the plain stage where in-place reads cannot stand in for it (permuted
   // phases, --code 1/2 planes, taping and replay), and for
   // stages below ~1e8 cell updates: those are bound by launch latency,
   // where the fused kernels' longer set-up costs more than the ghost planes
   // they save (FP32 runs of 8^3-16^3 blocks with 8-40 variables measured
   // 1-7% slower fused; the 16^3, 40-variable bandwidth-scale run 15-18%
   // faster)
   const double cells = (double) d->n_active * num * d->g.nx * d->g.ny * d->g.nz;
   if (!d->two_interiors || permute || d->whole_ghosts || (d->tape && d->tape->on) ||
       d->replaying || (d->fused_mode == 1 && cells < 1.0e8))
      return amr_dev_stage(d, start, num, stage, permute, sum_row);
   if (start < 0 || num < 0 || start + num > d->c.num_vars)
      return amr_set_error(d, AMR_ERR_ARG, "stage: variables out of range");
   if (sum_row >= d->sum_rows)
      return amr_set_error(d, AMR_ERR_ARG, "stage: checksum row %d not reserved", sum_row);
   if (!num)
      return AMR_OK;
   if (!d->timing_dirs)
      amr_span_begin(d, 0);
   for (int dir = 0; dir < 3; dir++) {
      if (d->timing_dirs)
         amr_span_begin(d, kSpanExchangeX + dir);
      int rc = amr_launch_exchange_fused_dir(d, dir, start, num);
      if (d->timing_dirs)
         amr_span_end(d);
      if (rc)
         return rc;
   }
   if (!d->timing_dirs)
      amr_span_end(d);
   amr_span_begin(d, 1);
   int rc = amr_launch_stencil_fused(d, start, num, sum_row >= 0);
   amr_span_end(d);
   if (rc)
      return rc;
   for (int v = start; v < start + num; v++)
      d->stale[v] = 1;
   if (sum_row < 0)
      return AMR_OK;
   amr_span_begin(d, 2);
   rc = amr_launch_reduce_sums(d, start, num,
                               d->d_sums + (size_t) sum_row * d->c.num_vars + start);
   amr_span_end(d);
   return rc;
}

int amr_dev_sync(amr_dev *d)
{
   if (!d)
      return AMR_ERR_ARG;
   HIP_TRY(d, hipStreamSynchronize(d->stream));
   return AMR_OK;
}

int amr_dev_set_timing(amr_dev *d, int on)
{
   if (!d)
      return AMR_ERR_ARG;
   d->timing = on != 0;
   d->timing_dirs = on == 2;
   d->timing_stencil_only = on == 3;
   return AMR_OK;
}

int amr_dev_read_timing(amr_dev *d, double ms[4])
{
   if (!d || !ms)
      return AMR_ERR_ARG;
   ms[0] = ms[1] = ms[2] = ms[3] = 0.0;
   HIP_TRY(d, hipStreamSynchronize(d->stream));
   for (auto &s : d->spans) {
      float t = 0.0f;
      if (s.a && s.b && hipEventElapsedTime(&t, s.a, s.b) == hipSuccess) {
         if (s.cat >= kSpanExchangeX) {
            ms[0] += t;                                   // exchange, by direction
            d->exchange_ms[s.cat - kSpanExchangeX] += t;
         } else {
            ms[s.cat] += t;
         }
      }
      if (s.a) d->ev_free.push_back(s.a);
      if (s.b) d->ev_free.push_back(s.b);
   }
   d->spans.clear();
   return AMR_OK;
}

int amr_dev_read_exchange_timing(amr_dev *d, double ms[3])
{
   if (!d || !ms)
      return AMR_ERR_ARG;
   for (int i = 0; i < 3; i++) {
      ms[i] = d->exchange_ms[i];
      d->exchange_ms[i] = 0.0;
   }
   return AMR_OK;
}

int amr_dev_memory(const amr_dev *d, size_t *bytes, int *buffers)
{
   if (!d || !bytes)
      return AMR_ERR_ARG;
   const size_t b[] = {
      d->pool ? (size_t) d->g.var_stride * d->c.num_vars * d->elem_bytes : 0,
      d->d_sorted ? sizeof(int32_t) * d->c.max_blocks : 0,
      sizeof(amr_face_job) * (size_t) d->jobs_cap,
      d->d_partials ? sizeof(double) * (size_t) d->c.num_vars * d->c.max_blocks : 0,
      sizeof(double) * (size_t) d->sum_rows * d->c.num_vars,
      d->d_stage ? sizeof(double) * (size_t) d->g.gx * d->g.gy * d->g.gz * d->c.num_vars : 0,
      d->scratch_bytes, d->jobbuf_bytes, d->h_pinned ? d->h_pinned_bytes : 0 };
   size_t tot = 0;
   int n = 0;
   for (size_t x : b) {
      tot += x;
      n += x > 0;
   }
   *bytes = tot;
   if (buffers)
      *buffers = n;
   return AMR_OK;
}

}  // extern "C"
