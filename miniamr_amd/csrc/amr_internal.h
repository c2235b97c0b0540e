// amr_internal.h -- device context and cell-record layout shared by the
// HIP sources of libminiamr_dev.so (gfx950).  Not part of the public ABI.
//
// Storage.  A block's cells are the reference's (nx+2)(ny+2)(nz+2) array
// with one ghost layer (ref/block.h:52), but in HBM each (variable, slot)
// owns one *record* whose parts are each contiguous:
//   I    interior [nx][ny][nz]                      (z fastest)
//   F0..F5 face ghosts, interior of the face only:  x faces [ny][nz],
//        y faces [nx][nz], z faces [nx][ny]
//   EXY  the 4 ghost lines along z where x and y are both ghosts,
//        [xs][ys][0..nz+1] (includes the 8 corners)
//   EXZ  the 4 ghost lines along y, [xs][zs][1..ny]
//   EYZ  the 4 ghost lines along x, [ys][zs][1..nx]
//   ZI, ZX, ZY  a copy of the cells at z = 1 and z = nz (zs = 0, 1):
//        ZI [zs][nx][ny] interior, ZX [zs][xs][0..ny+1] (x ghosts),
//        ZY [zs][ys][1..nx] (y ghosts).
// Every canonical cell (i, j, k) has exactly one home (home()); the Z*
// copies exist because z is the fastest axis of I: the z-direction ghost
// exchange reads other blocks' z = 1 / z = nz planes, which in I are one
// element per cache line.  Every writer of a cell at z = 1 or z = nz also
// writes its copy (store_cell()), so the copies always equal their homes.
// With this layout every access of the exchange and of the stencil is a
// contiguous run; the pool is [var][slot][record], zeroed at creation and
// never cleared, since ghost cells the exchange does not write keep what
// the slot held before (the 27-point stencil reads some of them, exactly
// as in the reference).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <array>
#include <vector>
#include "amr_dev.h"

struct Lay {
   int nx, ny, nz;          // interior sizes
   int gx, gy, gz;          // with ghosts
   int stencil;
   int oF[6];               // face planes
   int oEXY, oEXZ, oEYZ;    // edge lines
   int oZI, oZX, oZY;       // z = 1 / z = nz copies
   int oI;                  // the interior part in use (0, or the second
                            // interior of a two-interior record)
   int64_t rec;             // elements per (var, slot), padded to 256 bytes
   int64_t var_stride;      // rec * max_blocks
};

// A face job of a stage that reads same-level faces in place
// (amr_dev_stage_fused).  FJ_FULL fills the job's ghost cells as the plain
// exchange does; FJ_EXT_ONLY (a same-level y or z face whose face interior
// the stencil reads from the neighbour directly) fills only the cells
// outside the face interior: the 27-point edge extension.  alt[] redirects
// the job's reads of the source's own ghost rows that are themselves read
// in place to the block they would have been copied from: y jobs alt[0] /
// alt[1] for the source's x = 0 / x = nx+1 rows, z jobs the same plus
// alt[2] / alt[3] for the y = 0 / y = ny+1 rows; -1 reads the source.
enum { FJ_FULL = 0, FJ_EXT_ONLY = 1 };
struct FusedJob {
   amr_face_job jb;
   int32_t mode;
   int32_t alt[4];
   int32_t pad;
};

// Where a fused stage finds the cells of one ghost edge line or corner of a
// block: kEdgeSrcs entries per slot, in this order
//   0..3    lines along z, cells (x ghost, y ghost, k = 1..nz), xs*2 + ys
//   4..7    lines along y, cells (x ghost, j = 1..ny, z ghost), 4 + xs*2 + zs
//   8..11   lines along x, cells (i = 1..nx, y ghost, z ghost), 8 + ys*2 + zs
//   12..19  corners, 12 + xs*4 + ys*2 + zs
// (xs = 0 for x = 0, 1 for x = nx+1, and so on).  Cell t of a line (t = 0
// for the first) is at element slot*rec + off + t*stride of the variable,
// plus the offset of the interior copy being read for kind ES_I (an
// interior cell of slot) or of its z-plane copy for ES_ZI.  With edge
// folding an entry names the cell the 27-point extension of the exchange
// would have copied (the copy is then not made); otherwise the block's own
// ghost cell.
enum { ES_ABS = 0, ES_I = 1, ES_ZI = 2 };
struct EdgeSrc {
   int32_t slot, off, stride, kind;
};
constexpr int kEdgeSrcs = 20;

__host__ __device__ __forceinline__ int64_t edge_at(const EdgeSrc &e, const Lay &L, int t)
{
   const int part = e.kind == ES_I ? L.oI : e.kind == ES_ZI ? L.oZI : 0;
   return (int64_t) e.slot * L.rec + e.off + part + (int64_t) t * e.stride;
}

// Recorded data operations of a context (include/amr_adjoint.h): each op
// is {kind, args..., plan in force}; plans and job lists are host copies.
struct TapePlan {
   std::vector<int32_t> sorted;
   std::vector<amr_face_job> jobs;
   int32_t count[3];
};
struct Tape {
   bool on = false;
   std::vector<std::array<int32_t, 8>> ops;
   std::vector<TapePlan> plans;
   std::vector<amr_split_job> splits;
   std::vector<amr_merge_job> merges;
   int cur_plan = -1;
};

struct amr_dev {
   amr_dev_config c;
   Lay g;
   int elem_bytes;
   void *pool = nullptr;
   hipStream_t stream = nullptr;
   // refinement marking (no cell data): its own stream, so that it runs
   // beside the stages still queued on `stream` (default priority)
   hipStream_t mark_stream = nullptr;
   // pinned staging of split/consolidate job lists, copied in stream order;
   // jobstage_used bytes may still feed queued copies until the stream drains
   void *h_jobstage = nullptr;
   size_t jobstage_bytes = 0, jobstage_used = 0;
   bool owns_stream = true;         // false for batch members (batch's stream)
   int plan_version = 0;            // bumped when device plan buffers change

   int32_t *d_sorted = nullptr;     // active slots in sorted-list order
   int n_active = 0;
   amr_face_job *d_jobs = nullptr;  // x jobs, then y, then z
   int32_t count[3] = {0, 0, 0};
   int jobs_cap = 0;

   double *d_partials = nullptr;    // [num_vars][max_blocks] per-block sums
   double *d_sums = nullptr;        // checksum table [rows][num_vars]
   int sum_rows = 0;
   double *h_pinned = nullptr;      // pinned buffer for read-backs
   size_t h_pinned_bytes = 0;

   double *d_stage = nullptr;       // one slot in FP64 for upload/download
   void *d_scratch = nullptr;       // split/merge snapshot
   size_t scratch_bytes = 0;
   void *d_jobbuf = nullptr;        // split/merge job array
   size_t jobbuf_bytes = 0;
   // halo codec (codec.hip): message formats of every code, direction and
   // face case (built on first use), values per variable of each, and the
   // message list of the current call
   void *d_codec = nullptr;
   int codec_len[3 * 3 * 20 * 2] = {};
   void *d_msgs = nullptr;
   size_t msgs_bytes = 0;
   // amr_dev_compare_cells (compare.hip): host copy of the active slots and
   // the per-variable maxima
   std::vector<int32_t> h_sorted;
   unsigned long long *d_cmp = nullptr;

   // Two-interior records (unless MINIAMR_FUSED=0 at creation): a second copy of the
   // I and ZI parts, so a stage can read one and write the other.  par[v]
   // is the copy that holds variable v's current interior; oI_par / oZI_par
   // are the two copies' offsets in the record.
   bool whole_ghosts = false;       // the plan has --code 1/2 jobs (kind & AMR_FACE_GHOSTS)
   // --stencil 0 (amr_dev_set_materials): mat = num_vars / 4, the
   // coefficients a1 and a0[mat], the clamp counters and the work buffer of
   // the Jacobi updates
   int mat = 0;
   double a1 = 0.0;
   double *d_a0 = nullptr;
   unsigned long long *d_check_counts = nullptr;
   void *d_matwork = nullptr;
   size_t matwork_bytes = 0;
   // amr_dev_mark_refine's inputs and output flags
   void *mk_blocks = nullptr, *mk_parents = nullptr, *mk_objects = nullptr, *mk_levels = nullptr,
        *mk_out = nullptr;
   size_t mk_blocks_cap = 0, mk_parents_cap = 0, mk_objects_cap = 0, mk_levels_cap = 0,
          mk_out_cap = 0;
   bool two_interiors = false;
   int fused_mode = 1;              // MINIAMR_FUSED: 0 never, 1 large stages, 2 always
   int oI_par[2] = {0, 0}, oZI_par[2] = {0, 0};
   std::vector<uint8_t> par;

   // The plan of amr_dev_stage_fused, derived from the plain plan by
   // set_mesh: reduced job lists per direction, and per active block
   // (sorted order) the slots its six faces are read from in place (-1:
   // its own ghost plane).  stale[v]: variable v's in-place faces' ghost
   // planes have not been written since the last plain exchange.
   FusedJob *d_fjobs = nullptr;
   int32_t fcount[3] = {0, 0, 0};
   int fjobs_cap = 0;
   // the plain jobs of the faces a fused stage reads in place (same-level
   // jobs of blocks without a boundary face), per direction: rerun on the
   // interiors the last fused stage read, they write the ghost cells that
   // stage left unwritten (amr_require_fresh)
   amr_face_job *d_mjobs = nullptr;
   int32_t mcount[3] = {0, 0, 0};
   int mjobs_cap = 0;
   int32_t *d_redir = nullptr;
   // kEdgeSrcs entries per slot (EdgeSrc); fold_edges: the plan drops the
   // 27-point edge extensions of in-place faces (the stencil and the
   // remaining jobs read their cells where the extension would copy from)
   EdgeSrc *d_edges = nullptr;
   bool fold_edges = false;
   // pinned staging of set_mesh's plan uploads (stream-ordered copies)
   void *h_plan = nullptr;
   size_t h_plan_bytes = 0;
   std::vector<uint8_t> stale;

   Tape *tape = nullptr;            // recording of data operations, if any
   bool replaying = false;          // calls come from the tape: do not record

   bool timing = false;
   bool timing_dirs = false;        // exchange spans per direction (timing level 2)
   bool timing_stencil_only = false;   // stencil spans only (timing level 3)
   bool span_skipped = false;       // the last amr_span_begin recorded nothing
   double exchange_ms[3] = {0.0, 0.0, 0.0};   // per direction, read by read_timing
   std::vector<hipEvent_t> ev_free;
   struct Span { hipEvent_t a, b; int cat; };
   std::vector<Span> spans;

   char err[512] = {0};
};

// Storage helpers: T is double, float or uint16_t (bfloat16 bits); A is the
// accumulation type (double for FP64 storage, float otherwise).
template <typename T> struct Acc { typedef float type; };
template <> struct Acc<double> { typedef double type; };

__device__ __forceinline__ double to_acc(double v) { return v; }
__device__ __forceinline__ float to_acc(float v) { return v; }
__device__ __forceinline__ float to_acc(uint16_t v)
{
   return __uint_as_float(((uint32_t) v) << 16);
}

template <typename T> __device__ __forceinline__ T from_acc(typename Acc<T>::type v);
template <> __device__ __forceinline__ double from_acc<double>(double v) { return v; }
template <> __device__ __forceinline__ float from_acc<float>(float v) { return v; }
// float -> bfloat16, round to nearest even (NaN kept quiet).
template <> __device__ __forceinline__ uint16_t from_acc<uint16_t>(float v)
{
   uint32_t u = __float_as_uint(v);
   if ((u & 0x7fffffffu) > 0x7f800000u)
      return (uint16_t) ((u >> 16) | 0x40);
   u += 0x7fffu + ((u >> 16) & 1u);
   return (uint16_t) (u >> 16);
}

// Global-memory view of a device pointer.  A pointer that a kernel reads
// from memory (the batch's item records) is otherwise a generic address,
// accessed with flat instructions, which count against the LDS wait counter
// as well: every LDS wait would then also wait for the outstanding global
// loads.  Only for pointers into device allocations.
template <typename T> using gptr = __attribute__((address_space(1))) T *;
template <typename T> __device__ __forceinline__ gptr<T> global_view(T *p)
{
   return (gptr<T>) p;
}

// A pointer that is the same in every lane of the wavefront, moved to
// scalar registers.  Accesses then take the scalar base plus one 32-bit
// vector offset (ld / st below) instead of a 64-bit address per lane and
// per access.  Only for pointers that are wave-uniform.
template <typename T> __device__ __forceinline__ gptr<T> uniform_base(gptr<T> p)
{
   const uint64_t a = (uint64_t) p;
   const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t) a);
   const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t) (a >> 32));
   return (gptr<T>) (((uint64_t) hi << 32) | lo);
}

// Element i of a record-sized array (i * sizeof(T) < 2^32) at a base.
template <typename T> __device__ __forceinline__ T ld(gptr<const T> base, uint32_t i)
{
   return *(gptr<const T>) ((gptr<const char>) base + i * (uint32_t) sizeof(T));
}
template <typename T> __device__ __forceinline__ T ld(gptr<T> base, uint32_t i)
{
   return *(gptr<const T>) ((gptr<const char>) base + i * (uint32_t) sizeof(T));
}
template <typename T> __device__ __forceinline__ void st(gptr<T> base, uint32_t i, T v)
{
   *(gptr<T>) ((gptr<char>) base + i * (uint32_t) sizeof(T)) = v;
}

// Streaming accesses of the stencil (each element touched once per pass).
// Loads are cached: in fused stages a block's face planes are read again by
// its neighbours' workgroups, and cached FP64 loads made the stencil 4.4%
// faster there (3.399 -> 3.248 ms on the 16^3 workload; nontemporal loads
// had won by 5% before stages read faces in place).  Stores of new FP64
// values are nontemporal (whole 128-byte record rows); FP32 and BF16 rows
// are 64 / 32 bytes, a line's halves are written by different
// instructions, and nontemporal stores made them 12-22% slower, so they
// stay cached.
template <typename T> __device__ __forceinline__ T ld_stream(gptr<T> base, uint32_t i)
{
   return *(gptr<T>) ((gptr<char>) base + i * (uint32_t) sizeof(T));
}
template <typename T> __device__ __forceinline__ T ld_stream64(gptr<T> base, int64_t i)
{
   return base[i];
}
template <typename T> __device__ __forceinline__ void st_stream(gptr<T> base, uint32_t i, T v)
{
   const gptr<T> p = (gptr<T>) ((gptr<char>) base + i * (uint32_t) sizeof(T));
   if constexpr (sizeof(T) == 8)
      __builtin_nontemporal_store(v, p);
   else
      *p = v;
}

// Part offsets of a record of nx*ny*nz cells with elem_bytes per element,
// with a second I and ZI when two_interiors (oI_par[1] / oZI_par[1]; else
// both entries name the first copy).  Every part starts on a 128-byte line,
// so a 16-cell row of the interior or of a z-plane copy never straddles two
// lines (with the first ZI 64 bytes into a line, the fused FP64 stencil ran
// ~4% slower in every other stage, as the stages take the two interior
// copies in turn); the parts the stencil streams (I, F, ZI) come first.  rec is padded to 256 bytes; var_stride is left to the
// caller.
struct RecordLayout {
   Lay g;
   int oI_par[2], oZI_par[2];
};
inline RecordLayout make_record_layout(int nx, int ny, int nz, int stencil, int elem_bytes,
                                       bool two_interiors)
{
   RecordLayout r{};
   Lay &g = r.g;
   g.nx = nx; g.ny = ny; g.nz = nz;
   g.gx = nx + 2; g.gy = ny + 2; g.gz = nz + 2;
   g.stencil = stencil;
   const int line = 128 / elem_bytes;
   int o = 0;
   auto part = [&](int n) {
      o = (o + line - 1) / line * line;
      const int at = o;
      o += n;
      return at;
   };
   g.oI = part(nx * ny * nz);
   g.oF[0] = part(ny * nz);
   g.oF[1] = part(ny * nz);
   g.oF[2] = part(nx * nz);
   g.oF[3] = part(nx * nz);
   g.oF[4] = part(nx * ny);
   g.oF[5] = part(nx * ny);
   g.oZI = part(2 * nx * ny);
   g.oEXY = part(4 * g.gz);
   g.oEXZ = part(4 * ny);
   g.oEYZ = part(4 * nx);
   g.oZX = part(4 * g.gy);
   g.oZY = part(4 * nx);
   r.oI_par[0] = r.oI_par[1] = g.oI;
   r.oZI_par[0] = r.oZI_par[1] = g.oZI;
   if (two_interiors) {
      r.oI_par[1] = part(nx * ny * nz);
      r.oZI_par[1] = part(2 * nx * ny);
   }
   const int per256 = 256 / elem_bytes;
   g.rec = ((int64_t) o + per256 - 1) / per256 * per256;
   g.var_stride = 0;
   return r;
}

// Offset of canonical cell (i, j, k), 0 <= i <= nx+1 etc., in a record.
__host__ __device__ __forceinline__ int home(const Lay &L, int i, int j, int k)
{
   const int gi = (i == 0) | (i == L.nx + 1), gj = (j == 0) | (j == L.ny + 1),
             gk = (k == 0) | (k == L.nz + 1);
   const int xs = i != 0, ys = j != 0, zs = k != 0;
   switch (gi + gj + gk) {
   case 0:
      return L.oI + ((i - 1) * L.ny + (j - 1)) * L.nz + (k - 1);
   case 1:
      if (gi) return L.oF[xs] + (j - 1) * L.nz + (k - 1);
      if (gj) return L.oF[2 + ys] + (i - 1) * L.nz + (k - 1);
      return L.oF[4 + zs] + (i - 1) * L.ny + (j - 1);
   default:
      if (gi && gj) return L.oEXY + (xs * 2 + ys) * L.gz + k;
      if (gi) return L.oEXZ + (xs * 2 + zs) * L.ny + (j - 1);
      return L.oEYZ + (ys * 2 + zs) * L.nx + (i - 1);
   }
}

// Offset of the z-plane copy of cell (i, j, k) with k == 1 or k == nz.
__host__ __device__ __forceinline__ int zcopy(const Lay &L, int i, int j, int k)
{
   const int zs = k == L.nz;
   const int gi = (i == 0) | (i == L.nx + 1), gj = (j == 0) | (j == L.ny + 1);
   if (!gi && !gj) return L.oZI + (zs * L.nx + (i - 1)) * L.ny + (j - 1);
   if (gi) return L.oZX + (zs * 2 + (i != 0)) * L.gy + j;
   return L.oZY + (zs * 2 + (j != 0)) * L.nx + (i - 1);
}

// Write a cell and, on the z = 1 / z = nz planes, its copy.
template <typename T>
__device__ __forceinline__ void store_cell(T *rec, const Lay &L, int i, int j, int k, T v)
{
   rec[home(L, i, j, k)] = v;
   if (k == 1 || k == L.nz)
      rec[zcopy(L, i, j, k)] = v;
}

// Error helpers (context.hip).
int amr_set_error(amr_dev *d, int rc, const char *fmt, ...);
int amr_hip_check(amr_dev *d, hipError_t e, const char *what);
#define HIP_TRY(d, call)                                           \
   do {                                                            \
      hipError_t e_ = (call);                                      \
      if (e_ != hipSuccess) return amr_hip_check((d), e_, #call);  \
   } while (0)

// Timing spans (context.hip): category 0 exchange, 1 stencil, 2 checksum,
// 3 split/merge.
void amr_span_begin(amr_dev *d, int cat);
void amr_span_end(amr_dev *d);
// Span categories: 0 exchange (the batch's graph), 1 stencil, 2 checksum,
// 3 split/consolidate, kSpanExchangeX + dir one exchange direction.
constexpr int kSpanExchangeX = 4;

// Unrecorded implementations behind the public calls (context.hip), used
// by the tape's replay (adjoint.hip).
int amr_set_mesh_impl(amr_dev *d, const int32_t *sorted, int n_active, const amr_face_job *jobs,
                      const int32_t count[3]);
int amr_comm_impl(amr_dev *d, int start, int num, int stage, int permute);
int amr_stencil_impl(amr_dev *d, int var, int calc_stage);
int amr_stage_impl(amr_dev *d, int start, int num, int stage, int permute, int sum_row);
int amr_split_impl(amr_dev *d, const amr_split_job *jobs, int n);
int amr_merge_impl(amr_dev *d, const amr_merge_job *jobs, int n);
// Append an operation to the tape when recording (adjoint.hip).
void amr_tape_record(amr_dev *d, int kind, int a1 = 0, int a2 = 0, int a3 = 0, int a4 = 0);
void amr_tape_free(amr_dev *d);

// The layout of variables [start, start+num) with their current interior
// copy (AMR_ERR_STATE, via amr_set_error, if the range mixes copies).
int amr_lay(amr_dev *d, int start, int num, Lay *out);
// The same layout with the other interior copy (where a stage writes).
Lay amr_lay_other(const amr_dev *d, const Lay &L);

// Launchers implemented per source file; all enqueue on d->stream.
int amr_launch_exchange_dir(amr_dev *d, int dir, int start, int num);
int amr_launch_exchange_fused_dir(amr_dev *d, int dir, int start, int num);
int amr_launch_stencil(amr_dev *d, int start, int num, bool want_partials, int calc_stage);
int amr_launch_mat_stencil(amr_dev *d, int var, int calc_stage);
// Stencil of a fused stage: interiors read from the current copy, written
// to the other one (the variables' copy flips), same-level faces read in
// place.
int amr_launch_stencil_fused(amr_dev *d, int start, int num, bool want_partials);
// Brings the ghost planes of variables [start, start+num) that an
// amr_dev_stage_fused left unwritten (stale) up to date: they are set to what
// that stage's plain exchange would have written.
int amr_require_fresh(amr_dev *d, int start, int num, const char *what);
// Plain exchange jobs jobs[0..n) of direction-ordered phases over variables
// [start, start+num), reading and writing the copy L describes (exchange.hip).
int amr_launch_exchange_jobs(amr_dev *d, const amr_face_job *jobs, int n, const Lay &L,
                             int start, int num);
int amr_launch_block_sums(amr_dev *d, int start, int num);
int amr_launch_reduce_sums(amr_dev *d, int start, int num, double *dst_row);
int amr_launch_split(amr_dev *d, const amr_split_job *d_jobs, int n, void *scratch);
// Looks up this context's device copy of the static scratch array of
// amr_dev_split_noalloc, once per device (noalloc.hip).
int amr_noalloc_init(amr_dev *d);
int amr_launch_merge(amr_dev *d, const amr_merge_job *d_jobs, int n, void *scratch);
int amr_launch_convert_in(amr_dev *d, int slot);
int amr_launch_convert_out(amr_dev *d, int slot);
