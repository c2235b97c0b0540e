// amr_internal.h -- device context and helpers shared by the HIP sources of
// libminiamr_dev.so (gfx950).  Not part of the public ABI (include/amr_dev.h).
//
// Cell pool layout in HBM: one allocation, structure of arrays
//   pool[var][slot][x][y][z]     (z fastest, one ghost layer per side)
// Each (var, slot) slab holds gx*gy*gz values (g* = n* + 2) padded to a
// multiple of 256 bytes, so every slab starts on its own 256-byte boundary
// and one variable's slabs are contiguous: a stage that works on one
// variable streams one dense region.  The pool is zeroed at creation and
// slots are never cleared afterwards, because ghost cells that the exchange
// does not write keep whatever the slot held before -- the 27-point stencil
// reads some of them, exactly as the reference does.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <vector>
#include "amr_dev.h"

struct Geo {
   int nx, ny, nz;          // interior sizes
   int gx, gy, gz;          // with ghosts
   int64_t slab;            // elements per (var, slot), padded
   int64_t var_stride;      // elements per variable = slab * max_blocks
   int stencil;
};

struct amr_dev {
   amr_dev_config c;
   Geo g;
   int elem_bytes;
   void *pool = nullptr;
   hipStream_t stream = nullptr;

   int32_t *d_sorted = nullptr;     // active slots in sorted-list order
   int n_active = 0;
   amr_face_job *d_jobs = nullptr;  // x jobs, then y, then z
   int32_t count[3] = {0, 0, 0};
   int jobs_cap = 0;

   double *d_partials = nullptr;    // [num_vars][max_blocks] per-block sums
   double *d_sums = nullptr;        // checksum table [rows][num_vars]
   int sum_rows = 0;
   double *h_pinned = nullptr;      // small pinned buffer for read-backs
   size_t h_pinned_bytes = 0;

   double *d_stage = nullptr;       // one slot in FP64 for upload/download
   void *d_scratch = nullptr;       // split/merge snapshot
   size_t scratch_bytes = 0;
   void *d_jobbuf = nullptr;        // split/merge job array
   size_t jobbuf_bytes = 0;

   bool timing = false;
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

__device__ __forceinline__ int64_t cell_index(const Geo &g, int var, int slot, int i, int j,
                                              int k)
{
   return (int64_t) var * g.var_stride + (int64_t) slot * g.slab +
          ((int64_t) i * g.gy + j) * g.gz + k;
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

// Launchers implemented per source file; all enqueue on d->stream.
int amr_launch_exchange_dir(amr_dev *d, int dir, int start, int num);
int amr_launch_stencil(amr_dev *d, int start, int num, bool want_partials);
int amr_launch_block_sums(amr_dev *d, int start, int num);
int amr_launch_reduce_sums(amr_dev *d, int start, int num, double *dst_row);
int amr_launch_split(amr_dev *d, const amr_split_job *d_jobs, int n, void *scratch);
int amr_launch_merge(amr_dev *d, const amr_merge_job *d_jobs, int n, void *scratch);
int amr_launch_convert_in(amr_dev *d, int slot);
int amr_launch_convert_out(amr_dev *d, int slot);
