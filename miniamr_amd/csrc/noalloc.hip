// noalloc.hip -- amr_dev_split_noalloc: block split without device
// allocation, for callers that cannot call hipMalloc (code under HIP graph
// capture, serving loops, harnesses that forbid it).
//
// amr_dev_split stages its job list and the parent-octant snapshot in
// buffers the context grows with hipMalloc.  This variant stages both in
// one file-scope device array, `scratch`, of kScratchBytes per GPU (held by
// every GPU the library runs on until the process ends):
//   [0, kJobBytes)             the job list (at most kMaxJobs jobs);
//   [kJobBytes, kScratchBytes) parent-octant snapshots, one chunk of jobs
//                              at a time, in job order (the chunking of
//                              amr_dev_split, so the result is the same).
// Host code never names `scratch` as an address: amr_noalloc_init, called
// by amr_dev_create with the context's device current, looks up each
// device's copy once with hipGetSymbolAddress and keeps it per device.
// Every call on a GPU shares that GPU's array; nothing in it has to start
// at zero (the job list and each snapshot are written before they are
// read).  A call returns after its kernels have finished, so calls made one
// after another from one thread never overlap; calls that overlap on one
// GPU (from several threads, on any contexts of that GPU) race.
#include <mutex>

#include "amr_internal.h"

namespace {

constexpr size_t kScratchBytes = (size_t) 16 << 20;  // 16 MiB per GPU
constexpr size_t kJobBytes = (size_t) 1 << 20;       // job list staging
constexpr int kMaxJobs = (int) (kJobBytes / sizeof(amr_split_job));
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mutex;
void *g_scratch[kMaxDevices];   // device address of `scratch` per device

}  // namespace

int amr_noalloc_init(amr_dev *d)
{
   const int dev = d->c.device;
   if (dev < 0 || dev >= kMaxDevices)
      return amr_set_error(d, AMR_ERR_NODEV, "device %d: static scratch is kept for devices "
                           "0..%d", dev, kMaxDevices - 1);
   std::lock_guard<std::mutex> lock(g_mutex);
   if (g_scratch[dev])
      return AMR_OK;
   void *ptr = nullptr;
   hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
   if (e != hipSuccess || !ptr)
      return amr_set_error(d, AMR_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                           e != hipSuccess ? hipGetErrorString(e) : "null address");
   g_scratch[dev] = ptr;
   return AMR_OK;
}

extern "C" int amr_dev_split_noalloc(amr_dev *d, const amr_split_job *jobs, int n)
{
   if (!d || n < 0 || (n && !jobs))
      return amr_set_error(d, AMR_ERR_ARG, "split_noalloc: bad arguments");
   if (!n)
      return AMR_OK;
   if (d->tape && d->tape->on && !d->replaying)
      return amr_set_error(d, AMR_ERR_STATE, "split_noalloc: not while the tape records "
                           "(use amr_dev_split)");
   if (amr_require_fresh(d, 0, d->c.num_vars, "split_noalloc"))
      return AMR_ERR_STATE;
   if (n > kMaxJobs)
      return amr_set_error(d, AMR_ERR_ARG, "split_noalloc: %d jobs exceed the static job list "
                           "(%d)", n, kMaxJobs);
   for (int t = 0; t < n; t++) {
      const amr_split_job &j = jobs[t];
      if (j.parent < 0 || j.parent >= d->c.max_blocks || j.child < 0 ||
          j.child >= d->c.max_blocks || j.octant < 0 || j.octant > 7)
         return amr_set_error(d, AMR_ERR_ARG, "split_noalloc: bad job %d", t);
   }
   const size_t per_job = (size_t) d->c.num_vars * (d->g.nx / 2) * (d->g.ny / 2) *
                          (d->g.nz / 2) * d->elem_bytes;
   const size_t room = kScratchBytes - kJobBytes;
   if (per_job > room)
      return amr_set_error(d, AMR_ERR_ARG, "split_noalloc: one parent octant of all variables "
                           "is %zu bytes, more than the %zu-byte static scratch", per_job, room);
   void *ptr = g_scratch[d->c.device];
   if (!ptr)
      return amr_set_error(d, AMR_ERR_STATE, "split_noalloc: static scratch of device %d not "
                           "initialised", d->c.device);
   const int chunk = (int) std::min<size_t>((size_t) n, room / per_job);
   // earlier work of this context may still read a previous call's jobs
   HIP_TRY(d, hipStreamSynchronize(d->stream));
   HIP_TRY(d, hipMemcpyAsync(ptr, jobs, sizeof(amr_split_job) * n, hipMemcpyHostToDevice,
                             d->stream));
   const amr_split_job *dj = (const amr_split_job *) ptr;
   void *snap = (unsigned char *) ptr + kJobBytes;
   int rc = AMR_OK;
   for (int t0 = 0; t0 < n && !rc; t0 += chunk)
      rc = amr_launch_split(d, dj + t0, std::min(chunk, n - t0), snap);
   HIP_TRY(d, hipStreamSynchronize(d->stream));
   return rc;
}
