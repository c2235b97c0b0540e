// copy_probe.hip -- HBM calibration for the roofline figures: achieved
// bandwidth of a plain streaming copy and read on this device, and the
// reference point for scaling the FETCH_SIZE / WRITE_SIZE PMC counters
// (a copy of known size: 4 GiB read + 4 GiB written).
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
__global__ void copy_k(const double4* __restrict__ a, double4* __restrict__ b, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) b[i] = a[i];
}
__global__ void read_k(const double4* __restrict__ a, double* out, size_t n) {
  double s = 0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) { double4 v = a[i]; s += v.x + v.y + v.z + v.w; }
  if (s == 12345.678) out[0] = s;
}
int main() {
  size_t bytes = (size_t)4 << 30;  // 4 GiB each
  size_t n = bytes / sizeof(double4);
  double4 *a, *b; double* o;
  if (hipMalloc(&a, bytes) || hipMalloc(&b, bytes) || hipMalloc(&o, 64)) { printf("alloc failed\n"); return 1; }
  hipMemset(a, 0, bytes); hipMemset(b, 0, bytes);
  hipEvent_t e0, e1; hipEventCreate(&e0); hipEventCreate(&e1);
  int grid = 256 * 8 * 4;
  for (int rep = 0; rep < 3; rep++) {
    hipEventRecord(e0); copy_k<<<grid, 256>>>(a, b, n); hipEventRecord(e1); hipEventSynchronize(e1);
    float ms; hipEventElapsedTime(&ms, e0, e1);
    printf("copy kernel: %.3f ms, %.1f GB/s (read+write)\n", ms, 2.0 * bytes / ms / 1e6);
    hipEventRecord(e0); read_k<<<grid, 256>>>(a, o, n); hipEventRecord(e1); hipEventSynchronize(e1);
    hipEventElapsedTime(&ms, e0, e1);
    printf("read kernel: %.3f ms, %.1f GB/s\n", ms, 1.0 * bytes / ms / 1e6);
    hipEventRecord(e0); hipMemcpyAsync(b, a, bytes, hipMemcpyDeviceToDevice); hipEventRecord(e1); hipEventSynchronize(e1);
    hipEventElapsedTime(&ms, e0, e1);
    printf("hipMemcpy D2D: %.3f ms, %.1f GB/s (read+write)\n", ms, 2.0 * bytes / ms / 1e6);
  }
  // small working set (fits the last-level cache): 128 MiB copy repeated
  size_t sb = (size_t)128 << 20, sn = sb / sizeof(double4);
  for (int rep = 0; rep < 3; rep++) {
    hipEventRecord(e0); for (int r = 0; r < 20; r++) copy_k<<<grid, 256>>>(a, b, sn); hipEventRecord(e1); hipEventSynchronize(e1);
    float ms; hipEventElapsedTime(&ms, e0, e1);
    printf("copy 128MiB x20: %.3f ms, %.1f GB/s (read+write)\n", ms, 20 * 2.0 * sb / ms / 1e6);
  }
  return 0;
}
