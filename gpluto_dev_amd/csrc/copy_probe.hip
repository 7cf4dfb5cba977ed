// Byte-counter calibration for profiling: one plain device copy of a known
// size.  rocprofv3's FETCH_SIZE / WRITE_SIZE read on this kernel give the
// factor between the counters and real HBM bytes (src read once, dst written
// once).  Built as its own library (libgpl_probe.so); no stage-update code
// depends on it.
#include <hip/hip_runtime.h>
#include <cstddef>

__global__ __launch_bounds__(256)
void k_fill_probe(double *__restrict__ a, size_t n, double v)
{
  for (size_t i = blockIdx.x*(size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x*blockDim.x)
    a[i] = v + (double)(i & 1023);
}

__global__ __launch_bounds__(256)
void k_copy_probe(const double *__restrict__ src, double *__restrict__ dst, size_t n)
{
  for (size_t i = blockIdx.x*(size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x*blockDim.x)
    dst[i] = src[i];
}

// Copies nbytes (rounded down to doubles) once; returns 0 or a hipError_t.
extern "C" int gpl_probe_copy(size_t nbytes)
{
  const size_t n = nbytes/sizeof(double);
  if (n == 0) return (int)hipErrorInvalidValue;
  double *a = nullptr, *b = nullptr;
  hipError_t e = hipMalloc((void **)&a, n*sizeof(double));
  if (e != hipSuccess) return (int)e;
  e = hipMalloc((void **)&b, n*sizeof(double));
  if (e != hipSuccess) { (void)hipFree(a); return (int)e; }
  const int blocks = 256*8*4;          // grid-stride: 4 blocks of 256 per SIMD
  hipLaunchKernelGGL(k_fill_probe, dim3(blocks), dim3(256), 0, 0, a, n, 1.0);
  e = hipDeviceSynchronize();
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(256), 0, 0, a, b, n);
    e = hipDeviceSynchronize();
  }
  (void)hipFree(a);
  (void)hipFree(b);
  return (int)e;
}
