// The dbl dump image of a 3D MHD/RMHD box, gathered on the device in the
// file's own order (Src/write_data.cpp:136-159 with FillOutputBuffer
// :564-607): the eight primitives over the active zones, i fastest, then
// Bx1s, Bx2s, Bx3s with nx+1, ny+1, nz+1 faces along their direction.
#include "mhd3d.hpp"

namespace gpl {

namespace {

struct DumpBox {
  int n[3];        // active zones
  int lo[3];       // device index of active zone 0
  int per[3];      // periodic direction: face n is the image of face 0
};

template <typename R>
__global__ __launch_bounds__(256)
void k_dump_gather(Box3 g, DumpBox b, const R *__restrict__ V, const double *__restrict__ Vs,
                   double *__restrict__ dst)
{
  const int nx = b.n[0], ny = b.n[1], nz = b.n[2];
  const long long ex = nx + 1, ey = ny + 1, ez = nz + 1;
  const long long t = (long long)blockIdx.x*blockDim.x + threadIdx.x;
  if (t >= ex*ey*ez) return;
  const int i = (int)(t % ex), j = (int)((t/ex) % ey), k = (int)(t/(ex*ey));
  const size_t n = (size_t)g.ncell, na = (size_t)nx*ny*nz;
  // device index of active point (a, b, c) with a <= nx etc.
  auto dev = [&](int a, int bb, int c) -> size_t {
    if (a == nx && b.per[0]) a = 0;
    if (bb == ny && b.per[1]) bb = 0;
    if (c == nz && b.per[2]) c = 0;
    return (size_t)(a + b.lo[0]) + (size_t)g.nx*((size_t)(bb + b.lo[1]) + (size_t)g.ny*(c + b.lo[2]));
  };
  if (i < nx && j < ny && k < nz) {
    const size_t d = dev(i, j, k), o = ((size_t)k*ny + j)*nx + i;
#pragma unroll
    for (int q = 0; q < 8; q++) dst[q*na + o] = (double)V[q*n + d];
  }
  size_t base = 8*na;
  if (j < ny && k < nz) dst[base + ((size_t)k*ny + j)*(nx + 1) + i] = Vs[dev(i, j, k)];
  base += (size_t)(nx + 1)*ny*nz;
  if (i < nx && k < nz) dst[base + ((size_t)k*(ny + 1) + j)*nx + i] = Vs[n + dev(i, j, k)];
  base += (size_t)nx*(ny + 1)*nz;
  if (i < nx && j < ny) dst[base + ((size_t)k*ny + j)*nx + i] = Vs[2*n + dev(i, j, k)];
}

}  // namespace

template <typename R>
void mhd3d_dump_gather(const Box3 &g, const int na[3], const int periodic[3], const R *V,
                       const double *Vs, double *dst, hipStream_t s)
{
  DumpBox b;
  for (int d = 0; d < 3; d++) { b.n[d] = na[d]; b.lo[d] = g.lo[d]; b.per[d] = periodic[d]; }
  const long long pts = (long long)(na[0] + 1)*(na[1] + 1)*(na[2] + 1);
  hipLaunchKernelGGL(k_dump_gather<R>, dim3((unsigned)((pts + 255)/256)), dim3(256), 0, s, g, b, V, Vs, dst);
}

template void mhd3d_dump_gather<double>(const Box3 &, const int *, const int *, const double *,
                                        const double *, double *, hipStream_t);
template void mhd3d_dump_gather<float>(const Box3 &, const int *, const int *, const float *,
                                       const double *, double *, hipStream_t);

}  // namespace gpl
