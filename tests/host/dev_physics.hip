// Device build of the pointwise physics (gpluto_dev_amd/csrc/physics.hpp) and
// of the correctly rounded acos/sincos (csrc/crmath.hpp) for tests: batched
// evaluations on the GPU, one item per thread, compared with the host build
// of the same source (host_physics.cpp, host_crmath.cpp).  Test
// infrastructure only; built with the FP64 flags of the library.
#include "../../gpluto_dev_amd/csrc/physics.hpp"

#include <vector>

using namespace gpl;

namespace {

__global__ void __launch_bounds__(256) k_crmath(const double *x, int n, int mode, double *o1, double *o2)
{
  const int q = blockIdx.x*blockDim.x + threadIdx.x;
  if (q >= n) return;
  if (mode == 0) {
    o1[q] = cr::acos_cr(x[q]);
  } else {
    double s, c;
    cr::sincos_cr(x[q], &s, &c);
    o1[q] = s;
    o2[q] = c;
  }
}

template <int N>
__device__ void face_n(int phys, int solver, const double *vl, const double *vr, double gamma,
                       FaceFlux<double> &ff)
{
  if (phys == 3) {
    if (solver == 3)      rmhd_hlld<N>(vl, vr, gamma, 1.e-12, 1.e-12, ff);
    else if (solver == 2) rmhd_hllc<N>(vl, vr, gamma, ff);
    else                  rmhd_hll<N>(vl, vr, gamma, ff);
  } else if (solver == 3) {
    mhd_hlld<double, N>(vl, vr, gamma, ff);
  } else if (solver == 2) {
    mhd_hllc<double, N>(vl, vr, gamma, ff);
  } else {
    mhd_hll<double, N>(vl, vr, gamma, ff);
  }
}

__global__ void __launch_bounds__(256) k_face(int phys, int solver, const double *vl, const double *vr, int n, int dir,
                       double gamma, double *flux, double *sl, double *sr, uint32_t *err)
{
  const int q = blockIdx.x*blockDim.x + threadIdx.x;
  if (q >= n) return;
  FaceFlux<double> ff;
  const double *a = vl + 8*(size_t)q, *b = vr + 8*(size_t)q;
  if (dir == 0) face_n<0>(phys, solver, a, b, gamma, ff);
  else if (dir == 1) face_n<1>(phys, solver, a, b, gamma, ff);
  else face_n<2>(phys, solver, a, b, gamma, ff);
  for (int v = 0; v < 8; v++) flux[8*(size_t)q + v] = ff.f[v];
  sl[q] = ff.Sl;
  sr[q] = ff.Sr;
  err[q] = ff.err;
}

// the RMHD HLLC solver alone in its kernel (the solver of the stage kernels
// is instantiated alone per kernel too)
template <int N>
__global__ void __launch_bounds__(256) k_face_rmhd_hllc(const double *vl, const double *vr, int n, double gamma,
                                 double *flux, double *sl, double *sr, uint32_t *err)
{
  const int q = blockIdx.x*blockDim.x + threadIdx.x;
  if (q >= n) return;
  double a[8], b[8];
  for (int v = 0; v < 8; v++) { a[v] = vl[8*(size_t)q + v]; b[v] = vr[8*(size_t)q + v]; }
  FaceFlux<double> ff;
  rmhd_hllc<N>(a, b, gamma, ff);
  for (int v = 0; v < 8; v++) flux[8*(size_t)q + v] = ff.f[v];
  sl[q] = ff.Sl;
  sr[q] = ff.Sr;
  err[q] = ff.err;
}

__global__ void __launch_bounds__(256) k_c2p(double *u, double *v, int n, double gamma, uint32_t *err)
{
  const int q = blockIdx.x*blockDim.x + threadIdx.x;
  if (q >= n) return;
  double uu[8], vv[8];
  for (int a = 0; a < 8; a++) uu[a] = u[8*(size_t)q + a];
  uint32_t local = 0;
  rmhd_cons_to_prim(uu, vv, gamma, 1.e-12, 1.e-12, local);
  for (int a = 0; a < 8; a++) { u[8*(size_t)q + a] = uu[a]; v[8*(size_t)q + a] = vv[a]; }
  err[q] = local;
}

// device buffers for one call: copies in, runs, copies out, frees
struct Dev {
  std::vector<void *> ptrs;
  hipError_t rc = hipSuccess;
  template <typename T> T *in(const T *h, size_t cnt)
  {
    void *p = nullptr;
    if (rc == hipSuccess) rc = hipMalloc(&p, cnt*sizeof(T) + 8);
    if (rc == hipSuccess && h) rc = hipMemcpy(p, h, cnt*sizeof(T), hipMemcpyHostToDevice);
    ptrs.push_back(p);
    return (T *)p;
  }
  template <typename T> void out(T *h, const T *d, size_t cnt)
  {
    if (rc == hipSuccess) rc = hipMemcpy(h, d, cnt*sizeof(T), hipMemcpyDeviceToHost);
  }
  ~Dev() { for (void *p : ptrs) if (p) (void)hipFree(p); }
};

inline int blocks(int n) { return (n + 255)/256; }

}  // namespace

extern "C" int dp_crmath(const double *x, int n, int mode, double *o1, double *o2)
{
  if (n <= 0) return 0;
  Dev d;
  const double *dx = d.in(x, n);
  double *d1 = d.in<double>(nullptr, n), *d2 = d.in<double>(nullptr, n);
  if (d.rc != hipSuccess) return (int)d.rc;
  hipLaunchKernelGGL(k_crmath, dim3(blocks(n)), dim3(256), 0, 0, dx, n, mode, d1, d2);
  d.rc = hipDeviceSynchronize();
  d.out(o1, d1, n);
  if (mode != 0) d.out(o2, d2, n);
  return (int)d.rc;
}

extern "C" int dp_face(int phys, int solver, const double *vl, const double *vr, int n, int dir,
                       double gamma, double *flux, double *sl, double *sr, uint32_t *err)
{
  if (n <= 0) return 0;
  Dev d;
  const double *a = d.in(vl, 8*(size_t)n), *b = d.in(vr, 8*(size_t)n);
  double *f = d.in<double>(nullptr, 8*(size_t)n), *l = d.in<double>(nullptr, n);
  double *r = d.in<double>(nullptr, n);
  uint32_t *e = d.in<uint32_t>(nullptr, n);
  if (d.rc != hipSuccess) return (int)d.rc;
  hipLaunchKernelGGL(k_face, dim3(blocks(n)), dim3(256), 0, 0, phys, solver, a, b, n, dir, gamma,
                     f, l, r, e);
  d.rc = hipDeviceSynchronize();
  d.out(flux, f, 8*(size_t)n);
  d.out(sl, l, n);
  d.out(sr, r, n);
  d.out(err, e, n);
  return (int)d.rc;
}

extern "C" int dp_face_rmhd_hllc(const double *vl, const double *vr, int n, int dir, double gamma,
                                 double *flux, double *sl, double *sr, uint32_t *err)
{
  if (n <= 0) return 0;
  Dev d;
  const double *a = d.in(vl, 8*(size_t)n), *b = d.in(vr, 8*(size_t)n);
  double *f = d.in<double>(nullptr, 8*(size_t)n), *l = d.in<double>(nullptr, n);
  double *r = d.in<double>(nullptr, n);
  uint32_t *e = d.in<uint32_t>(nullptr, n);
  if (d.rc != hipSuccess) return (int)d.rc;
  if (dir == 0) hipLaunchKernelGGL(k_face_rmhd_hllc<0>, dim3(blocks(n)), dim3(256), 0, 0, a, b, n, gamma, f, l, r, e);
  else if (dir == 1) hipLaunchKernelGGL(k_face_rmhd_hllc<1>, dim3(blocks(n)), dim3(256), 0, 0, a, b, n, gamma, f, l, r, e);
  else hipLaunchKernelGGL(k_face_rmhd_hllc<2>, dim3(blocks(n)), dim3(256), 0, 0, a, b, n, gamma, f, l, r, e);
  d.rc = hipDeviceSynchronize();
  d.out(flux, f, 8*(size_t)n);
  d.out(sl, l, n);
  d.out(sr, r, n);
  d.out(err, e, n);
  return (int)d.rc;
}

extern "C" int dp_rmhd_c2p(double *u, double *v, int n, double gamma, uint32_t *err)
{
  if (n <= 0) return 0;
  Dev d;
  double *du = d.in(u, 8*(size_t)n), *dv = d.in<double>(nullptr, 8*(size_t)n);
  uint32_t *e = d.in<uint32_t>(nullptr, n);
  if (d.rc != hipSuccess) return (int)d.rc;
  hipLaunchKernelGGL(k_c2p, dim3(blocks(n)), dim3(256), 0, 0, du, dv, n, gamma, e);
  d.rc = hipDeviceSynchronize();
  d.out(u, du, 8*(size_t)n);
  d.out(v, dv, 8*(size_t)n);
  d.out(err, e, n);
  return (int)d.rc;
}
