// Ghost-zone fill of a 3D MHD/RMHD box padded with ghost layers: the
// reference's Boundary (Src/boundary.cpp:67-303) side by side, X1_BEG ..
// X3_END, on the reference's own index ranges.  The padded box is the
// reference's array: cells 0 .. ntot-1 (two ghost layers, the active cells
// lo .. hi, two ghost layers) plus one spare layer that holds the outermost
// face (index ntot) of every staggered component.
//   periodic side: PeriodicBoundary (:502-669), images of cells and faces;
//   outflow side:  OutflowBoundary (:677-760), the edge layer's cells and
//                  tangential faces, then CT_FillStaggeredFields
//                  (Src/MHD/CT/ct_fill_mag_field.cpp:39-330) for the normal
//                  faces from div B = 0;
//   reflective side: ReflectiveBoundary (:931-1100), mirror images with the
//                  normal velocity and normal field negated (FlipSign);
// then the ghost cells' field average (:275-303 with
// Src/MHD/CT/ct_field_average.cpp:182-184).
#include "mhd3d.hpp"
#include "../../include/gpluto_amd.h"

namespace gpl {

namespace {

__device__ __forceinline__ size_t pid(const Box3 &g, int i, int j, int k)
{
  return (size_t)i + (size_t)g.nx*((size_t)j + (size_t)g.ny*k);
}

// one thread per (layer, a, b): layer 0, 1 the two ghost layers, layer 2 the
// extra normal face of an upper periodic side
template <typename R>
__global__ __launch_bounds__(256)
void k_bnd_copy(Box3 g, int side, int type, R *__restrict__ V, double *__restrict__ Vs)
{
  const int dir = side >> 1, upper = side & 1;
  const int t1 = (dir + 1) % 3, t2 = (dir + 2) % 3;
  const int P[3] = {g.nx, g.ny, g.nz};
  const int na = P[t1], nb = P[t2];
  const int t = blockIdx.x*blockDim.x + threadIdx.x;
  if (t >= 3*na*nb) return;
  const int layer = t / (na*nb), r = t - layer*na*nb, b = r / na, a = r - b*na;
  const int n = g.hi[dir] - g.lo[dir] + 1;
  const int gi = upper ? g.hi[dir] + 1 + layer : layer;          // ghost index along dir
  // reflective: the mirror image i0 - i, i0 = 2*lo - 1 or 2*hi + 1 (Src/boundary.cpp:939-945)
  const int i0 = upper ? 2*g.hi[dir] + 1 : 2*g.lo[dir] - 1;
  const int src = type == GPL_PERIODIC ? gi + (upper ? -n : n)
                : type == GPL_REFLECTIVE ? i0 - gi : (upper ? g.hi[dir] : g.lo[dir]);
  int p[3], q[3];
  p[dir] = gi; p[t1] = a; p[t2] = b;
  q[dir] = src; q[t1] = a; q[t2] = b;
  const size_t N = (size_t)g.ncell;
  const size_t cp = pid(g, p[0], p[1], p[2]), cq = pid(g, q[0], q[1], q[2]);
  const bool cell = layer < 2 && a <= na - 2 && b <= nb - 2;    // BOX_LOOP over the ghost box
  if (cell) {
    for (int nv = 0; nv < 8; nv++) {
      const R v = V[nv*N + cq];
      // FlipSign (Src/boundary.cpp:840-925): a wall negates the normal velocity and field
      V[nv*N + cp] = type == GPL_REFLECTIVE && (nv == 1 + dir || nv == 4 + dir) ? -v : v;
    }
  }
  if (layer < 2) {
    // tangential components: one more face along the component (BOX{I,J,K}_LOOP)
    if (a <= na - 1 && b <= nb - 2) Vs[t1*N + cp] = Vs[t1*N + cq];
    if (a <= na - 2 && b <= nb - 1) Vs[t2*N + cp] = Vs[t2*N + cq];
  }
  // periodic normal component: BOX_LOOP on the lower side, BOXI/J/K_LOOP (one
  // more face, the spare layer) on the upper side
  if (type == GPL_PERIODIC && a <= na - 2 && b <= nb - 2 && (layer < 2 || upper))
    Vs[dir*N + cp] = Vs[dir*N + cq];
  // reflective normal component, mirrored about the boundary face and
  // negated: face i from i0 - i + 1 on the lower side, face i + 1 from i0 - i
  // on the upper side
  if (type == GPL_REFLECTIVE && layer < 2 && a <= na - 2 && b <= nb - 2) {
    int t[3] = {p[0], p[1], p[2]}, f[3] = {q[0], q[1], q[2]};
    if (!upper) f[dir] += 1;
    else        t[dir] += 1;
    Vs[dir*N + pid(g, t[0], t[1], t[2])] = -Vs[dir*N + pid(g, f[0], f[1], f[2])];
  }
}

// one thread per transverse (a, b) of the ghost box: the outflow side's
// normal faces outward from the boundary, then the ghost cells' field average
template <typename R>
__global__ __launch_bounds__(256)
void k_bnd_fill(Box3 g, int side, int type, R *__restrict__ V, double *__restrict__ Vs)
{
  const int dir = side >> 1, upper = side & 1;
  const int t1 = (dir + 1) % 3, t2 = (dir + 2) % 3;
  const int P[3] = {g.nx, g.ny, g.nz};
  const int na = P[t1] - 1, nb = P[t2] - 1;                       // cells 0 .. ntot-1
  const int t = blockIdx.x*blockDim.x + threadIdx.x;
  if (t >= na*nb) return;
  const int b = t / na, a = t - b*na;
  const size_t N = (size_t)g.ncell;
  // face areas of the uniform Cartesian grid (Src/set_geometry.cpp:124,147,166)
  const double A[3] = {1.0*g.dy*g.dz, g.dx*1.0*g.dz, g.dx*g.dy*1.0};
  const double *B[3] = {Vs, Vs + N, Vs + 2*N};
  double *Bn = Vs + dir*N;
  if (type == GPL_OUTFLOW) {
    for (int l = 0; l < 2; l++) {
      const int i = upper ? g.hi[dir] + 1 + l : g.lo[dir] - 1 - l;
      int p[3];
      p[dir] = i; p[t1] = a; p[t2] = b;
      double dB[3];
#pragma unroll
      for (int d = 0; d < 3; d++) {
        int q[3] = {p[0], p[1], p[2]};
        q[d] += 1;
        dB[d] = A[d]*B[d][pid(g, q[0], q[1], q[2])] - A[d]*B[d][pid(g, p[0], p[1], p[2])];
      }
      // the two transverse terms in x, y, z order
      const double ta = dir == 0 ? dB[1] : dB[0], tb = dir == 2 ? dB[1] : dB[2];
      int q[3] = {p[0], p[1], p[2]};
      q[dir] += 1;
      if (!upper) Bn[pid(g, p[0], p[1], p[2])] = (A[dir]*Bn[pid(g, q[0], q[1], q[2])] + ta + tb)/A[dir];
      else        Bn[pid(g, q[0], q[1], q[2])] = (A[dir]*Bn[pid(g, p[0], p[1], p[2])] - (ta + tb))/A[dir];
    }
  }
  for (int l = 0; l < 2; l++) {
    const int i = upper ? g.hi[dir] + 1 + l : l;
    int p[3];
    p[dir] = i; p[t1] = a; p[t2] = b;
    const size_t c = pid(g, p[0], p[1], p[2]);
#pragma unroll
    for (int d = 0; d < 3; d++) {
      int q[3] = {p[0], p[1], p[2]};
      q[d] += 1;
      V[(4 + d)*N + c] = (R)(0.5*(B[d][c] + B[d][pid(g, q[0], q[1], q[2])]));
    }
  }
}

}  // namespace

template <typename R>
void mhd3d_boundary(const Box3 &g, const int type[6], R *V, double *Vs, hipStream_t s)
{
  for (int side = 0; side < 6; side++) {
    const int dir = side >> 1;
    const int P[3] = {g.nx, g.ny, g.nz};
    const int na = P[(dir + 1) % 3], nb = P[(dir + 2) % 3];
    hipLaunchKernelGGL((k_bnd_copy<R>), dim3((3*na*nb + 255)/256), dim3(256), 0, s,
                       g, side, type[side], V, Vs);
    hipLaunchKernelGGL((k_bnd_fill<R>), dim3(((na - 1)*(nb - 1) + 255)/256), dim3(256), 0, s,
                       g, side, type[side], V, Vs);
  }
}

template void mhd3d_boundary<double>(const Box3 &, const int *, double *, double *, hipStream_t);
template void mhd3d_boundary<float>(const Box3 &, const int *, float *, double *, hipStream_t);

}  // namespace gpl
