// Ghost-zone fill of the batched 2D MHD boxes with an outflow or reflective
// side: the reference's Boundary (Src/boundary.cpp:67-303) side by side,
// X1_BEG, X1_END, X2_BEG, X2_END, on every box of the batch in one launch per
// step of a side.  A box is the reference's array for DIMENSIONS 2: cells
// 0 .. ntot-1 along each direction (two ghost layers, the active cells
// lo .. hi, two ghost layers) plus one spare layer holding the outermost face
// (index ntot) of each staggered component.  B1 and B2 live on faces, B3 is
// a cell variable.
//   periodic side: PeriodicBoundary (:502-669), images of cells and faces;
//   outflow side:  OutflowBoundary (:677-760), the edge layer's cells and
//                  tangential faces, then CT_FillStaggeredFields
//                  (Src/MHD/CT/ct_fill_mag_field.cpp:39-330) for the normal
//                  faces from div B = 0;
//   reflective side: ReflectiveBoundary (:931-1100), mirror images with the
//                  normal velocity and normal field negated (FlipSign :840-925);
// then the ghost cells' field average of B1, B2 (:275-303 with
// Src/MHD/CT/ct_field_average.cpp:182-184).
#include "batch2d.hpp"
#include "../../include/gpluto_amd.h"

namespace gpl {

namespace {

__device__ __forceinline__ size_t bid(const BoxDesc &d, int i, int j)
{
  return (size_t)d.off + (size_t)i + (size_t)d.nx*j;
}

// one thread per (layer, a) of one box (blockIdx.y): layer 0, 1 the two ghost
// layers, layer 2 the extra normal face of an upper periodic side; a runs
// along the other direction
template <typename R>
__global__ __launch_bounds__(256)
void k2_bnd_copy(const BoxDesc *__restrict__ box, int side, int type, R *__restrict__ V,
                 double *__restrict__ Vs, size_t N)
{
  const BoxDesc d = box[blockIdx.y];
  const int dir = side >> 1, upper = side & 1, tr = 1 - dir;
  const int P[2] = {d.nx, d.ny};
  const int na = P[tr];
  const int t = blockIdx.x*blockDim.x + threadIdx.x;
  if (t >= 3*na) return;
  const int layer = t / na, a = t - layer*na;
  const int n = d.hi[dir] - d.lo[dir] + 1;
  const int gi = upper ? d.hi[dir] + 1 + layer : layer;          // ghost index along dir
  // reflective: the mirror image i0 - i, i0 = 2*lo - 1 or 2*hi + 1 (Src/boundary.cpp:939-945)
  const int i0 = upper ? 2*d.hi[dir] + 1 : 2*d.lo[dir] - 1;
  const int src = type == GPL_PERIODIC ? gi + (upper ? -n : n)
                : type == GPL_REFLECTIVE ? i0 - gi : (upper ? d.hi[dir] : d.lo[dir]);
  int p[2], q[2];
  p[dir] = gi; p[tr] = a;
  q[dir] = src; q[tr] = a;
  const size_t cp = bid(d, p[0], p[1]), cq = bid(d, q[0], q[1]);
  if (layer < 2 && a <= na - 2) {                                 // BOX_LOOP over the ghost box
    for (int nv = 0; nv < 8; nv++) {
      const R v = V[nv*N + cq];
      V[nv*N + cp] = type == GPL_REFLECTIVE && (nv == 1 + dir || nv == 4 + dir) ? -v : v;
    }
  }
  // the tangential component: one more face along it (BOX{I,J}_LOOP)
  if (layer < 2) Vs[tr*N + cp] = Vs[tr*N + cq];
  // periodic normal component: BOX_LOOP on the lower side, one more face
  // (the spare layer) on the upper side
  if (type == GPL_PERIODIC && a <= na - 2 && (layer < 2 || upper)) Vs[dir*N + cp] = Vs[dir*N + cq];
  // reflective normal component, mirrored about the boundary face and
  // negated: face i from i0 - i + 1 on the lower side, face i + 1 from i0 - i
  // on the upper side
  if (type == GPL_REFLECTIVE && layer < 2 && a <= na - 2) {
    int w[2] = {p[0], p[1]}, f[2] = {q[0], q[1]};
    if (!upper) f[dir] += 1;
    else        w[dir] += 1;
    Vs[dir*N + bid(d, w[0], w[1])] = -Vs[dir*N + bid(d, f[0], f[1])];
  }
}

// one thread per cell a across the side of one box (blockIdx.y): the outflow
// side's normal faces outward from the boundary face, then the ghost cells'
// field average
template <typename R>
__global__ __launch_bounds__(256)
void k2_bnd_fill(const BoxDesc *__restrict__ box, int side, int type, R *__restrict__ V,
                 double *__restrict__ Vs, size_t N)
{
  const BoxDesc d = box[blockIdx.y];
  const int dir = side >> 1, upper = side & 1, tr = 1 - dir;
  const int P[2] = {d.nx, d.ny};
  const int na = P[tr] - 1;                                       // cells 0 .. ntot-1
  const int a = blockIdx.x*blockDim.x + threadIdx.x;
  if (a >= na) return;
  // face areas of the uniform Cartesian grid in 2D (Src/set_geometry.cpp:124,147)
  const double A[2] = {1.0*d.dy, d.dx*1.0};
  const double *B[2] = {Vs, Vs + N};
  double *Bn = Vs + dir*N;
  if (type == GPL_OUTFLOW) {
    for (int l = 0; l < 2; l++) {
      const int i = upper ? d.hi[dir] + 1 + l : d.lo[dir] - 1 - l;
      int p[2], q[2];
      p[dir] = i; p[tr] = a;
      q[dir] = i; q[tr] = a + 1;
      // the transverse term, then the missing direction's 0.0 (x, y, z order)
      const double ta = A[tr]*B[tr][bid(d, q[0], q[1])] - A[tr]*B[tr][bid(d, p[0], p[1])];
      const double tb = 0.0;
      q[dir] = i + 1; q[tr] = a;
      if (!upper) Bn[bid(d, p[0], p[1])] = (A[dir]*Bn[bid(d, q[0], q[1])] + ta + tb)/A[dir];
      else        Bn[bid(d, q[0], q[1])] = (A[dir]*Bn[bid(d, p[0], p[1])] - (ta + tb))/A[dir];
    }
  }
  for (int l = 0; l < 2; l++) {
    const int i = upper ? d.hi[dir] + 1 + l : l;
    int p[2];
    p[dir] = i; p[tr] = a;
    const size_t c = bid(d, p[0], p[1]);
    V[4*N + c] = (R)(0.5*(B[0][c] + B[0][bid(d, p[0] + 1, p[1])]));
    V[5*N + c] = (R)(0.5*(B[1][c] + B[1][bid(d, p[0], p[1] + 1)]));
  }
}

}  // namespace

template <typename R>
void batch2d_boundary(const BoxDesc *box, int nbox, int maxp, const int type[4], R *V, double *Vs,
                      size_t ncell, hipStream_t s)
{
  for (int side = 0; side < 4; side++) {
    hipLaunchKernelGGL((k2_bnd_copy<R>), dim3((3*maxp + 255)/256, nbox), dim3(256), 0, s,
                       box, side, type[side], V, Vs, ncell);
    hipLaunchKernelGGL((k2_bnd_fill<R>), dim3((maxp + 255)/256, nbox), dim3(256), 0, s,
                       box, side, type[side], V, Vs, ncell);
  }
}

template void batch2d_boundary<double>(const BoxDesc *, int, int, const int *, double *, double *,
                                       size_t, hipStream_t);
template void batch2d_boundary<float>(const BoxDesc *, int, int, const int *, float *, double *,
                                      size_t, hipStream_t);

}  // namespace gpl
