// Host build of the device physics (gpluto_dev_amd/csrc/physics.hpp) for
// tests: the same source the GPU kernels compile, evaluated on the CPU
// against the host libm, so that any difference from the oracle is an
// arithmetic difference and not a math-library one.
#include "../../gpluto_dev_amd/csrc/physics.hpp"

using namespace gpl;

extern "C" uint32_t hp_rmhd_face(const double *vl, const double *vr, int dir, double gamma,
                                 double *flux, double *sl, double *sr)
{
  FaceFlux<double> ff;
  if (dir == 0) rmhd_hll<0>(vl, vr, gamma, ff);
  else if (dir == 1) rmhd_hll<1>(vl, vr, gamma, ff);
  else rmhd_hll<2>(vl, vr, gamma, ff);
  for (int q = 0; q < 8; q++) flux[q] = ff.f[q];
  *sl = ff.Sl; *sr = ff.Sr;
  return ff.err;
}

template <int N>
static void face_n(int phys, int solver, const double *vl, const double *vr, double gamma,
                   FaceFlux<double> &ff)
{
  if (phys == 3) {                       // RMHD (floors: the default 1e-12)
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

extern "C" uint32_t hp_face(int phys, int solver, const double *vl, const double *vr, int dir,
                            double gamma, double *flux, double *sl, double *sr)
{
  FaceFlux<double> ff;
  if (dir == 0) face_n<0>(phys, solver, vl, vr, gamma, ff);
  else if (dir == 1) face_n<1>(phys, solver, vl, vr, gamma, ff);
  else face_n<2>(phys, solver, vl, vr, gamma, ff);
  for (int q = 0; q < 8; q++) flux[q] = ff.f[q];
  *sl = ff.Sl; *sr = ff.Sr;
  return ff.err;
}

extern "C" uint32_t hp_rmhd_c2p(double *u, double *v, double gamma, double small_rho,
                                double small_prs)
{
  uint32_t local = 0;
  rmhd_cons_to_prim(u, v, gamma, small_rho, small_prs, local);
  return local;
}
