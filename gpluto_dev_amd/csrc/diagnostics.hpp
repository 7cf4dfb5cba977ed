// On-device diagnostics of a stage-update state: the reference's DataMinMax
// (Src/main.cpp:752-860) and CT_CheckDivB (Src/MHD/CT/ct_update.cpp:580-635)
// as two-pass reductions, so that a caller checks a run without downloading
// the state.
#pragma once
#include <hip/hip_runtime.h>

namespace gpl {

// Reduction inputs of one periodic box (ghost-free SoA, i fastest).
struct DiagBox {
  int nx, ny, nz, ncell;
  int nv;          // primitives per cell (HD 5, MHD/RMHD 8); the pressure is the last
  double ax, ay, az;   // face areas of the uniform Cartesian grid
  int lo[3], hi[3];    // the active cells (a padded box's ghost layers are skipped)
  int nb;              // staggered components: 3, or 2 in 2D (div B without a z term)
};

// Number of partial results the first pass writes (each 2*nv + 2 doubles).
constexpr int DIAG_BLOCKS = 1024;
constexpr int DIAG_MAXQ = 2*8 + 2;

// out[0..nv) = min, out[nv..2nv) = max of every primitive, out[2nv] = the
// largest Mach number, out[2nv+1] = the largest |div B| (0 without faces).
// part: DIAG_BLOCKS*DIAG_MAXQ doubles of device scratch; out: device memory.
template <typename R>
void launch_diagnostics(const DiagBox &b, const R *V, const double *Vs, double *part, double *out,
                        hipStream_t s);

}  // namespace gpl
