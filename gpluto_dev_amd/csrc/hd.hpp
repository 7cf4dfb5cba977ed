// Declarations of the hydrodynamics (HD) stage kernels: 1D, 2D or 3D boxes,
// periodic, outflow or reflective sides, PLM + HLLC/HLL + RK2/RK3, FP64.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/gpluto_amd.h"
#include "common.hpp"

namespace gpl {

struct HdBox {
  int nx, ny, nz;     // active zones (1 along unused directions)
  int ncell;
  int dims;           // DIMENSIONS
  int lo[3], hi[3];   // side types per direction: GPL_PERIODIC / GPL_OUTFLOW / GPL_REFLECTIVE
  double dx, dy, dz;
};

struct HdBufs {
  const double *Vin;  // [5][n] primitives at stage start
  double *Vout;       // [5][n] primitives after the stage
  double *U;          // [5][n] conserved
  double *U0;         // [5][n] stage-start copy (RK)
  double *dU;         // [5][n] flux-divergence accumulator (dims > 1)
  double *Cdt;        // [n]    C_dt accumulator (dims > 1)
  double *cdt_part;   // per-block maxima of C_dt (hd_blocks entries)
  uint32_t *err;
  ZoneRep zr;         // this stage's zone-by-zone reports (zr.err null: off)
};

// Launch the sweep along direction dir (0..dims-1).  The last direction also
// applies the RK combination and ConsToPrim and writes Vout.
void hd_launch(int dir, const HdBufs &b, const HdBox &g, const StepCtl *ctl, RKW w,
               int first_stage, int want_dt, int limiter, int solver, double gamma,
               double small_rho, double small_prs, hipStream_t s);

int hd_block_size(const HdBox &g);
int hd_blocks(const HdBox &g);

}  // namespace gpl
