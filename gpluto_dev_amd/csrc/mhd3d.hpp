// Shared declarations of the 3D MHD stage kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.hpp"

namespace gpl {

struct NetPack;

struct Box3 {
  int nx, ny, nz;
  int ncell;
  double dx, dy, dz;
  // active cells lo..hi per direction: the whole box for a periodic box; a
  // box padded with ghost layers (non-periodic sides) reduces C_dt and
  // reports error bits from the reference's cells and faces only
  int lo[3], hi[3];
  // density and pressure floors of ConsToPrim (run_config.small_rho /
  // small_prs), read by the RMHD HLLD's con2prim of the HLL state
  double small_rho, small_prs;
  // planes kb .. ke-1 of the launch (the whole box: 0, nz) for the fused x
  // sweep, the y and z marches, the EMF tiles and the finish kernel, and the
  // first C_dt partial slot of the z march's blocks; a slab of planes starts
  // at a multiple of the march lengths
  int kb, ke, part0;
};

template <typename R>
struct Mhd3dBufs {
  R *V;            // [8][n]
  // Conserved state: the stage reads Uin (and from its second stage on U0,
  // the state at the start of the step) and writes U, cell by cell; U may be
  // Uin or U0, never another cell's input.
  const double *Uin; // [5][n]
  double *U;       // [5][n]
  const double *U0;  // [5][n]
  // Staggered faces: three distinct buffers, so that no stage overwrites a
  // face another thread still reads.
  const double *Vin; // [3][n] faces at stage start
  double *Vout;    // [3][n] faces after the stage
  const double *Bs0; // [3][n] faces at the start of the step (second stage on)
  R *F;            // [3][5][n]
  R *CM;           // [3][n]
  R *FD;           // [3][4][n] UCT face records ([3][6][n] with UCT_HLLD)
  R *E;            // [3][n]
  R *dU;           // [5][n] flux-divergence accumulator within a stage (fused path)
  R *Cdt;          // [n]    C_dt accumulator within a stage (fused path)
  double *cdt_part;
  uint32_t *err;
  int emf_march;   // EMF kernel: 0 one edge per thread, n > 0 register z-march of n, -n LDS tiles marching n
  int phys;        // 0: ideal MHD, 1: relativistic MHD (FP64, fused kernels)
  const NetPack *net; // RMHD: learned con2prim guess (nullptr: the reference's A27 guess)
  ZoneRep zr;       // this stage's zone-by-zone reports (zr.err null: off)
};

template <typename R>
void mhd3d_launch(int which, const Mhd3dBufs<R> &b, const Box3 &g, const StepCtl *ctl, RKW w,
                  int first_stage, int want_dt, int limiter, int solver, int march, double gamma,
                  double small_rho, double small_prs, hipStream_t s);

// Ghost-zone fill of a padded box (non-periodic sides), side by side;
// type[s] = GPL_PERIODIC, GPL_OUTFLOW or GPL_REFLECTIVE for X1-beg, ..., X3-end.
// dbl dump image of the box (active cells, faces 0..n along each direction;
// the upper face of a periodic direction is the image of face 0) into dst
template <typename R>
void mhd3d_dump_gather(const Box3 &g, const int na[3], const int periodic[3], const R *V,
                       const double *Vs, double *dst, hipStream_t s);

template <typename R>
void mhd3d_boundary(const Box3 &g, const int type[6], R *V, double *Vs, hipStream_t s);

int mhd3d_blocks(const Box3 &g);
int mhd3d_march_blocks(const Box3 &g, int march);

}  // namespace gpl
