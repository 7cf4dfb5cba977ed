// Declarations of the batched 2D MHD stage kernels: many independent boxes
// of uneven sizes advanced together, one launch per phase and group of boxes,
// each launch a persistent grid striding over the group's slice of a work
// list.  Boxes are periodic,
// or, with an outflow or reflective side, stored as the reference's own array
// (two ghost layers per side and one spare layer for the outermost face) and
// computed on as a periodic box of that size, with the ghosts refilled at
// every stage and C_dt and error bits taken from the reference's cells and
// faces only.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.hpp"

namespace gpl {

// One stored nx x ny box: cells off .. off + nx*ny - 1 of every batch
// array, c = off + i + nx*j.  Active cells lo[d] .. hi[d]: the whole box when
// periodic, else nx = active + 5 with lo = 2.
struct BoxDesc {
  int nx, ny;
  int off;
  double dx, dy;
  int lo[2], hi[2];
};

// Work items.  Sweep items are one wavefront each:
//   x: {box, row j, first cell i0, cells in the segment (<= 63)}
//   y: {box, first column i0 (64 lanes), segment index, unused}
// Cell tiles (EMF, face update): {box, first cell within the box}, 256 cells.
struct Item4 { int box, a, b, c; };
struct Item2 { int box, start; };

template <typename R>
struct Batch2dBufs {
  R *V;              // [8][ncell] primitives (B1, B2 = face averages, B3 cell-centred)
  // conserved rho, m1, m2, m3, B3, E [6][ncell]: the stage reads Uin (and
  // from its second stage on U0, the state at the start of the step) and
  // writes U, cell by cell; U may be Uin or U0
  const double *Uin;
  double *U;
  const double *U0;
  const double *Vin; // [2][ncell] lower x- and y-faces at stage start
  double *Vout;      // [2][ncell] faces after the stage (a third buffer)
  const double *Bs0; // [2][ncell] faces at the start of the step (second stage on)
  R *FD;             // [2][4][ncell] UCT face data et, eb, min(0,Sl), max(0,Sr); UCT_HLLD:
                     // [2][6][ncell] et, eb, aL, aR, dL, dR
  R *E;              // [ncell] Ez at the lower-left corner of every cell
  R *dU;             // [6][ncell] x-sweep flux divergence
  R *Cdt;            // [ncell]    x-sweep C_dt
  double *cdt_part;  // one C_dt maximum per y item
  uint32_t *err;
  StepCtl *ctl;      // [nbox]
  const BoxDesc *box;
  const Item4 *xitems;
  const Item4 *yitems;
  const Item2 *tiles;
  int nx_items, ny_items, ntiles;
  int ncell;
  ZoneRep zr;        // this stage's zone-by-zone reports (zr.err null: off)
};

enum Batch2dPhase { B2_SWEEP_X = 0, B2_SWEEP_Y, B2_EMF, B2_FINISH };

// Launch one phase as a persistent grid of 'blocks' workgroups of 256
// threads; 'march' is the y-segment length the y items were built for.
// solver: 1 HLL, 2 HLLC, 3 HLLD (UCT_HLL records), 4 HLLD with UCT_HLLD.
template <typename R>
void batch2d_launch(int phase, const Batch2dBufs<R> &b, RKW w, int first_stage, int want_dt,
                    int limiter, int solver, int march, int blocks, double gamma,
                    double small_rho, double small_prs, hipStream_t s);

// Boundary (Src/boundary.cpp:67-303) of every box, sides X1_BEG .. X2_END
// in turn; type[4] per side (GPL_PERIODIC / GPL_OUTFLOW / GPL_REFLECTIVE),
// maxp the largest stored box side.  Faces are FP64, cells R.
template <typename R>
void batch2d_boundary(const BoxDesc *box, int nbox, int maxp, const int type[4], R *V, double *Vs,
                      size_t ncell, hipStream_t s);

// The dbl dump image of one box into dst (device): the eight primitives over
// the active zones, then Bx1s (nx+1 faces along x) and Bx2s (ny+1 along y);
// per[d]: periodic direction (face n = the image of face 0).
template <typename R>
void batch2d_dump_gather(const BoxDesc &d, const int per[2], const R *V, const double *Vs, size_t ncell,
                         double *dst, hipStream_t s);

// Per-box dt: invDt_hyp = max over the box's y items / 2, then optionally
// NextTimeStep and time += dt (one workgroup per box).
void batch2d_launch_dt(const double *cdt_part, const int *part_beg, int nbox, StepCtl *ctl,
                       double cfl, double cfl_max_var, int advance, hipStream_t s);

}  // namespace gpl
