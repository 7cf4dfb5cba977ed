// Shared declarations of the 3D MHD stage kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gpl {

struct Box3 {
  int nx, ny, nz;
  int ncell;
  double dx, dy, dz;
};

// Device-resident time-step control: the step loop never needs the host.
struct StepCtl {
  double dt;       // current time step (timeStep::dt)
  double time;     // timeStep::time
  double inv_dt;   // timeStep::invDt_hyp of the last stage-1 update
  long long nstep;
};

// RK weights of one stage: U = w0*U0 + w1*U + wR*R  (rk_step.cpp:20-51)
struct RKW { double w0, w1, wR; };

template <typename R>
struct Mhd3dBufs {
  R *V;            // [8][n]
  double *U;       // [5][n]
  double *U0;      // [5][n]
  const double *Vin; // [3][n] faces at stage start
  double *Vout;    // [3][n] faces after the stage
  double *Bs0;     // [3][n]
  R *F;            // [3][5][n]
  R *CM;           // [3][n]
  R *FD;           // [3][5][n]
  R *E;            // [3][n]
  double *dU;      // [5][n] flux-divergence accumulator (fused path)
  double *Cdt;     // [n]    C_dt accumulator (fused path)
  double *cdt_part;
  uint32_t *err;
};

enum KernelId { K_SWEEP_X = 0, K_SWEEP_Y, K_SWEEP_Z, K_RHS, K_SWEEPX_F, K_SWEEPY_F, K_SWEEPZ_F,
                K_EMF, K_FINISH, K_DT, K_COUNT };

template <typename R>
void mhd3d_launch(int which, const Mhd3dBufs<R> &b, const Box3 &g, const StepCtl *ctl, RKW w,
                  int first_stage, int want_dt, int limiter, int solver, int march, double gamma,
                  double small_rho, double small_prs, hipStream_t s);

int mhd3d_blocks(const Box3 &g);
int mhd3d_march_blocks(const Box3 &g, int march);

void launch_dt(const double *part, int nparts, double dims, StepCtl *ctl, double cfl,
               double cfl_max_var, int advance, hipStream_t s);

}  // namespace gpl
