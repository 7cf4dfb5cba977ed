// Declarations shared by every stage-update engine: the device-resident
// time-step control block, RK stage weights, kernel ids for timing, and the
// dt reduction kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gpl {

// Device-resident time-step control: the step loop never needs the host.
struct StepCtl {
  double dt;       // current time step (timeStep::dt)
  double time;     // timeStep::time
  double inv_dt;   // timeStep::invDt_hyp of the last stage-1 update
  long long nstep;
};

// RK weights of one stage: U = w0*U0 + w1*U + wR*R  (rk_step.cpp:20-51)
struct RKW { double w0, w1, wR; };

// Every kernel the library launches, for the per-kernel timing tables.
enum KernelId { K_SWEEP_X = 0, K_SWEEP_Y, K_SWEEP_Z, K_RHS, K_SWEEPX_F, K_SWEEPY_F, K_SWEEPZ_F,
                K_EMF, K_FINISH, K_DT, K_HD_X, K_HD_Y, K_HD_Z, K_SWEEPXY_F, K_COUNT };

// invDt_hyp = max over the partial maxima / DIMENSIONS; with advance != 0
// also NextTimeStep and time += dt on the device.
void launch_dt(const double *part, int nparts, double dims, StepCtl *ctl, double cfl,
               double cfl_max_var, int advance, hipStream_t s);

}  // namespace gpl
