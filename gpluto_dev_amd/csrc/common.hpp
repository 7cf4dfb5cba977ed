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

// Zone-by-zone error reports: the reference's Data::err (Src/structs.hpp:678)
// as ConsToPrim3D writes it in every zone it converts (Src/mappers3D.cpp:115,
// 121, 173-176) and the flagged-zone counts Where() adds to Data::err_count
// (Src/debug_tools.cpp:587-593).  err is this stage's array over the device
// cells, count this stage's counter, total the running count; err == nullptr
// when the reports are off.
struct ZoneRep {
  uint32_t *err;
  unsigned long long *count, *total;
};

__device__ __forceinline__ void zone_report(const ZoneRep &z, size_t c, uint32_t e)
{
  if (!z.err) return;
  z.err[c] = e;
  if (e) { atomicAdd(z.count, 1ull); atomicAdd(z.total, 1ull); }
}

// counts[0..2] = 0 at the start of a step (Src/main.cpp:423); a kernel, so a
// captured step replays it
void launch_zone_reset(unsigned long long *counts, hipStream_t s);

// invDt_hyp = max over the partial maxima / DIMENSIONS; with advance != 0
// also NextTimeStep and time += dt on the device.
void launch_dt(const double *part, int nparts, double dims, StepCtl *ctl, double cfl,
               double cfl_max_var, int advance, hipStream_t s);

}  // namespace gpl
