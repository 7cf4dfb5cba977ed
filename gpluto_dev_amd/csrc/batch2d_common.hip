// Per-box time-step control of the batched 2D update.
#include "batch2d.hpp"

namespace gpl {

// invDt_hyp = max(C_dt)/DIMENSIONS over the box's y items
// (update_stage.cpp:422-425); optionally NextTimeStep
// (next_time_step.cpp:80,123-125) and time += dt (main.cpp:458).
__global__ __launch_bounds__(256)
void k2_dt(const double *__restrict__ part, const int *__restrict__ beg, StepCtl *ctl, double cfl,
           double cfl_max_var, int advance)
{
  __shared__ double red[4];
  const int bx = blockIdx.x;
  double m = -1.0;
  for (int q = beg[bx] + threadIdx.x; q < beg[bx + 1]; q += blockDim.x) m = m >= part[q] ? m : part[q];
  for (int o = 32; o > 0; o >>= 1) {
    const double w = __shfl_xor(m, o, 64);
    m = m >= w ? m : w;
  }
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < 4; q++) m = m >= red[q] ? m : red[q];
    m = red[0] >= m ? red[0] : m;
    const double inv = m/2.0;
    StepCtl &c = ctl[bx];
    c.inv_dt = inv;
    if (advance) {
      const double dt = c.dt;
      c.time += dt;
      const double dt_new = cfl/inv;
      c.dt = dt_new <= cfl_max_var*dt ? dt_new : cfl_max_var*dt;
      c.nstep += 1;
    }
  }
}

void batch2d_launch_dt(const double *cdt_part, const int *part_beg, int nbox, StepCtl *ctl,
                       double cfl, double cfl_max_var, int advance, hipStream_t s)
{
  hipLaunchKernelGGL(k2_dt, dim3(nbox), dim3(256), 0, s, cdt_part, part_beg, ctl, cfl, cfl_max_var,
                     advance);
}

}  // namespace gpl
