// Time-step control kernel and launch-geometry helpers shared by both
// precisions of the 3D MHD stage update.
#include "mhd3d.hpp"

namespace gpl {

template <typename T>
__device__ __forceinline__ T block_max_dt(T v)
{
  __shared__ T red[32];
  for (int o = 32; o > 0; o >>= 1) {
    T w = __shfl_xor(v, o, 64);
    v = v >= w ? v : w;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    T m = red[0];
    for (int q = 1; q < (int)((blockDim.x + 63) >> 6); q++) m = m >= red[q] ? m : red[q];
    red[0] = m;
  }
  __syncthreads();
  return red[0];
}

// invDt_hyp = max(C_dt)/DIMENSIONS (update_stage.cpp:422-425); optionally the
// NextTimeStep recurrence (next_time_step.cpp:80,123-125) and time += dt
// (main.cpp:458).
__global__ __launch_bounds__(1024)
void k_dt(const double *__restrict__ part, int nparts, double dims, StepCtl *ctl,
          double cfl, double cfl_max_var, int advance)
{
  double m = -1.0;
  for (int q = threadIdx.x; q < nparts; q += blockDim.x) m = m >= part[q] ? m : part[q];
  m = block_max_dt(m);
  if (threadIdx.x == 0) {
    const double inv = m/dims;
    ctl->inv_dt = inv;
    if (advance) {
      const double dt = ctl->dt;
      ctl->time += dt;
      const double dt_new = cfl/inv;
      ctl->dt = dt_new <= cfl_max_var*dt ? dt_new : cfl_max_var*dt;
      ctl->nstep += 1;
    }
  }
}

__global__ void k_zone_reset(unsigned long long *counts)
{
  if (threadIdx.x < 3) counts[threadIdx.x] = 0ull;
}

void launch_zone_reset(unsigned long long *counts, hipStream_t s)
{
  hipLaunchKernelGGL(k_zone_reset, dim3(1), dim3(64), 0, s, counts);
}

int mhd3d_blocks(const Box3 &g) { return (g.ncell + 255)/256; }

// blocks of the z march (one C_dt partial maximum each)
int mhd3d_march_blocks(const Box3 &g, int march)
{
  const long long nt = (long long)g.nx*g.ny*((g.nz + march - 1)/march);
  return (int)((nt + 255)/256);
}

void launch_dt(const double *part, int nparts, double dims, StepCtl *ctl, double cfl,
               double cfl_max_var, int advance, hipStream_t s)
{
  hipLaunchKernelGGL(k_dt, dim3(1), dim3(1024), 0, s, part, nparts, dims, ctl, cfl, cfl_max_var, advance);
}

}  // namespace gpl
