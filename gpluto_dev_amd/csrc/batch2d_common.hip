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

// The dbl dump image of one box in the file's own order
// (Src/write_data.cpp:136-159 with FillOutputBuffer :564-607): the eight
// primitives over the active zones, i fastest, then Bx1s with nx+1 faces
// along x and Bx2s with ny+1 faces along y.  A periodic direction's face n is
// the image of face 0; an outflow or reflective side's is its own.
template <typename R>
__global__ __launch_bounds__(256)
void k2_dump_gather(BoxDesc d, int px, int py, const R *__restrict__ V, const double *__restrict__ Vs,
                    size_t N, double *__restrict__ dst)
{
  const int nx = d.hi[0] - d.lo[0] + 1, ny = d.hi[1] - d.lo[1] + 1;
  const long long ex = nx + 1, ey = ny + 1;
  const long long t = (long long)blockIdx.x*blockDim.x + threadIdx.x;
  if (t >= ex*ey) return;
  const int i = (int)(t % ex), j = (int)(t / ex);
  const size_t na = (size_t)nx*ny;
  auto dev = [&](int a, int bb) -> size_t {
    if (a == nx && px) a = 0;
    if (bb == ny && py) bb = 0;
    return (size_t)d.off + (size_t)(a + d.lo[0]) + (size_t)d.nx*(bb + d.lo[1]);
  };
  if (i < nx && j < ny) {
    const size_t c = dev(i, j), o = (size_t)j*nx + i;
#pragma unroll
    for (int q = 0; q < 8; q++) dst[q*na + o] = (double)V[q*N + c];
  }
  size_t base = 8*na;
  if (j < ny) dst[base + (size_t)j*(nx + 1) + i] = Vs[dev(i, j)];
  base += (size_t)(nx + 1)*ny;
  if (i < nx) dst[base + (size_t)j*nx + i] = Vs[N + dev(i, j)];
}

template <typename R>
void batch2d_dump_gather(const BoxDesc &d, const int per[2], const R *V, const double *Vs, size_t ncell,
                         double *dst, hipStream_t s)
{
  const long long pts = (long long)(d.hi[0] - d.lo[0] + 2)*(d.hi[1] - d.lo[1] + 2);
  hipLaunchKernelGGL(k2_dump_gather<R>, dim3((unsigned)((pts + 255)/256)), dim3(256), 0, s,
                     d, per[0], per[1], V, Vs, ncell, dst);
}

template void batch2d_dump_gather<double>(const BoxDesc &, const int *, const double *, const double *,
                                          size_t, double *, hipStream_t);
template void batch2d_dump_gather<float>(const BoxDesc &, const int *, const float *, const double *,
                                         size_t, double *, hipStream_t);

}  // namespace gpl
