// Diagnostics reductions (DataMinMax, Src/main.cpp:752-860; CT_CheckDivB,
// Src/MHD/CT/ct_update.cpp:580-635).  Min, max and maxima of non-negative
// values do not depend on the order of the reduction, so the results are the
// reference's to the bit for the same state.
#include "diagnostics.hpp"

#include <cmath>

namespace gpl {

namespace {

// the reference's MAX / MIN (Src/macros.hpp:191-194)
__device__ __forceinline__ double dmax(double a, double b) { return a >= b ? a : b; }
__device__ __forceinline__ double dmin(double a, double b) { return a <= b ? a : b; }

// reduce q[0..NQ) over the block, the first NLO entries by min and the rest
// by max; the block's results land in res[0..NQ)
template <int NQ, int NLO>
__device__ void block_reduce(const double *q, double *sh, double *res)
{
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int a = 0; a < NQ; a++) {
    double v = q[a];
    for (int o = 32; o > 0; o >>= 1) {
      const double u = __shfl_xor(v, o, 64);
      v = a < NLO ? dmin(v, u) : dmax(v, u);
    }
    if (lane == 0) sh[a*8 + wv] = v;
  }
  __syncthreads();
  if (threadIdx.x < NQ) {
    const int a = threadIdx.x;
    double v = sh[a*8];
    for (int w = 1; w < nw; w++) v = a < NLO ? dmin(v, sh[a*8 + w]) : dmax(v, sh[a*8 + w]);
    res[a] = v;
  }
}

template <typename R, int NV>
__global__ __launch_bounds__(256)
void k_diag(DiagBox b, const R *__restrict__ V, const double *__restrict__ Vs,
            double *__restrict__ part)
{
  constexpr int NQ = 2*NV + 2;
  __shared__ double sh[NQ*8];
  const size_t n = (size_t)b.ncell;
  double q[NQ];
#pragma unroll
  for (int a = 0; a < NV; a++) { q[a] = 1e18; q[NV + a] = -1e18; }   // Src/main.cpp:820-823
  q[2*NV] = 0.0;
  q[2*NV + 1] = 0.0;
  for (int c = blockIdx.x*blockDim.x + threadIdx.x; c < b.ncell; c += gridDim.x*blockDim.x) {
    {
      const int i = c % b.nx, t = c / b.nx, j = t % b.ny, k = t / b.ny;
      if (i < b.lo[0] || i > b.hi[0] || j < b.lo[1] || j > b.hi[1] || k < b.lo[2] || k > b.hi[2])
        continue;
    }
    double v[NV];
#pragma unroll
    for (int a = 0; a < NV; a++) {
      v[a] = (double)V[a*n + c];
      q[a] = dmin(q[a], v[a]);
      q[NV + a] = dmax(q[NV + a], v[a]);
    }
    // Max(MACH), Src/main.cpp:786-801
    double cs = v[NV - 1]/v[0];
    cs = sqrt(cs);
    const double mi = fabs(v[1])/cs, mj = fabs(v[2])/cs, mk = fabs(v[3])/cs;
    double lm = dmax(mi, mj);
    lm = dmax(lm, mk);
    q[2*NV] = dmax(q[2*NV], lm);
    if (Vs) {
      // div B with the face areas (Src/MHD/CT/ct_update.cpp:600-617)
      const int i = c % b.nx, t = c / b.nx, j = t % b.ny, k = t / b.ny;
      const int ip = i + 1 == b.nx ? 0 : i + 1, jp = j + 1 == b.ny ? 0 : j + 1;
      const int kp = k + 1 == b.nz ? 0 : k + 1;
      const double *Bx = Vs, *By = Vs + n, *Bz = Vs + 2*n;
      const double dbx = b.ax*Bx[ip + b.nx*(j + b.ny*k)] - b.ax*Bx[c];
      const double dby = b.ay*By[i + b.nx*(jp + b.ny*k)] - b.ay*By[c];
      // DIM_EXPAND: no z term in 2D, whose faces hold two components
      const double dbz = b.nb == 3 ? b.az*Bz[i + b.nx*(j + b.ny*kp)] - b.az*Bz[c] : 0.0;
      const double divb = b.nb == 3 ? dbx + dby + dbz : dbx + dby;
      q[2*NV + 1] = dmax(q[2*NV + 1], fabs(divb));
    }
  }
  __shared__ double res[NQ];
  block_reduce<NQ, NV>(q, sh, res);
  if (threadIdx.x < NQ) part[(size_t)blockIdx.x*DIAG_MAXQ + threadIdx.x] = res[threadIdx.x];
}

template <int NV>
__global__ __launch_bounds__(256)
void k_diag_final(int nparts, const double *__restrict__ part, double *__restrict__ out)
{
  constexpr int NQ = 2*NV + 2;
  __shared__ double sh[NQ*8];
  double q[NQ];
#pragma unroll
  for (int a = 0; a < NV; a++) { q[a] = 1e18; q[NV + a] = -1e18; }
  q[2*NV] = 0.0;
  q[2*NV + 1] = 0.0;
  for (int p = threadIdx.x; p < nparts; p += blockDim.x) {
#pragma unroll
    for (int a = 0; a < NQ; a++) {
      const double v = part[(size_t)p*DIAG_MAXQ + a];
      q[a] = a < NV ? dmin(q[a], v) : dmax(q[a], v);
    }
  }
  __shared__ double res[NQ];
  block_reduce<NQ, NV>(q, sh, res);
  if (threadIdx.x < NQ) out[threadIdx.x] = res[threadIdx.x];
}

}  // namespace

template <typename R>
void launch_diagnostics(const DiagBox &b, const R *V, const double *Vs, double *part, double *out,
                        hipStream_t s)
{
  int nb = (b.ncell + 255)/256;
  if (nb > DIAG_BLOCKS) nb = DIAG_BLOCKS;
  if (b.nv == 8) {
    hipLaunchKernelGGL((k_diag<R, 8>), dim3(nb), dim3(256), 0, s, b, V, Vs, part);
    hipLaunchKernelGGL(k_diag_final<8>, dim3(1), dim3(256), 0, s, nb, (const double *)part, out);
  } else {
    hipLaunchKernelGGL((k_diag<R, 5>), dim3(nb), dim3(256), 0, s, b, V, Vs, part);
    hipLaunchKernelGGL(k_diag_final<5>, dim3(1), dim3(256), 0, s, nb, (const double *)part, out);
  }
}

template void launch_diagnostics<double>(const DiagBox &, const double *, const double *, double *,
                                         double *, hipStream_t);
template void launch_diagnostics<float>(const DiagBox &, const float *, const double *, double *,
                                        double *, hipStream_t);

}  // namespace gpl
