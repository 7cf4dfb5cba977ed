// HD stage update (PLM + HLLC/HLL + RK2/RK3) on 1D/2D/3D boxes with periodic
// or outflow sides, FP64, for gfx950.
//
// Device data model: V[5][n] (rho, v1..3, prs), U/U0[5][n] (rho, m1..3, E),
// no ghost zones.  A neighbour outside the box is the wrapped cell
// (periodic), the nearest active cell (outflow: OutflowBoundary,
// boundary.cpp:677-760, copies the edge cell into every ghost zone) or the
// mirrored cell with its normal velocity negated (reflective:
// ReflectiveBoundary, boundary.cpp:931-1000, ghost i = vsign*Vc(i0 - i)),
// which is bitwise what the reference's ghost fill supplies to the sweep.
//
// One kernel per direction and stage, one thread per cell: the thread
// reconstructs cells -1, 0, +1 along the sweep, solves the Riemann problem on
// both faces of its cell (a face shared by two cells is solved identically by
// both) and forms rhs = -dt/dx (F(i+1) - F(i)) (rhs.cpp:386,671,900).  The
// first direction stores dUc = rhs, later ones dUc += rhs
// (update_stage.cpp:173-184,272-284,369-380); the last direction adds its
// term in registers, applies the RK combination (rk_step.cpp:366-374,509-518)
// and ConsToPrim3D (mappers3D.cpp:104-194), and writes the primitives of the
// next stage to the other V buffer.
#include "hd.hpp"
#include "physics.hpp"

namespace gpl {

template <typename T>
__device__ __forceinline__ T hd_block_max(T v)
{
  __shared__ T red[16];
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

template <int D, bool LAST, int LIM, int SOLVER>
__global__ __launch_bounds__(256)
void k_hd_sweep(HdBox g, const double *__restrict__ V, double *__restrict__ Vout,
                double *__restrict__ U, double *__restrict__ U0, double *__restrict__ dU,
                double *__restrict__ Cdt, const StepCtl *__restrict__ ctl, RKW w,
                int first_stage, int want_cdt, double *__restrict__ cdt_part, double gamma,
                double small_rho, double small_prs, uint32_t *__restrict__ err, ZoneRep zr)
{
  const int c = blockIdx.x*blockDim.x + threadIdx.x;
  double cdt_max = -1.0;
  uint32_t e = 0;
  if (c < g.ncell) {
    const size_t n = (size_t)g.ncell;
    const int nd = D == 0 ? g.nx : (D == 1 ? g.ny : g.nz);
    const int sd = D == 0 ? 1 : (D == 1 ? g.nx : g.nx*g.ny);
    const int p = D == 0 ? c % g.nx : (D == 1 ? (c/g.nx) % g.ny : c/(g.nx*g.ny));
    const int base = c - p*sd;
    int idx[5];
    bool flip[5];                   // a mirrored ghost: its normal velocity changes sign
#pragma unroll
    for (int o = 0; o < 5; o++) {
      int q = p + o - 2;
      flip[o] = false;
      if (q < 0) {
        const int t = g.lo[D];
        flip[o] = t == GPL_REFLECTIVE;
        q = t == GPL_PERIODIC ? q + nd : (t == GPL_OUTFLOW ? 0 : -1 - q);
      } else if (q >= nd) {
        const int t = g.hi[D];
        flip[o] = t == GPL_REFLECTIVE;
        q = t == GPL_PERIODIC ? q - nd : (t == GPL_OUTFLOW ? nd - 1 : 2*nd - 1 - q);
      }
      idx[o] = base + q*sd;
    }
    double v[5][5];                 // [cell -2..+2][variable]
#pragma unroll
    for (int o = 0; o < 5; o++) {
#pragma unroll
      for (int nv = 0; nv < 5; nv++) v[o][nv] = V[nv*n + idx[o]];
      if (flip[o]) v[o][1 + D] = -v[o][1 + D];     // vsign = -1 (FlipSign, boundary.cpp:840-925)
    }
    // Reconstruct (plm_states.cpp:121-138): vL(i+1) = v + dv/2, vR(i) = v - dv/2
    double vlo_l[5], vlo_r[5], vhi_l[5], vhi_r[5];
#pragma unroll
    for (int nv = 0; nv < 5; nv++) {
      const double dm = plm_slope<double, LIM>(v[0][nv], v[1][nv], v[2][nv]);
      const double d0 = plm_slope<double, LIM>(v[1][nv], v[2][nv], v[3][nv]);
      const double dp = plm_slope<double, LIM>(v[2][nv], v[3][nv], v[4][nv]);
      vlo_l[nv] = v[1][nv] + 0.5*dm;
      vlo_r[nv] = v[2][nv] - 0.5*d0;
      vhi_l[nv] = v[2][nv] + 0.5*d0;
      vhi_r[nv] = v[3][nv] - 0.5*dp;
    }
    FaceFlux<double> flo, fhi;
    if (SOLVER == 2) {
      hd_hllc<double, D>(vlo_l, vlo_r, gamma, flo);
      hd_hllc<double, D>(vhi_l, vhi_r, gamma, fhi);
    } else {
      hd_hll<double, D>(vlo_l, vlo_r, gamma, flo);
      hd_hll<double, D>(vhi_l, vhi_r, gamma, fhi);
    }
    // each face's NaN is counted by both of its cells; the bits are the same
    e |= flo.err | fhi.err;
    const double ddx = D == 0 ? g.dx : (D == 1 ? g.dy : g.dz);
    const double dtdx = ctl->dt/ddx;
    double du[5];
#pragma unroll
    for (int nv = 0; nv < 5; nv++) {
      double rhs = -dtdx*(fhi.f[nv] - flo.f[nv]);
      rhs += 0.0;
      if (D == 0) du[nv] = rhs;
      else { du[nv] = dU[nv*n + c]; du[nv] += rhs; }
    }
    const double cmlo = rmax(rabs(flo.Sl), rabs(flo.Sr));
    const double cmhi = rmax(rabs(fhi.Sl), rabs(fhi.Sr));
    double cd = 0.0;
    if (want_cdt) {
      const double t = 0.5*(cmlo + cmhi)/ddx;        // rhs.cpp:523,756,981
      if (D == 0) cd = t;
      else { cd = Cdt[c]; cd += t; }
    }
    if (!LAST) {
#pragma unroll
      for (int nv = 0; nv < 5; nv++) dU[nv*n + c] = du[nv];
      if (want_cdt) Cdt[c] = cd;
    } else {
      double u[5], vn[5];
#pragma unroll
      for (int nv = 0; nv < 5; nv++) {
        const size_t o = nv*n + c;
        if (first_stage) {
          const double u0 = U[o];
          U0[o] = u0;
          u[nv] = u0 + w.wR*du[nv];
        } else {
          u[nv] = w.w0*U0[o] + w.w1*U[o] + w.wR*du[nv];
        }
      }
      const uint32_t ce = hd_cons_to_prim(u, gamma - 1.0, small_rho, small_prs, vn);
      e |= ce;
      zone_report(zr, c, ce);
#pragma unroll
      for (int nv = 0; nv < 5; nv++) { U[nv*n + c] = u[nv]; Vout[nv*n + c] = vn[nv]; }
      if (want_cdt) cdt_max = cd;
    }
  }
  if (e) atomicOr(err, e);
  if (LAST && want_cdt) {
    const double m = hd_block_max(cdt_max);
    if (threadIdx.x == 0) cdt_part[blockIdx.x] = m;
  }
}

int hd_block_size(const HdBox &g) { return g.ncell <= (1 << 16) ? 64 : 256; }
int hd_blocks(const HdBox &g) { const int b = hd_block_size(g); return (g.ncell + b - 1)/b; }

template <int D, bool LAST, int LIM, int SOLVER>
static void launch_t(const HdBufs &b, const HdBox &g, const StepCtl *ctl, RKW w, int first_stage,
                     int want_dt, double gamma, double small_rho, double small_prs, hipStream_t s)
{
  hipLaunchKernelGGL((k_hd_sweep<D, LAST, LIM, SOLVER>), dim3(hd_blocks(g)), dim3(hd_block_size(g)),
                     0, s, g, b.Vin, b.Vout, b.U, b.U0, b.dU, b.Cdt, ctl, w, first_stage, want_dt,
                     b.cdt_part, gamma, small_rho, small_prs, b.err, b.zr);
}

template <int D, bool LAST>
static void launch_d(const HdBufs &b, const HdBox &g, const StepCtl *ctl, RKW w, int first_stage,
                     int want_dt, int limiter, int solver, double gamma, double small_rho,
                     double small_prs, hipStream_t s)
{
  if (limiter == MC_LIM) {
    if (solver == 2) launch_t<D, LAST, MC_LIM, 2>(b, g, ctl, w, first_stage, want_dt, gamma, small_rho, small_prs, s);
    else             launch_t<D, LAST, MC_LIM, 1>(b, g, ctl, w, first_stage, want_dt, gamma, small_rho, small_prs, s);
  } else {
    if (solver == 2) launch_t<D, LAST, VANLEER_LIM, 2>(b, g, ctl, w, first_stage, want_dt, gamma, small_rho, small_prs, s);
    else             launch_t<D, LAST, VANLEER_LIM, 1>(b, g, ctl, w, first_stage, want_dt, gamma, small_rho, small_prs, s);
  }
}

void hd_launch(int dir, const HdBufs &b, const HdBox &g, const StepCtl *ctl, RKW w,
               int first_stage, int want_dt, int limiter, int solver, double gamma,
               double small_rho, double small_prs, hipStream_t s)
{
  const bool last = dir == g.dims - 1;
  if (dir == 0) {
    if (last) launch_d<0, true>(b, g, ctl, w, first_stage, want_dt, limiter, solver, gamma, small_rho, small_prs, s);
    else      launch_d<0, false>(b, g, ctl, w, first_stage, want_dt, limiter, solver, gamma, small_rho, small_prs, s);
  } else if (dir == 1) {
    if (last) launch_d<1, true>(b, g, ctl, w, first_stage, want_dt, limiter, solver, gamma, small_rho, small_prs, s);
    else      launch_d<1, false>(b, g, ctl, w, first_stage, want_dt, limiter, solver, gamma, small_rho, small_prs, s);
  } else {
    launch_d<2, true>(b, g, ctl, w, first_stage, want_dt, limiter, solver, gamma, small_rho, small_prs, s);
  }
}

}  // namespace gpl
