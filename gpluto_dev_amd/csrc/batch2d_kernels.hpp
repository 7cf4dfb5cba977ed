// Batched 2D ideal-MHD stage update: PLM + HLLD/HLLC/HLL + UCT_HLL or UCT_HLLD CT
// + RK2/RK3 on many independent boxes of uneven sizes, for gfx950.  A box with
// an outflow or reflective side is stored with the reference's ghost layers
// and computed on as a periodic box of that size (batch2d_boundary.hip
// refills the ghosts before every stage); C_dt and error bits come from the
// reference's cells and faces only.
//
// Every phase is one launch over a group of boxes: a persistent grid whose
// workgroups stride over the group's slice of a work list (sweep items are
// one wavefront each, cell tiles one workgroup each), so boxes from 32^2 to
// 512^2 share the chip in a single launch; the host runs the groups on
// streams of their own and the step is one HIP graph.
// The per-cell arithmetic is the 3D path's, restricted to DIMENSIONS 2 with
// three field components (reference 2D MHD, Src/MHD/mhd.hpp:21-41): B1, B2
// live on faces (CT), B3 is cell-centred and evolved by its flux.
//
//  x sweep: a wavefront owns 63 cells of a row and solves its 64 faces (the
//    last one is the upper face of cell 62); the flux of face i+1 reaches the
//    lane of cell i by a wavefront shuffle, so no flux touches memory.
//    dUc = rhs_x, C_dt = x term (update_stage.cpp:173-184, rhs.cpp:384-386,523).
//  y sweep: each lane marches L cells of a column with a register window and
//    completes dUc += rhs_y (rhs.cpp:671,756), then the RK combination of U
//    (rk_step.cpp:366-374,509-518) and a wavefront maximum of C_dt.
//  EMF: Ez = X1a + X2b of CT_EMF_Riemann2D (ct_emf_average.cpp:535-989).
//  finish: CT_Update (ct_update.cpp:181-258), the staggered RK combination
//    (rk_step.cpp:378-385,540-552), B1/B2 = face averages
//    (ct_field_average.cpp:182-184) and ConsToPrim (mappers3D.cpp:104-194).
// This header holds the templates; batch2d_fp64.hip and batch2d_mixed.hip
// instantiate them with their own compiler flags.
#pragma once
#include "batch2d.hpp"
#include "physics.hpp"

namespace gpl {

namespace b2 {

__device__ __forceinline__ int wrapm(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

template <typename T>
__device__ __forceinline__ T wave_max(T v)
{
  for (int o = 32; o > 0; o >>= 1) {
    T w = __shfl_xor(v, o, 64);
    v = v >= w ? v : w;
  }
  return v;
}

// PLM of a scalar column (array_reconstruct.cpp:167-205): left state at the
// interface between q1 and q2 from (q0,q1,q2), right state from (q1,q2,q3).
template <typename R, int LIM>
__device__ __forceinline__ void rec4(R q0, R q1, R q2, R q3, R &L, R &Rr)
{
  L  = q1 + R(0.5)*plm_slope<R, LIM>(q0, q1, q2);
  Rr = q2 - R(0.5)*plm_slope<R, LIM>(q1, q2, q3);
}

// Face records per direction: et, eb and the clipped speeds min(0, Sl),
// max(0, Sr) for UCT_HLL, from which the EMF kernel recomputes aL, aR and
// d = dL = dR (uct_hll_coef, the reference's expressions); et, eb, aL, aR,
// dL, dR for UCT_HLLD (SOLVER 4: HLLD with CT_Store_UCTHlld,
// Src/MHD/CT/ct_store.hpp:254-354).
template <int SOLVER> __host__ __device__ constexpr int nrec() { return SOLVER == 4 ? 6 : 4; }

template <typename R, int N, int SOLVER>
__device__ __forceinline__ void store_rec(const R *vl, const R *vr, R gamma, const FaceFlux<R> &ff,
                                          R *fd, int c, size_t n)
{
  if constexpr (SOLVER == 4) {
    R et, eb, aL, aR, dL, dR;
    uct_hlld_store<R, N>(vl, vr, gamma, ff.Sl, ff.Sr, et, eb, aL, aR, dL, dR);
    fd[0*n + c] = et; fd[1*n + c] = eb; fd[2*n + c] = aL; fd[3*n + c] = aR;
    fd[4*n + c] = dL; fd[5*n + c] = dR;
  } else {
    R et, eb, aL, aR, dd;
    uct_hll_store<R, N>(vl, vr, ff.Sl, ff.Sr, et, eb, aL, aR, dd);
    fd[0*n + c] = et; fd[1*n + c] = eb;
    fd[2*n + c] = rmin(R(0), ff.Sl); fd[3*n + c] = rmax(R(0), ff.Sr);
  }
}

template <typename R, int N, int SOLVER>
__device__ __forceinline__ void solve(const R *vl, const R *vr, R gamma, FaceFlux<R> &ff)
{
  if constexpr (SOLVER == 3 || SOLVER == 4) mhd_hlld<R, N>(vl, vr, gamma, ff);
  else if constexpr (SOLVER == 2) mhd_hllc<R, N>(vl, vr, gamma, ff);
  else mhd_hll<R, N>(vl, vr, gamma, ff);
}

// one UCT contribution (ct_emf_average.cpp:636-643), as in the 3D kernels;
// NQ = 4 coefficient records (aL, aR, dL, dR) for UCT_HLLD, 3 (dR = dL) for UCT_HLL
template <typename R, int LIMV, int NQ>
__device__ __forceinline__ R uct_val(const R *q, const R *e, const R *fa, const R *fb, int sgn)
{
  R bL, bR, vL, vR;
  rec4<R, MC_LIM>(q[0], q[1], q[2], q[3], bL, bR);
  rec4<R, LIMV>(e[0], e[1], e[2], e[3], vL, vR);
  const R aL = R(0.5)*(fa[0] + fb[0]);
  const R aR = R(0.5)*(fa[1] + fb[1]);
  const R dL = R(0.5)*(fa[2] + fb[2]);
  R dR = dL;
  if constexpr (NQ == 4) dR = R(0.5)*(fa[3] + fb[3]);
  const R phi = dR*bR - dL*bL;
  const R s = aL*vL*bL + aR*vR*bR;
  return sgn > 0 ? s + phi : s - phi;
}

// flux-vector index of conserved slot q of the 2D batch (rho, m1, m2, m3, B3, E)
__host__ __device__ constexpr int fidx(int q) { return q < 4 ? q : q + 2; }

// The reference's active cells (the whole box when periodic).
__device__ __forceinline__ bool active(const BoxDesc &bx, int i, int j)
{
  return i >= bx.lo[0] && i <= bx.hi[0] && j >= bx.lo[1] && j <= bx.hi[1];
}

// The reference solves the faces of the active cells and the upper boundary
// face along a sweep, over boxes widened by two ghost layers across it
// (update_stage.cpp:80-99); error bits count from those faces only.  (i, j):
// the face's lower-face cell index, i along x; for a periodic box the face
// index n along the sweep is the image of face 0.
template <int D>
__device__ __forceinline__ bool face_reported(const BoxDesc &bx, int i, int j)
{
  const int p[2] = {i, j};
#pragma unroll
  for (int d = 0; d < 2; d++) {
    const int lo = d == D ? bx.lo[d] : bx.lo[d] - 2, hi = d == D ? bx.hi[d] + 1 : bx.hi[d] + 2;
    if (p[d] < lo || p[d] > hi) return false;
  }
  return true;
}

}  // namespace b2

// ---------------------------------------------------------------------------
// x sweep
// ---------------------------------------------------------------------------
template <typename R, int LIM, int SOLVER>
__global__ __launch_bounds__(256)
void k2_sweepx(Batch2dBufs<R> b, R gamma, int want_cdt)
{
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const size_t n = (size_t)b.ncell;
  constexpr int VAR[7] = {0, 1, 2, 3, 5, 6, 7};   // B1 comes from the face
  const int ngroups = (b.nx_items + 3) >> 2;
  uint32_t e = 0;
  for (int gi = blockIdx.x; gi < ngroups; gi += gridDim.x) {
    const int it = 4*gi + wv;
    if (it >= b.nx_items) continue;               // uniform across the wavefront
    const Item4 w4 = b.xitems[it];
    const BoxDesc bx = b.box[w4.box];
    const int N = bx.nx, nc = w4.c;
    const int row = bx.off + N*w4.a;
    const int f = w4.b + (lane < nc ? lane : nc);  // this lane's face (lower face of cell f), f <= N
    const int cm2 = row + b2::wrapm(f - 2, N), cm1 = row + b2::wrapm(f - 1, N);
    const int c0 = row + b2::wrapm(f, N), cp1 = row + b2::wrapm(f + 1, N);
    R vl[8], vr[8];
#pragma unroll
    for (int q = 0; q < 7; q++) {
      const R *v = b.V + VAR[q]*n;
      const R a0 = v[cm2], a1 = v[cm1], a2 = v[c0], a3 = v[cp1];
      vl[VAR[q]] = a1 + R(0.5)*plm_slope<R, LIM>(a0, a1, a2);
      vr[VAR[q]] = a2 - R(0.5)*plm_slope<R, LIM>(a1, a2, a3);
    }
    vl[4] = vr[4] = (R)b.Vin[c0];
    FaceFlux<R> ff;
    b2::solve<R, 0, SOLVER>(vl, vr, gamma, ff);
    if (b2::face_reported<0>(bx, f, w4.a)) e |= ff.err;
    if (lane < nc) b2::store_rec<R, 0, SOLVER>(vl, vr, gamma, ff, b.FD, c0, n);
    R F[6], Fn[6];
#pragma unroll
    for (int q = 0; q < 6; q++) { F[q] = ff.f[b2::fidx(q)]; Fn[q] = __shfl_down(F[q], 1, 64); }
    const R cm = rmax(rabs(ff.Sl), rabs(ff.Sr));
    const R cmn = __shfl_down(cm, 1, 64);
    if (lane < nc) {
      const double dtdx = b.ctl[w4.box].dt/bx.dx;
#pragma unroll
      for (int q = 0; q < 6; q++) {
        double r = -dtdx*((double)Fn[q] - (double)F[q]);
        r += 0.0;
        b.dU[q*n + c0] = (R)r;
      }
      if (want_cdt) b.Cdt[c0] = (R)(0.5*((double)cm + (double)cmn)/bx.dx);
    }
  }
  if (e) atomicOr(b.err, e);
}

// ---------------------------------------------------------------------------
// y sweep + RK combination of U + C_dt maximum per item
// ---------------------------------------------------------------------------
template <typename R, int LIM, int SOLVER, int L>
__global__ __launch_bounds__(256)
void k2_sweepy(Batch2dBufs<R> b, RKW w, int first_stage, int want_cdt, R gamma)
{
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const size_t n = (size_t)b.ncell;
  constexpr int VAR[7] = {0, 1, 2, 3, 4, 6, 7};   // B2 comes from the face
  const int ngroups = (b.ny_items + 3) >> 2;
  uint32_t e = 0;
  for (int gi = blockIdx.x; gi < ngroups; gi += gridDim.x) {
    const int it = 4*gi + wv;
    if (it >= b.ny_items) continue;               // uniform across the wavefront
    const Item4 w4 = b.yitems[it];
    const BoxDesc bx = b.box[w4.box];
    const int NX = bx.nx, N = bx.ny, i = w4.a + lane;
    double cdt_max = -1.0;
    if (i < NX) {
      const int base = bx.off + i;
      const int s0 = w4.b*L, s1 = s0 + L < N ? s0 + L : N;
      const double dtdx = b.ctl[w4.box].dt/bx.dy;
      R Vm[7], Vc[7], Vp[7], vlc[7];
      {
        const int cm2 = base + b2::wrapm(s0 - 2, N)*NX, cm1 = base + b2::wrapm(s0 - 1, N)*NX;
        const int c0 = base + b2::wrapm(s0, N)*NX;
#pragma unroll
        for (int q = 0; q < 7; q++) {
          Vm[q] = b.V[VAR[q]*n + cm2]; Vc[q] = b.V[VAR[q]*n + cm1]; Vp[q] = b.V[VAR[q]*n + c0];
        }
#pragma unroll
        for (int q = 0; q < 7; q++) {
          vlc[q] = Vc[q] + R(0.5)*plm_slope<R, LIM>(Vm[q], Vc[q], Vp[q]);
          Vm[q] = Vc[q]; Vc[q] = Vp[q];
        }
      }
      double Fp[6] = {0, 0, 0, 0, 0, 0}, cmp = 0.0;
      for (int f = s0; f <= s1; f++) {
        const int cf = base + b2::wrapm(f, N)*NX;
        const int cn = base + b2::wrapm(f + 1, N)*NX;
        const int cc = base + b2::wrapm(f - 1, N)*NX;   // the cell completed by face f
        // accumulator and state of cell f-1 loaded before the Riemann solve
        // (U0, needed from stage 2 on, after it: register budget)
        double acc[6], u_old[6], cacc = 0.0;
        if (f > s0) {
#pragma unroll
          for (int q = 0; q < 6; q++) { acc[q] = (double)b.dU[q*n + cc]; u_old[q] = b.Uin[q*n + cc]; }
          if (want_cdt) cacc = (double)b.Cdt[cc];
        }
        R vl[8], vr[8];
#pragma unroll
        for (int q = 0; q < 7; q++) Vp[q] = b.V[VAR[q]*n + cn];
#pragma unroll
        for (int q = 0; q < 7; q++) {
          const R dv = plm_slope<R, LIM>(Vm[q], Vc[q], Vp[q]);
          vl[VAR[q]] = vlc[q];
          vr[VAR[q]] = Vc[q] - R(0.5)*dv;
          vlc[q] = Vc[q] + R(0.5)*dv;
        }
        vl[5] = vr[5] = (R)b.Vin[n + cf];
        FaceFlux<R> ff;
        b2::solve<R, 1, SOLVER>(vl, vr, gamma, ff);
        if (b2::face_reported<1>(bx, i, f)) e |= ff.err;
        if (f < s1) b2::store_rec<R, 1, SOLVER>(vl, vr, gamma, ff, b.FD + b2::nrec<SOLVER>()*n, cf, n);
        double F[6];
#pragma unroll
        for (int q = 0; q < 6; q++) F[q] = (double)ff.f[b2::fidx(q)];
        const double cm = (double)rmax(rabs(ff.Sl), rabs(ff.Sr));
        if (f > s0) {
#pragma unroll
          for (int q = 0; q < 6; q++) {
            double rr = -dtdx*(F[q] - Fp[q]);
            rr += 0.0;
            double du = acc[q];
            du += rr;
            const size_t o = q*n + cc;
            if (first_stage) {
              b.U[o] = u_old[q] + w.wR*du;
            } else {
              b.U[o] = w.w0*b.U0[o] + w.w1*u_old[q] + w.wR*du;
            }
          }
          if (want_cdt && b2::active(bx, i, f - 1)) {
            double cd = cacc;
            cd += 0.5*(cmp + cm)/bx.dy;
            cdt_max = cdt_max >= cd ? cdt_max : cd;
          }
        }
#pragma unroll
        for (int q = 0; q < 6; q++) Fp[q] = F[q];
        cmp = cm;
#pragma unroll
        for (int q = 0; q < 7; q++) { Vm[q] = Vc[q]; Vc[q] = Vp[q]; }
      }
    }
    if (want_cdt) {
      const double m = b2::wave_max(cdt_max);
      if (lane == 0) b.cdt_part[it] = m;
    }
  }
  if (e) atomicOr(b.err, e);
}

// ---------------------------------------------------------------------------
// Ez at the lower-left corner of every cell: X1a + X2b
// ---------------------------------------------------------------------------
template <typename R, int LIMV, int NQ>
__global__ __launch_bounds__(256)
void k2_emf(Batch2dBufs<R> b)
{
  const size_t n = (size_t)b.ncell;
  const double *Bx = b.Vin, *By = b.Vin + n;
  const R *FDx = b.FD, *FDy = b.FD + (NQ == 4 ? 6 : 4)*n;
  for (int gi = blockIdx.x; gi < b.ntiles; gi += gridDim.x) {
    const Item2 t2 = b.tiles[gi];
    const BoxDesc bx = b.box[t2.box];
    const int NX = bx.nx, NY = bx.ny, q = t2.start + threadIdx.x;
    if (q >= NX*NY) continue;
    const int i = q % NX, j = q / NX, c = bx.off + q;
    int xo[4], yo[4];
#pragma unroll
    for (int o = 0; o < 4; o++) {
      xo[o] = bx.off + b2::wrapm(i + o - 2, NX) + NX*j;
      yo[o] = bx.off + i + NX*b2::wrapm(j + o - 2, NY);
    }
    R cx[NQ], cy[NQ], cxj[NQ], cyi[NQ];
    if constexpr (NQ == 4) {
#pragma unroll
      for (int s = 0; s < NQ; s++) {
        cx[s] = FDx[(2 + s)*n + c];  cy[s] = FDy[(2 + s)*n + c];
        cxj[s] = FDx[(2 + s)*n + yo[1]];
        cyi[s] = FDy[(2 + s)*n + xo[1]];
      }
    } else {
      uct_hll_coef<R>(FDx[2*n + c], FDx[3*n + c], cx[0], cx[1], cx[2]);
      uct_hll_coef<R>(FDy[2*n + c], FDy[3*n + c], cy[0], cy[1], cy[2]);
      uct_hll_coef<R>(FDx[2*n + yo[1]], FDx[3*n + yo[1]], cxj[0], cxj[1], cxj[2]);
      uct_hll_coef<R>(FDy[2*n + xo[1]], FDy[3*n + xo[1]], cyi[0], cyi[1], cyi[2]);
    }
    R q4[4], e4[4];
#pragma unroll
    for (int o = 0; o < 4; o++) { q4[o] = (R)By[xo[o]]; e4[o] = FDy[1*n + xo[o]]; }
    R ez = b2::uct_val<R, LIMV, NQ>(q4, e4, cxj, cx, +1);             // X1a
#pragma unroll
    for (int o = 0; o < 4; o++) { q4[o] = (R)Bx[yo[o]]; e4[o] = FDx[0*n + yo[o]]; }
    ez += b2::uct_val<R, LIMV, NQ>(q4, e4, cyi, cy, -1);              // X2b
    b.E[c] = ez;
  }
}

// ---------------------------------------------------------------------------
// Face update + cell-centred field + ConsToPrim
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(256)
void k2_finish(Batch2dBufs<R> b, RKW w, int first_stage, double gmm1, double small_rho,
               double small_prs)
{
  const size_t n = (size_t)b.ncell;
  uint32_t eall = 0;
  for (int gi = blockIdx.x; gi < b.ntiles; gi += gridDim.x) {
    const Item2 t2 = b.tiles[gi];
    const BoxDesc bx = b.box[t2.box];
    const int NX = bx.nx, NY = bx.ny, q = t2.start + threadIdx.x;
    if (q >= NX*NY) continue;
    const int i = q % NX, j = q / NX, c = bx.off + q;
    const int ip = b2::wrapm(i + 1, NX), jp = b2::wrapm(j + 1, NY);
    const int cxp = bx.off + ip + NX*j, cyp = bx.off + i + NX*jp, cxyp = bx.off + ip + NX*jp;
    const double e00 = b.E[c], e10 = b.E[cxp], e01 = b.E[cyp], e11 = b.E[cxyp];
    const double dt = b.ctl[t2.box].dt;
    const double dtdx1 = dt/bx.dx, dtdx2 = dt/bx.dy;
    // CT_Update, Cartesian 2D (ct_update.cpp:189-258)
    double curlE = 0.0 + dtdx2*(e01 - e00);
    const double rbx_lo = -curlE;
    curlE = 0.0 + dtdx2*(e11 - e10);
    const double rbx_hi = -curlE;                 // face (i+1, j)
    curlE = - dtdx1*(e10 - e00);
    const double rby_lo = -curlE;
    curlE = - dtdx1*(e11 - e01);
    const double rby_hi = -curlE;                 // face (i, j+1)
    const double rl[2] = {rbx_lo, rby_lo}, rh[2] = {rbx_hi, rby_hi};
    const int up[2] = {cxp, cyp};
    double b_lo[2], b_hi[2];
#pragma unroll
    for (int d = 0; d < 2; d++) {
      const double *vin = b.Vin + d*n;
      if (first_stage) {
        b_lo[d] = vin[c] + w.wR*rl[d];
        b_hi[d] = vin[up[d]] + w.wR*rh[d];
      } else {
        const double *bs0 = b.Bs0 + d*n;
        b_lo[d] = w.w0*bs0[c] + w.w1*vin[c] + w.wR*rl[d];
        b_hi[d] = w.w0*bs0[up[d]] + w.w1*vin[up[d]] + w.wR*rh[d];
      }
      b.Vout[d*n + c] = b_lo[d];
    }
    double bb[3], u[5], v[8];
    bb[0] = 0.5*(b_lo[0] + b_hi[0]);
    bb[1] = 0.5*(b_lo[1] + b_hi[1]);
    bb[2] = b.U[4*n + c];
    u[0] = b.U[0*n + c]; u[1] = b.U[1*n + c]; u[2] = b.U[2*n + c]; u[3] = b.U[3*n + c];
    u[4] = b.U[5*n + c];
    const uint32_t e = mhd_cons_to_prim(u, bb, gmm1, small_rho, small_prs, v);
    if (e) {
      b.U[0*n + c] = u[0]; b.U[1*n + c] = u[1]; b.U[2*n + c] = u[2]; b.U[3*n + c] = u[3];
      b.U[5*n + c] = u[4];
      if (b2::active(bx, i, j)) eall |= e;
    }
    if (b2::active(bx, i, j)) zone_report(b.zr, c, e);
#pragma unroll
    for (int s = 0; s < 8; s++) b.V[s*n + c] = (R)v[s];
  }
  if (eall) atomicOr(b.err, eall);
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
template <typename R, int LIM, int SOLVER, int L>
static void b2_launch_t(int phase, const Batch2dBufs<R> &b, RKW w, int first_stage, int want_dt,
                        int blocks, double gamma, double small_rho, double small_prs, hipStream_t s)
{
  switch (phase) {
  case B2_SWEEP_X:
    hipLaunchKernelGGL((k2_sweepx<R, LIM, SOLVER>), dim3(blocks), dim3(256), 0, s, b, (R)gamma, want_dt);
    break;
  case B2_SWEEP_Y:
    hipLaunchKernelGGL((k2_sweepy<R, LIM, SOLVER, L>), dim3(blocks), dim3(256), 0, s, b, w, first_stage,
                       want_dt, (R)gamma);
    break;
  case B2_EMF:
    hipLaunchKernelGGL((k2_emf<R, LIM, SOLVER == 4 ? 4 : 3>), dim3(blocks), dim3(256), 0, s, b);
    break;
  case B2_FINISH:
    hipLaunchKernelGGL((k2_finish<R>), dim3(blocks), dim3(256), 0, s, b, w, first_stage, gamma - 1.0,
                       small_rho, small_prs);
    break;
  default:
    break;
  }
}

template <typename R, int LIM, int SOLVER>
static void b2_launch_l(int phase, const Batch2dBufs<R> &b, RKW w, int first_stage, int want_dt,
                        int march, int blocks, double gamma, double small_rho, double small_prs,
                        hipStream_t s)
{
  if (march == 8)       b2_launch_t<R, LIM, SOLVER, 8>(phase, b, w, first_stage, want_dt, blocks, gamma, small_rho, small_prs, s);
  else if (march == 32) b2_launch_t<R, LIM, SOLVER, 32>(phase, b, w, first_stage, want_dt, blocks, gamma, small_rho, small_prs, s);
  else                  b2_launch_t<R, LIM, SOLVER, 16>(phase, b, w, first_stage, want_dt, blocks, gamma, small_rho, small_prs, s);
}

template <typename R>
void batch2d_launch(int phase, const Batch2dBufs<R> &b, RKW w, int first_stage, int want_dt,
                    int limiter, int solver, int march, int blocks, double gamma,
                    double small_rho, double small_prs, hipStream_t s)
{
  if (limiter == MC_LIM) {
    if (solver == 4)      b2_launch_l<R, MC_LIM, 4>(phase, b, w, first_stage, want_dt, march, blocks, gamma, small_rho, small_prs, s);
    else if (solver == 3) b2_launch_l<R, MC_LIM, 3>(phase, b, w, first_stage, want_dt, march, blocks, gamma, small_rho, small_prs, s);
    else if (solver == 2) b2_launch_l<R, MC_LIM, 2>(phase, b, w, first_stage, want_dt, march, blocks, gamma, small_rho, small_prs, s);
    else                  b2_launch_l<R, MC_LIM, 1>(phase, b, w, first_stage, want_dt, march, blocks, gamma, small_rho, small_prs, s);
  } else {
    if (solver == 4)      b2_launch_l<R, VANLEER_LIM, 4>(phase, b, w, first_stage, want_dt, march, blocks, gamma, small_rho, small_prs, s);
    else if (solver == 3) b2_launch_l<R, VANLEER_LIM, 3>(phase, b, w, first_stage, want_dt, march, blocks, gamma, small_rho, small_prs, s);
    else if (solver == 2) b2_launch_l<R, VANLEER_LIM, 2>(phase, b, w, first_stage, want_dt, march, blocks, gamma, small_rho, small_prs, s);
    else                  b2_launch_l<R, VANLEER_LIM, 1>(phase, b, w, first_stage, want_dt, march, blocks, gamma, small_rho, small_prs, s);
  }
}

}  // namespace gpl
