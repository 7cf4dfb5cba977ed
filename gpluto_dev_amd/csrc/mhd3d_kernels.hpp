// 3D ideal-MHD stage update on a periodic box: PLM + HLLD/HLL + UCT_HLL
// constrained transport + RK2/RK3, for gfx950.
//
// Device data model (one GPU, whole periodic box, no ghost zones: periodic
// neighbours are addressed by wrapped indices, which is bitwise equivalent to
// the reference's ghost copies because every ghost is an exact image):
//   V [8][ncell]   primitives rho, v1..3, B1..3 (face averages), prs   (type R)
//   U [5][ncell]   conserved rho, m1..3, E (FP64); the cell-centred field of
//                  Uc is always the face average (ct_field_average.cpp:182),
//                  so it is never stored
//   Vs[3][ncell]   face-normal field of the lower face of every cell (FP64)
//   F [3][5][ncell], CM[3][ncell]   interface flux / max speed  (type R)
//   FD[3][4][ncell] UCT face data et, eb, min(0,Sl), max(0,Sr)  (type R)
//                   (UCT_HLLD: FD[3][6][ncell], et, eb, aL, aR, dL, dR)
//   E [3][ncell]   edge EMFs Ex(i,j-1/2,k-1/2), Ey(i-1/2,j,k-1/2),
//                  Ez(i-1/2,j-1/2,k)                            (type R)
// Cell index c = i + nx*(j + ny*k) (reference Ary4D order, i fastest).
//
// Two kernel paths compute the same numbers:
//  * fused (nx <= 512): one kernel per sweep direction reconstructs, solves
//    the Riemann problem and accumulates the flux divergence in place, with no
//    interface flux written to HBM -- x: one block per periodic x-row, the row
//    staged in LDS; y, z: each thread marches a register window along the
//    column; the z kernel also applies the RK combination of U;
//  * split: one thread per interface, fluxes through HBM, a separate
//    divergence/RK kernel (any box).
// This header holds the templates; mhd3d_fp64.hip and mhd3d_mixed.hip
// instantiate them with their own compiler flags.
#pragma once
#include <type_traits>
#include "c2p_net.hpp"
#include "physics.hpp"
#include "mhd3d.hpp"

namespace gpl {

__device__ __forceinline__ int wrapm(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

// base[idx] with a 32-bit byte offset from a uniform base: the load / store
// form with the array base in scalar registers and one 32-bit offset per
// lane, instead of a 64-bit address per lane and array.  Every array slot of
// a box is below 4 GiB (the engine refuses 2^29 cells or more).
template <typename T>
__device__ __forceinline__ T ldo(const T *base, unsigned idx)
{
  return *(const T *)((const char *)base + idx*(unsigned)sizeof(T));
}
template <typename T>
__device__ __forceinline__ void sto(T *base, unsigned idx, T v)
{
  *(T *)((char *)base + idx*(unsigned)sizeof(T)) = v;
}

struct Cell { int i, j, k; };

__device__ __forceinline__ bool is_active(const Box3 &g, int i, int j, int k)
{
  return i >= g.lo[0] && i <= g.hi[0] && j >= g.lo[1] && j <= g.hi[1] && k >= g.lo[2] && k <= g.hi[2];
}

// The reference solves the faces of the active cells and the upper boundary
// face along a sweep, over boxes widened by two ghost layers across it
// (update_stage.cpp:80-99); error bits count from those faces only.
template <int D>
__device__ __forceinline__ bool face_reported(const Box3 &g, int i, int j, int k)
{
  const int p[3] = {i, j, k};
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const int lo = d == D ? g.lo[d] : g.lo[d] - 2, hi = d == D ? g.hi[d] + 1 : g.hi[d] + 2;
    if (p[d] < lo || p[d] > hi) return false;
  }
  return true;
}

__device__ __forceinline__ Cell cell_of(int c, const Box3 &g)
{
  Cell r;
  r.i = c % g.nx;
  int t = c / g.nx;
  r.j = t % g.ny;
  r.k = t / g.ny;
  return r;
}

__device__ __forceinline__ int cid(int i, int j, int k, const Box3 &g)
{
  return i + g.nx*(j + g.ny*k);
}

// index of the cell displaced by o along direction D (periodic)
template <int D>
__device__ __forceinline__ int shift(const Cell &p, int o, const Box3 &g)
{
  if (D == 0) return cid(wrapm(p.i + o, g.nx), p.j, p.k, g);
  if (D == 1) return cid(p.i, wrapm(p.j + o, g.ny), p.k, g);
  return cid(p.i, p.j, wrapm(p.k + o, g.nz), g);
}

template <typename T>
__device__ __forceinline__ T block_max(T v)
{
  __shared__ T red[32];
  for (int o = 32; o > 0; o >>= 1) {
    T w = __shfl_xor(v, o, 64);
    v = v >= w ? v : w;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  const int nw = (blockDim.x + 63) >> 6;
  if (threadIdx.x == 0) {
    T m = red[0];
    for (int w = 1; w < nw; w++) m = m >= red[w] ? m : red[w];
    red[0] = m;
  }
  __syncthreads();
  return red[0];
}

// Riemann solver of the physics module: PH 0 = MHD (SOLVER 3 HLLD, 2 HLLC,
// 1 HLL), PH 1 = RMHD (SOLVER 3 HLLD, 2 HLLC, 1 HLL, all with the exact
// quartic fast speeds; FP64 only).
template <typename R, int N, int SOLVER, int PH>
__device__ __forceinline__ void riemann(const R *vl, const R *vr, R gamma, FaceFlux<R> &ff,
                                        const Box3 &g)
{
  if constexpr (PH == 1 && SOLVER == 3) rmhd_hlld<N>(vl, vr, gamma, g.small_rho, g.small_prs, ff);
  else if constexpr (PH == 1 && SOLVER == 2) rmhd_hllc<N>(vl, vr, gamma, ff);
  else if constexpr (PH == 1) rmhd_hll<N>(vl, vr, gamma, ff);
  else if constexpr (SOLVER == 3 || SOLVER == 4) mhd_hlld<R, N>(vl, vr, gamma, ff);
  else if constexpr (SOLVER == 2) mhd_hllc<R, N>(vl, vr, gamma, ff);
  else mhd_hll<R, N>(vl, vr, gamma, ff);
}

// Face records of the UCT EMF average per direction: et, eb and the two
// clipped signal speeds aL_1 = min(0, Sl), aR_1 = max(0, Sr) for UCT_HLL,
// from which the EMF kernels recompute the weights aL, aR and d (= dL = dR)
// with CT_Store_UCTHll's own expressions (uct_hll_coef: the same operations
// on the same values, so the same bits, and one record per face less through
// HBM); et, eb, aL, aR, dL, dR for UCT_HLLD (SOLVER 4: HLLD with
// CT_Store_UCTHlld), whose weights come from the HLLD states.
template <int SOLVER> __host__ __device__ constexpr int uct_nrec() { return SOLVER == 4 ? 6 : 4; }

// The NQ weights of one face for the EMF average from its records fdir[2..]:
// computed from the clipped speeds (UCT_HLL, NQ = 3) or loaded (UCT_HLLD).
template <typename R, int NQ>
__device__ __forceinline__ void uct_weights(const R *fdir, unsigned c, size_t n, R *w)
{
  if constexpr (NQ == 4) {
#pragma unroll
    for (int q = 0; q < 4; q++) w[q] = ldo(fdir + (2 + q)*n, c);
  } else {
    uct_hll_coef<R>(ldo(fdir + 2*n, c), ldo(fdir + 3*n, c), w[0], w[1], w[2]);
  }
}

template <typename R, int N, int SOLVER>
__device__ __forceinline__ void uct_store(const R *vl, const R *vr, R gamma, const FaceFlux<R> &ff,
                                          R *fd, unsigned c, size_t n)
{
  if constexpr (SOLVER == 4) {
    R et, eb, aL, aR, dL, dR;
    uct_hlld_store<R, N>(vl, vr, gamma, ff.Sl, ff.Sr, et, eb, aL, aR, dL, dR);
    sto(fd, c, et); sto(fd + n, c, eb); sto(fd + 2*n, c, aL); sto(fd + 3*n, c, aR);
    sto(fd + 4*n, c, dL); sto(fd + 5*n, c, dR);
  } else {
    R et, eb, aL, aR, dd;
    uct_hll_store<R, N>(vl, vr, ff.Sl, ff.Sr, et, eb, aL, aR, dd);
    sto(fd, c, et); sto(fd + n, c, eb);
    sto(fd + 2*n, c, rmin(R(0), ff.Sl)); sto(fd + 3*n, c, rmax(R(0), ff.Sr));
  }
}

// VelocityLimiter (Src/vel_limiter.hpp:14-31, applied in plm_states.cpp:158-
// 161): a cell whose reconstructed velocity is superluminal on either face
// is flattened (both states = the cell value).  lo = vR(i), hi = vL(i+1),
// indices 1..3 are the velocity components.
template <typename R>
__device__ __forceinline__ bool superluminal(const R *lo, const R *hi)
{
  const R v2m = lo[1]*lo[1] + lo[2]*lo[2] + lo[3]*lo[3];
  const R v2p = hi[1]*hi[1] + hi[2]*hi[2] + hi[3]*hi[3];
  return v2m >= R(1) || v2p >= R(1);
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2: block b runs on XCD b % 8.  Renumbering gives every XCD a contiguous
// range of logical tiles, so tiles that share stencil rows share an L2.
__device__ __forceinline__ int xcd_tile(int b, int nb)
{
  constexpr int NXCD = 8;
  const int x = b % NXCD, q = nb / NXCD, r = nb % NXCD;
  return x*q + (x < r ? x : r) + b / NXCD;
}

// Runs of RUN consecutive logical blocks per XCD: block b (on XCD b % 8) of
// group g = b / (8 RUN) takes logical block g 8 RUN + (b % 8) RUN + (b / 8) % RUN.
// A one-row block then reads the next row of its plane (RUN - 1 times in
// RUN) and the same row of the next plane (when a plane holds a multiple of
// 8 RUN blocks) through its own L2.  The last partial group keeps its numbers.
template <int RUN>
__device__ __forceinline__ int xcd_runs(int b, int nb)
{
  constexpr int G = 8*RUN;
  if (b >= nb - nb % G) return b;
  const int g = b / G, r = b - g*G;
  return g*G + (r % 8)*RUN + r / 8;
}

// ---------------------------------------------------------------------------
// Sweep along D: one thread per interface (lower face of cell c).
// DataToSweep + Reconstruct (plm_states.cpp:99-189) + Riemann solver
// (hlld.cpp:100-500) + CT_Store_UCTHll (ct_store.hpp:156).
// ---------------------------------------------------------------------------
template <typename R, int D, int LIM, int SOLVER>
__global__ __launch_bounds__(256)
void k_sweep(Box3 g, const R *__restrict__ V, const double *__restrict__ Vs,
             R *__restrict__ F, R *__restrict__ CM, R *__restrict__ FD,
             R gamma, uint32_t *__restrict__ err)
{
  const int c = blockIdx.x*blockDim.x + threadIdx.x;
  if (c >= g.ncell) return;
  const size_t n = (size_t)g.ncell;
  const Cell p = cell_of(c, g);
  const int cm2 = shift<D>(p, -2, g), cm1 = shift<D>(p, -1, g), cp1 = shift<D>(p, 1, g);
  R vl[8], vr[8];
#pragma unroll
  for (int nv = 0; nv < 8; nv++) {
    if (nv == 4 + D) continue;   // overwritten by the face value below
    const R a = V[nv*n + cm2], b = V[nv*n + cm1], q = V[nv*n + c], d = V[nv*n + cp1];
    const R sl = plm_slope<R, LIM>(a, b, q);
    const R sr = plm_slope<R, LIM>(b, q, d);
    vl[nv] = b + R(0.5)*sl;
    vr[nv] = q - R(0.5)*sr;
  }
  // normal field from the face (plm_states.cpp:184-189)
  vl[4 + D] = vr[4 + D] = (R)Vs[D*n + c];
  FaceFlux<R> ff;
  riemann<R, D, SOLVER, 0>(vl, vr, gamma, ff, g);
  R *Fd = F + (size_t)D*5*n;
  Fd[0*n + c] = ff.f[0];
  Fd[1*n + c] = ff.f[1];
  Fd[2*n + c] = ff.f[2];
  Fd[3*n + c] = ff.f[3];
  Fd[4*n + c] = ff.f[7];
  CM[(size_t)D*n + c] = rmax(rabs(ff.Sl), rabs(ff.Sr));
  uct_store<R, D, SOLVER>(vl, vr, gamma, ff, FD + (size_t)D*uct_nrec<SOLVER>()*n, c, n);
  if (ff.err && face_reported<D>(g, p.i, p.j, p.k)) atomicOr(err, ff.err);
}

// ---------------------------------------------------------------------------
// Flux divergence + RK combination of the hydro conserved variables.
// RightHandSide (rhs.cpp:384-386,523; :671,756; :900,981), dUc accumulation
// in x, y, z order (update_stage.cpp:173-184,272-284,369-380) and the RK
// stage update (rk_step.cpp:366-374,509-518,670-708).  Stage 1 also emits the
// per-block maximum of C_dt for getMax (update_stage.cpp:446-458).
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(256)
void k_rhs_update(Box3 g, const R *__restrict__ F, const R *__restrict__ CM,
                  const double *Uin, double *U, const double *U0,
                  const StepCtl *__restrict__ ctl, RKW w, int first_stage,
                  double *__restrict__ cdt_part)
{
  const int c = blockIdx.x*blockDim.x + threadIdx.x;
  double cdt = -1.0;
  if (c < g.ncell) {
    const size_t n = (size_t)g.ncell;
    const Cell p = cell_of(c, g);
    const int ci = shift<0>(p, 1, g), cj = shift<1>(p, 1, g), ck = shift<2>(p, 1, g);
    const double dt = ctl->dt;
    const double dtdx = dt/g.dx, dtdy = dt/g.dy, dtdz = dt/g.dz;
    const R *Fx = F, *Fy = F + 5*n, *Fz = F + 10*n;
#pragma unroll
    for (int nv = 0; nv < 5; nv++) {
      const size_t o = nv*n;
      double rx = -dtdx*((double)Fx[o + ci] - (double)Fx[o + c]);
      rx += 0.0;
      double ry = -dtdy*((double)Fy[o + cj] - (double)Fy[o + c]);
      ry += 0.0;
      double rz = -dtdz*((double)Fz[o + ck] - (double)Fz[o + c]);
      rz += 0.0;
      double du = rx;
      du += ry;
      du += rz;
      if (first_stage) U[o + c] = Uin[o + c] + w.wR*du;
      else             U[o + c] = w.w0*U0[o + c] + w.w1*Uin[o + c] + w.wR*du;
    }
    if (cdt_part) {
      cdt = 0.5*((double)CM[c] + (double)CM[ci])/g.dx;
      cdt += 0.5*((double)CM[n + c] + (double)CM[n + cj])/g.dy;
      cdt += 0.5*((double)CM[2*n + c] + (double)CM[2*n + ck])/g.dz;
    }
  }
  if (cdt_part) {
    const double m = block_max(cdt);
    if (threadIdx.x == 0) cdt_part[blockIdx.x] = m;
  }
}

// ---------------------------------------------------------------------------
// Fused x sweep: one block per periodic x-row (blockDim.x == nx <= 512).
// DataToSweep/Reconstruct (plm_states.cpp:99-189) with every cell's slope
// computed once and shared through LDS, HLLD (hlld.cpp:100-500), CT store,
// and RightHandSide's x part (rhs.cpp:384-386, 523): dUc = rhs_x
// (update_stage.cpp:173-184), C_dt = 0.5 (cmax_i + cmax_i+1)/dx.
// ---------------------------------------------------------------------------
template <typename R, int LIM, int SOLVER, int PH>
__global__ __launch_bounds__(512)
void k_sweepx_fused(Box3 g, const R *__restrict__ V, const double *__restrict__ Vs,
                    R *__restrict__ FD, R *__restrict__ dU, R *__restrict__ Cdt,
                    const StepCtl *__restrict__ ctl, R gamma, int want_cdt,
                    uint32_t *__restrict__ err)
{
  extern __shared__ __align__(16) unsigned char gpl_smem[];
  const int nx = g.nx, i = threadIdx.x;
  R *sV = (R *)gpl_smem;           // [7][nx] primitives of the row
  R *sD = sV + 7*nx;               // [7][nx] left state vL(i+1) made by cell i
  const size_t n = (size_t)g.ncell;
  const unsigned row = blockIdx.x + (unsigned)g.ny*g.kb;   // j + ny*k
  const unsigned c = i + nx*row;
  const int im = i == 0 ? nx - 1 : i - 1, ip = i == nx - 1 ? 0 : i + 1;
  constexpr int VAR[7] = {0, 1, 2, 3, 5, 6, 7};   // B1 comes from the face
  R v0[7], dv[7];
#pragma unroll
  for (int q = 0; q < 7; q++) { v0[q] = ldo(V + VAR[q]*n, c); sV[q*nx + i] = v0[q]; }
  const R bn = (R)ldo(Vs, c);
  __syncthreads();
  R lo[7], hi[7];                  // this cell's vR(i) and vL(i+1)
#pragma unroll
  for (int q = 0; q < 7; q++) {
    dv[q] = plm_slope<R, LIM>(sV[q*nx + im], v0[q], sV[q*nx + ip]);
    lo[q] = v0[q] - R(0.5)*dv[q];
    hi[q] = v0[q] + R(0.5)*dv[q];
  }
  if (PH == 1 && superluminal(lo, hi)) {
#pragma unroll
    for (int q = 0; q < 7; q++) lo[q] = hi[q] = v0[q];
  }
#pragma unroll
  for (int q = 0; q < 7; q++) sD[q*nx + i] = hi[q];
  __syncthreads();
  R vl[8], vr[8];
#pragma unroll
  for (int q = 0; q < 7; q++) {
    vl[VAR[q]] = sD[q*nx + im];
    vr[VAR[q]] = lo[q];
  }
  vl[4] = vr[4] = bn;
  FaceFlux<R> ff;
  riemann<R, 0, SOLVER, PH>(vl, vr, gamma, ff, g);
  uct_store<R, 0, SOLVER>(vl, vr, gamma, ff, FD, c, n);
  __syncthreads();                 // everyone is done with sV/sD
  R *sF = sV;                      // [6][nx]: hydro flux and cmax of face i
  sF[0*nx + i] = ff.f[0]; sF[1*nx + i] = ff.f[1]; sF[2*nx + i] = ff.f[2];
  sF[3*nx + i] = ff.f[3]; sF[4*nx + i] = ff.f[7];
  sF[5*nx + i] = rmax(rabs(ff.Sl), rabs(ff.Sr));
  __syncthreads();
  const double dtdx = ctl->dt/g.dx;
#pragma unroll
  for (int nv = 0; nv < 5; nv++) {
    double r = -dtdx*((double)sF[nv*nx + ip] - (double)sF[nv*nx + i]);
    if (PH == 0) r += 0.0;         // MHD adds src[] = 0 (rhs.cpp:519); RMHD+CT has no source
    sto(dU + nv*n, c, (R)r);
  }
  if (want_cdt) sto(Cdt, c, (R)(0.5*((double)sF[5*nx + i] + (double)sF[5*nx + ip])/g.dx));
  if (ff.err && face_reported<0>(g, i, (int)(row % g.ny), (int)(row / g.ny)))
    atomicOr(err, ff.err);
}

// ---------------------------------------------------------------------------
// Fused x + y sweep: one block per (plane k, y-segment of L rows) with one
// thread per x-cell (blockDim.x == nx).  At row f the block does
//  * the x sweep of row f through LDS, as k_sweepx_fused: every slope once,
//    the left state of face i+1 made by cell i, HLLD, UCT store;
//  * the y face f from the register window, as k_sweep_march;
// and completes cell (i, f-1) from rhs_x of row f-1, kept in registers:
// dUc = rhs_x; dUc += rhs_y and C_dt likewise (update_stage.cpp:173-184,
// 272-284), the reference's order.  dU and C_dt are written once per stage
// instead of written, read and written again, and V is read once.  The
// three LDS rows are disjoint, so a row needs three barriers.
// ---------------------------------------------------------------------------
template <typename R, int LIM, int SOLVER, int L, int PH>
__global__ __launch_bounds__(512)
void k_sweep_xy(Box3 g, const R *__restrict__ V, const double *__restrict__ Vs,
                R *__restrict__ FD, R *__restrict__ dU, R *__restrict__ Cdt,
                const StepCtl *__restrict__ ctl, R gamma, int want_cdt,
                uint32_t *__restrict__ err)
{
  extern __shared__ __align__(16) unsigned char gpl_smem[];
  const int nx = g.nx, ny = g.ny, i = threadIdx.x;
  R *sV = (R *)gpl_smem;           // [7][nx] primitives of row f
  R *sD = sV + 7*nx;               // [7][nx] left state vL(i+1) made by cell i
  R *sF = sD + 7*nx;               // [6][nx] x flux and cmax of face i
  const size_t n = (size_t)g.ncell;
  const int nseg = (ny + L - 1)/L;
  const int k = blockIdx.x / nseg, seg = blockIdx.x - k*nseg;
  const int s0 = seg*L, s1 = s0 + L < ny ? s0 + L : ny;
  const int im = i == 0 ? nx - 1 : i - 1, ip = i == nx - 1 ? 0 : i + 1;
  const unsigned base = i + (unsigned)nx*ny*k;
  // y window order: B1 is a cell value, B2 comes from the face
  constexpr int VY[7] = {0, 1, 2, 3, 4, 6, 7};
  const R *Vq[7];
#pragma unroll
  for (int q = 0; q < 7; q++) Vq[q] = V + VY[q]*n;
  const R *V5 = V + 5*n;
  const double *Bxf = Vs, *Byf = Vs + n;
  R *fdx = FD, *fdy = FD + uct_nrec<SOLVER>()*n;
  R *dUq[5];
#pragma unroll
  for (int nv = 0; nv < 5; nv++) dUq[nv] = dU + nv*n;
  const double dtdx = ctl->dt/g.dx, dtdy = ctl->dt/g.dy;
  uint32_t e = 0;
  R Vm[7], Vc[7], Vp[7], vlc[7];
  {
    const unsigned cm2 = base + wrapm(s0 - 2, ny)*nx, cm1 = base + wrapm(s0 - 1, ny)*nx;
    const unsigned c0 = base + wrapm(s0, ny)*nx;
#pragma unroll
    for (int q = 0; q < 7; q++) { Vm[q] = ldo(Vq[q], cm2); Vc[q] = ldo(Vq[q], cm1); Vp[q] = ldo(Vq[q], c0); }
    R lo[7];
#pragma unroll
    for (int q = 0; q < 7; q++) {
      const R dv = plm_slope<R, LIM>(Vm[q], Vc[q], Vp[q]);
      lo[q] = Vc[q] - R(0.5)*dv;
      vlc[q] = Vc[q] + R(0.5)*dv;
    }
    if (PH == 1 && superluminal(lo, vlc)) {
#pragma unroll
      for (int q = 0; q < 7; q++) vlc[q] = Vc[q];
    }
#pragma unroll
    for (int q = 0; q < 7; q++) { Vm[q] = Vc[q]; Vc[q] = Vp[q]; }
  }
  double Fp[5] = {0, 0, 0, 0, 0}, cmp = 0.0;
  R rx[5] = {0, 0, 0, 0, 0}, cx = 0;   // rhs_x and the x part of C_dt of row f-1
  for (int f = s0; f <= s1; f++) {
    const unsigned cf = base + wrapm(f, ny)*nx;
    const unsigned cn = base + wrapm(f + 1, ny)*nx;
    const unsigned cc = base + wrapm(f - 1, ny)*nx;    // the cell completed by face f
    // the next row and the y face are loaded before the x sweep's barriers
#pragma unroll
    for (int q = 0; q < 7; q++) Vp[q] = ldo(Vq[q], cn);
    const R byf = (R)ldo(Byf, cf);
    R rxn[5], cxn = 0;
    if (f < s1) {                      // block-uniform
      // ---- x sweep of row f (Vc holds V(f)) ----
      R v0[7];                         // x order: 0 1 2 3 5 6 7
      v0[0] = Vc[0]; v0[1] = Vc[1]; v0[2] = Vc[2]; v0[3] = Vc[3];
      v0[4] = ldo(V5, cf); v0[5] = Vc[5]; v0[6] = Vc[6];
      const R bn = (R)ldo(Bxf, cf);
#pragma unroll
      for (int q = 0; q < 7; q++) sV[q*nx + i] = v0[q];
      __syncthreads();
      R lo[7], hi[7];
#pragma unroll
      for (int q = 0; q < 7; q++) {
        const R dv = plm_slope<R, LIM>(sV[q*nx + im], v0[q], sV[q*nx + ip]);
        lo[q] = v0[q] - R(0.5)*dv;
        hi[q] = v0[q] + R(0.5)*dv;
      }
      if (PH == 1 && superluminal(lo, hi)) {
#pragma unroll
        for (int q = 0; q < 7; q++) lo[q] = hi[q] = v0[q];
      }
#pragma unroll
      for (int q = 0; q < 7; q++) sD[q*nx + i] = hi[q];
      __syncthreads();
      constexpr int VX[7] = {0, 1, 2, 3, 5, 6, 7};
      R vl[8], vr[8];
#pragma unroll
      for (int q = 0; q < 7; q++) { vl[VX[q]] = sD[q*nx + im]; vr[VX[q]] = lo[q]; }
      vl[4] = vr[4] = bn;
      FaceFlux<R> ff;
      riemann<R, 0, SOLVER, PH>(vl, vr, gamma, ff, g);
      if (face_reported<0>(g, i, f, k)) e |= ff.err;
      uct_store<R, 0, SOLVER>(vl, vr, gamma, ff, fdx, cf, n);
      sF[0*nx + i] = ff.f[0]; sF[1*nx + i] = ff.f[1]; sF[2*nx + i] = ff.f[2];
      sF[3*nx + i] = ff.f[3]; sF[4*nx + i] = ff.f[7];
      sF[5*nx + i] = rmax(rabs(ff.Sl), rabs(ff.Sr));
      __syncthreads();
#pragma unroll
      for (int nv = 0; nv < 5; nv++) {
        double r = -dtdx*((double)sF[nv*nx + ip] - (double)sF[nv*nx + i]);
        if (PH == 0) r += 0.0;         // MHD adds src[] = 0 (rhs.cpp:519); RMHD+CT has no source
        rxn[nv] = (R)r;
      }
      if (want_cdt) cxn = (R)(0.5*((double)sF[5*nx + i] + (double)sF[5*nx + ip])/g.dx);
    }
    // ---- y face f ----
    R vl[8], vr[8];
#pragma unroll
    for (int q = 0; q < 7; q++) {
      const R dv = plm_slope<R, LIM>(Vm[q], Vc[q], Vp[q]);
      vl[VY[q]] = vlc[q];
      vr[VY[q]] = Vc[q] - R(0.5)*dv;
      vlc[q] = Vc[q] + R(0.5)*dv;
    }
    if (PH == 1) {
      const R lo[4] = {R(0), vr[1], vr[2], vr[3]};
      if (superluminal(lo, vlc)) {
#pragma unroll
        for (int q = 0; q < 7; q++) { vr[VY[q]] = Vc[q]; vlc[q] = Vc[q]; }
      }
    }
    vl[5] = vr[5] = byf;
    FaceFlux<R> ff;
    riemann<R, 1, SOLVER, PH>(vl, vr, gamma, ff, g);
    if (face_reported<1>(g, i, wrapm(f, ny), k)) e |= ff.err;
    if (f < s1) uct_store<R, 1, SOLVER>(vl, vr, gamma, ff, fdy, cf, n);
    const double F[5] = {(double)ff.f[0], (double)ff.f[1], (double)ff.f[2], (double)ff.f[3],
                         (double)ff.f[7]};
    const double cm = (double)rmax(rabs(ff.Sl), rabs(ff.Sr));
    if (f > s0) {
#pragma unroll
      for (int nv = 0; nv < 5; nv++) {
        double rr = -dtdy*(F[nv] - Fp[nv]);
        if (PH == 0) rr += 0.0;
        double du = (double)rx[nv];
        du += rr;
        sto(dUq[nv], cc, (R)du);
      }
      if (want_cdt) {
        double cd = (double)cx;
        cd += 0.5*(cmp + cm)/g.dy;
        sto(Cdt, cc, (R)cd);
      }
    }
#pragma unroll
    for (int nv = 0; nv < 5; nv++) { Fp[nv] = F[nv]; rx[nv] = rxn[nv]; }
    cx = cxn;
    cmp = cm;
#pragma unroll
    for (int q = 0; q < 7; q++) { Vm[q] = Vc[q]; Vc[q] = Vp[q]; }
  }
  if (e) atomicOr(err, e);
}

// ---------------------------------------------------------------------------
// Fused y / z sweep: each thread marches a segment of L cells along D with a
// register window (V(f-1), V(f), left state of face f), computes faces
// s0..s0+L (the last one is the first face of the next segment) and
// accumulates the divergence of cell f-1 as soon as face f is known.  The
// accumulator (and for z the conserved state) of cell f-1 is loaded at the
// top of the step, before the Riemann solve, so its latency hides behind it.
// y: dUc += rhs_y, C_dt += ... (update_stage.cpp:272-284, rhs.cpp:671,756);
// z: dUc += rhs_z (:369-380, rhs.cpp:900,981), then the RK combination of U
// (rk_step.cpp:366-374,509-518,670-708) and the per-block max of C_dt
// (getMax, update_stage.cpp:446-458).
// ---------------------------------------------------------------------------
template <typename R, int D, int LIM, int SOLVER, int L, int PH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 && D == 1 ? 4 : (PH == 1 && SOLVER == 1 ? 2 : 1))))
void k_sweep_march(Box3 g, const R *__restrict__ V, const double *__restrict__ Vs,
                   R *__restrict__ FD, R *__restrict__ dU, R *__restrict__ Cdt,
                   const double *Uin, double *U, const double *U0,
                   const StepCtl *__restrict__ ctl, RKW w, int first_stage, int want_cdt,
                   double *__restrict__ cdt_part, R gamma, uint32_t *__restrict__ err)
{
  static_assert(D == 1 || D == 2, "march along y or z");
  constexpr int BXn = 4 + D;
  constexpr int VAR[7] = {0, 1, 2, 3, D == 1 ? 4 : 4, D == 1 ? 6 : 5, 7};
  // block order: x fastest, then the other transverse index, then the
  // segment, so that concurrent blocks stream whole planes
  // planes kb .. ke-1: for y the transverse index a, for z whole segments
  const int t = blockIdx.x*blockDim.x + threadIdx.x;
  const int nd = D == 1 ? g.ny : g.nz;          // cells along the sweep
  const int na = D == 1 ? g.ke - g.kb : g.ny;   // the other transverse extent
  const int nseg = D == 1 ? (nd + L - 1)/L : (g.ke - g.kb + L - 1)/L;
  double cdt_max = -1.0;
  uint32_t e = 0;
  if (t < g.nx*na*nseg) {
    const size_t n = (size_t)g.ncell;
    // uniform array bases (scalar registers) + 32-bit cell offsets
    const R *Vq[7];
#pragma unroll
    for (int q = 0; q < 7; q++) Vq[q] = V + VAR[q]*n;
    const double *Bn = Vs + D*n;
    R *dUq[5];
    double *Uq[5];
    const double *Uiq[5], *U0q[5];
#pragma unroll
    for (int nv = 0; nv < 5; nv++) {
      dUq[nv] = dU + nv*n; Uq[nv] = U + nv*n; Uiq[nv] = Uin + nv*n; U0q[nv] = U0 + nv*n;
    }
    R *fd0 = FD + (size_t)D*uct_nrec<SOLVER>()*n;
    const int i = t % g.nx;
    const int r = t / g.nx;
    const int a = (D == 1 ? g.kb : 0) + r % na, seg = (D == 1 ? 0 : g.kb/L) + r / na;
    const int sd = D == 1 ? g.nx : g.nx*g.ny;     // stride along D
    const int base = D == 1 ? i + g.nx*g.ny*a : i + g.nx*a;
    const int s0 = seg*L, s1 = s0 + L < nd ? s0 + L : nd;
    const double dtdx = ctl->dt/(D == 1 ? g.dy : g.dz);
    const double ddx = D == 1 ? g.dy : g.dz;
    R Vm[7], Vc[7], Vp[7], vlc[7];
    {
      const unsigned cm2 = base + wrapm(s0 - 2, nd)*sd, cm1 = base + wrapm(s0 - 1, nd)*sd;
      const unsigned c0 = base + wrapm(s0, nd)*sd;
#pragma unroll
      for (int q = 0; q < 7; q++) { Vm[q] = ldo(Vq[q], cm2); Vc[q] = ldo(Vq[q], cm1); Vp[q] = ldo(Vq[q], c0); }
      R lo[7];
#pragma unroll
      for (int q = 0; q < 7; q++) {
        const R dv = plm_slope<R, LIM>(Vm[q], Vc[q], Vp[q]);
        lo[q] = Vc[q] - R(0.5)*dv;
        vlc[q] = Vc[q] + R(0.5)*dv;
      }
      if (PH == 1 && superluminal(lo, vlc)) {
#pragma unroll
        for (int q = 0; q < 7; q++) vlc[q] = Vc[q];
      }
#pragma unroll
      for (int q = 0; q < 7; q++) { Vm[q] = Vc[q]; Vc[q] = Vp[q]; }
    }
    double Fp[5] = {0, 0, 0, 0, 0}, cmp = 0.0;
    for (int f = s0; f <= s1; f++) {
      const unsigned cf = base + wrapm(f, nd)*sd;
      const unsigned cn = base + wrapm(f + 1, nd)*sd;
      const unsigned cc = base + wrapm(f - 1, nd)*sd;    // the cell completed by face f
      // early loads for cell f-1: accumulator, C_dt, and for z the state
      double acc[5], u_old[5], u0_old[5], cacc = 0.0;
      if (f > s0) {
#pragma unroll
        for (int nv = 0; nv < 5; nv++) acc[nv] = (double)ldo((const R *)dUq[nv], cc);
        if (want_cdt) cacc = (double)ldo((const R *)Cdt, cc);
        if (D == 2) {
#pragma unroll
          for (int nv = 0; nv < 5; nv++) u_old[nv] = ldo(Uiq[nv], cc);
          if (!first_stage) {
#pragma unroll
            for (int nv = 0; nv < 5; nv++) u0_old[nv] = ldo(U0q[nv], cc);
          }
        }
      }
      R vl[8], vr[8];
#pragma unroll
      for (int q = 0; q < 7; q++) Vp[q] = ldo(Vq[q], cn);
#pragma unroll
      for (int q = 0; q < 7; q++) {
        const R dv = plm_slope<R, LIM>(Vm[q], Vc[q], Vp[q]);
        vl[VAR[q]] = vlc[q];
        vr[VAR[q]] = Vc[q] - R(0.5)*dv;
        vlc[q] = Vc[q] + R(0.5)*dv;
      }
      if (PH == 1) {
        const R lo[4] = {R(0), vr[1], vr[2], vr[3]};
        if (superluminal(lo, vlc)) {
#pragma unroll
          for (int q = 0; q < 7; q++) { vr[VAR[q]] = Vc[q]; vlc[q] = Vc[q]; }
        }
      }
      vl[BXn] = vr[BXn] = (R)ldo(Bn, cf);
      FaceFlux<R> ff;
      riemann<R, D, SOLVER, PH>(vl, vr, gamma, ff, g);
      {
        const int ff_ = wrapm(f, nd);
        if (D == 1 ? face_reported<1>(g, i, ff_, a) : face_reported<2>(g, i, a, ff_)) e |= ff.err;
      }
      if (f < s1) uct_store<R, D, SOLVER>(vl, vr, gamma, ff, fd0, cf, n);
      const double F[5] = {(double)ff.f[0], (double)ff.f[1], (double)ff.f[2], (double)ff.f[3],
                           (double)ff.f[7]};
      const double cm = (double)rmax(rabs(ff.Sl), rabs(ff.Sr));
      if (f > s0) {
#pragma unroll
        for (int nv = 0; nv < 5; nv++) {
          double rr = -dtdx*(F[nv] - Fp[nv]);
          if (PH == 0) rr += 0.0;
          double du = acc[nv];
          du += rr;
          if (D == 1) {
            sto(dUq[nv], cc, (R)du);
          } else if (first_stage) {
            sto(Uq[nv], cc, u_old[nv] + w.wR*du);
          } else {
            sto(Uq[nv], cc, w.w0*u0_old[nv] + w.w1*u_old[nv] + w.wR*du);
          }
        }
        if (want_cdt) {
          double cd = cacc;
          cd += 0.5*(cmp + cm)/ddx;
          if (D == 1) sto(Cdt, cc, (R)cd);
          else if (is_active(g, i, a, wrapm(f - 1, nd))) cdt_max = cdt_max >= cd ? cdt_max : cd;
        }
      }
#pragma unroll
      for (int nv = 0; nv < 5; nv++) Fp[nv] = F[nv];
      cmp = cm;
#pragma unroll
      for (int q = 0; q < 7; q++) { Vm[q] = Vc[q]; Vc[q] = Vp[q]; }
    }
  }
  if (e) atomicOr(err, e);
  if (D == 2 && want_cdt) {
    const double m = block_max(cdt_max);
    if (threadIdx.x == 0) cdt_part[g.part0 + blockIdx.x] = m;
  }
}

// PLM of a scalar column for the EMF average (array_reconstruct.cpp:167-205):
// left state at the interface between q1 and q2 from (q0,q1,q2), right state
// from (q1,q2,q3).
template <typename R, int LIM>
__device__ __forceinline__ void rec4(R q0, R q1, R q2, R q3, R &L, R &Rr)
{
  L  = q1 + R(0.5)*plm_slope<R, LIM>(q0, q1, q2);
  Rr = q2 - R(0.5)*plm_slope<R, LIM>(q1, q2, q3);
}

// one UCT contribution (ct_emf_average.cpp:636-643):
// (aL vL bL + aR vR bR) +/- (d bR - d bL), from preloaded values: face-field stencil q[0..3] and
// transverse-velocity stencil e[0..3] along the reconstruction direction, and
// the (aL, aR, d) or (aL, aR, dL, dR) weights of the two faces the edge
// averages over (fa is the displaced face, fb the face of the edge's own cell).
template <typename R, int LIMV, int NQ = 3>
__device__ __forceinline__ R uct_val(const R *q, const R *e, const R *fa, const R *fb, int sgn)
{
  R bL, bR, vL, vR;
  rec4<R, MC_LIM>(q[0], q[1], q[2], q[3], bL, bR);   // recB = MC_LIM
  rec4<R, LIMV>(e[0], e[1], e[2], e[3], vL, vR);     // recV = LIMITER
  const R aL = R(0.5)*(fa[0] + fb[0]);
  const R aR = R(0.5)*(fa[1] + fb[1]);
  const R dL = R(0.5)*(fa[2] + fb[2]);
  // UCT_HLLD (NQ = 4 records aL, aR, dL, dR) keeps dR apart; UCT_HLL has dR = dL
  R dR = dL;
  if constexpr (NQ == 4) dR = R(0.5)*(fa[3] + fb[3]);
  const R phi = dR*bR - dL*bL;
  const R s = aL*vL*bL + aR*vR*bR;
  return sgn > 0 ? s + phi : s - phi;
}

// ---------------------------------------------------------------------------
// Edge EMFs, CT_EMF_Riemann2D for UCT_HLL (ct_emf_average.cpp:535-989):
// Ez = X1a + X2b, Ey = X1b + X3a, Ex = X2a + X3b, same order as the reference.
// A block is a 64 x 4 (x, y) tile that marches L edges along z: the
// z-stencils (X3a, X3b) and the face records of plane k-1 (X1b, X2a) stay in
// registers, so every HBM byte along z is read once; x- and y-stencils come
// from the plane being processed, shared through L1/L2 by the tile and its
// neighbours on the same XCD.
// ---------------------------------------------------------------------------
template <typename R, int LIMV, int L>
__global__ __launch_bounds__(256)
void k_emf_march(Box3 g, const double *__restrict__ Vs, const R *__restrict__ FD, R *__restrict__ E)
{
  const int ntx = (g.nx + 63) >> 6, nty = (g.ny + 3) >> 2;
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int tx = t % ntx, rr = t / ntx, ty = rr % nty, seg = rr / nty;
  const int i = tx*64 + (threadIdx.x & 63), j = ty*4 + (threadIdx.x >> 6);
  if (i >= g.nx || j >= g.ny) return;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const size_t n = (size_t)g.ncell;
  const int pl = nx*ny;
  const int k0 = seg*L, k1 = k0 + L < nz ? k0 + L : nz;
  const double *Bx = Vs, *By = Vs + n, *Bz = Vs + 2*n;
  const R *FDx = FD, *FDy = FD + 4*n, *FDz = FD + 8*n;   // UCT_HLL: 4 records per face
  int xo[4], yo[4];               // in-plane offsets of the x- and y-stencils (m2, m1, 0, p1)
#pragma unroll
  for (int o = 0; o < 4; o++) {
    xo[o] = wrapm(i + o - 2, nx) + nx*j;
    yo[o] = i + nx*wrapm(j + o - 2, ny);
  }
  const int col = i + nx*j;
  const int colim = xo[1], coljm = yo[1];
  R zbx[4], zex[4], zby[4], zey[4];   // z-stencils of X3a (Bx, e[J][I]) and X3b (By, e[I][J])
  R cxm[3], cym[3];                   // x- and y-face records of plane k-1
#pragma unroll
  for (int o = 0; o < 3; o++) {
    const int c = col + wrapm(k0 + o - 2, nz)*pl;
    zbx[o] = (R)Bx[c]; zex[o] = FDx[1*n + c];
    zby[o] = (R)By[c]; zey[o] = FDy[0*n + c];
  }
  {
    const int c = col + wrapm(k0 - 1, nz)*pl;
    uct_weights<R, 3>(FDx, c, n, cxm);
    uct_weights<R, 3>(FDy, c, n, cym);
  }
  for (int k = k0; k < k1; k++) {
    const int base = k*pl;
    {
      const int c = col + wrapm(k + 1, nz)*pl;
      zbx[3] = (R)Bx[c]; zex[3] = FDx[1*n + c];
      zby[3] = (R)By[c]; zey[3] = FDy[0*n + c];
    }
    const int c = col + base;
    R cx[3], cy[3], cz[3], cxj[3], cyi[3], czi[3], czj[3];
    uct_weights<R, 3>(FDx, c, n, cx);  uct_weights<R, 3>(FDy, c, n, cy);  uct_weights<R, 3>(FDz, c, n, cz);
    uct_weights<R, 3>(FDx, coljm + base, n, cxj);
    uct_weights<R, 3>(FDy, colim + base, n, cyi);
    uct_weights<R, 3>(FDz, colim + base, n, czi);
    uct_weights<R, 3>(FDz, coljm + base, n, czj);
    R q4[4], e4[4];
    // X1a: By along x, e[K][J] (FDy slot 1), x-face records at (j-1, j)
#pragma unroll
    for (int o = 0; o < 4; o++) { q4[o] = (R)By[xo[o] + base]; e4[o] = FDy[1*n + xo[o] + base]; }
    R ez = uct_val<R, LIMV>(q4, e4, cxj, cx, +1);
    // X2b: Bx along y, e[K][I] (FDx slot 0), y-face records at (i-1, i)
#pragma unroll
    for (int o = 0; o < 4; o++) { q4[o] = (R)Bx[yo[o] + base]; e4[o] = FDx[0*n + yo[o] + base]; }
    ez += uct_val<R, LIMV>(q4, e4, cyi, cy, -1);
    // X1b: Bz along x, e[J][K] (FDz slot 0), x-face records at (k-1, k)
#pragma unroll
    for (int o = 0; o < 4; o++) { q4[o] = (R)Bz[xo[o] + base]; e4[o] = FDz[0*n + xo[o] + base]; }
    R ey = uct_val<R, LIMV>(q4, e4, cxm, cx, -1);
    // X3a: Bx along z, e[J][I] (FDx slot 1), z-face records at (i-1, i)
    ey += uct_val<R, LIMV>(zbx, zex, czi, cz, +1);
    // X2a: Bz along y, e[I][K] (FDz slot 1), y-face records at (k-1, k)
#pragma unroll
    for (int o = 0; o < 4; o++) { q4[o] = (R)Bz[yo[o] + base]; e4[o] = FDz[1*n + yo[o] + base]; }
    R ex = uct_val<R, LIMV>(q4, e4, cym, cy, +1);
    // X3b: By along z, e[I][J] (FDy slot 0), z-face records at (j-1, j)
    ex += uct_val<R, LIMV>(zby, zey, czj, cz, -1);
    E[c] = ex;
    E[n + c] = ey;
    E[2*n + c] = ez;
#pragma unroll
    for (int o = 0; o < 3; o++) { zbx[o] = zbx[o + 1]; zex[o] = zex[o + 1]; zby[o] = zby[o + 1]; zey[o] = zey[o + 1]; }
#pragma unroll
    for (int q = 0; q < 3; q++) { cxm[q] = cx[q]; cym[q] = cy[q]; }
  }
}

// The same EMFs with each plane of a 64 x 4 tile staged through LDS.  Wave
// w owns row y0 + w: per plane every thread loads its own cell of the seven
// in-plane stencil arrays and the nine face records once; the halo (rows
// y0-2, y0-1, y0+4 with their face records, columns x0-2, x0-1, x0+64 and
// the (x0-1, y0-1) corner) is one task per lane.  The loads of plane k+1
// are issued before plane k is computed, so their latency hides behind it.
// The z-stencils and the face records of plane k-1 stay in registers.
template <typename R, int LIMV, int L>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 8 ? 3 : 1)))
void k_emf_tile(Box3 g, const double *__restrict__ Vs, const R *__restrict__ FD, R *__restrict__ E)
{
  constexpr int TX = 64, TY = 4;
  __shared__ R xs[4][TY][TX + 3];        // By, e[K][J], Bz, e[J][K] at x0-2 .. x0+64
  __shared__ R ys[4][TY + 3][TX];        // Bx, e[K][I], Bz, e[I][K] at y0-2 .. y0+4
  __shared__ R cxs[3][TY + 1][TX];       // x-face records at y0-1 ..
  __shared__ R cys[3][TY][TX + 1];       // y-face records at x0-1 ..
  __shared__ R czs[3][TY + 1][TX + 1];   // z-face records at (x0-1 .., y0-1 ..)
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int ntx = (nx + TX - 1)/TX, nty = (ny + TY - 1)/TY;
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int txb = t % ntx, rr = t / ntx, tyb = rr % nty, seg = g.kb/L + rr / nty;
  const int x0 = txb*TX, y0 = tyb*TY;
  const int tw = nx - x0 < TX ? nx - x0 : TX, th = ny - y0 < TY ? ny - y0 : TY;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const bool act = tx < tw && ty < th;
  const size_t n = (size_t)g.ncell;
  const unsigned pl = nx*ny;
  const int k0 = seg*L, k1 = k0 + L < g.ke ? k0 + L : g.ke;
  // uniform array bases (scalar registers) + 32-bit cell offsets
  const double *Bx = Vs, *By = Vs + n, *Bz = Vs + 2*n;
  // UCT_HLL records per face: et, eb, aL_1, aR_1 (the weights are computed when
  // a face is staged)
  const R *ex0 = FD, *ex1 = FD + n, *fx2 = FD + 2*n, *fx3 = FD + 3*n;
  const R *ey0 = FD + 4*n, *ey1 = FD + 5*n, *fy2 = FD + 6*n, *fy3 = FD + 7*n;
  const R *ez0 = FD + 8*n, *ez1 = FD + 9*n, *fz2 = FD + 10*n, *fz3 = FD + 11*n;
  R *Ex = E, *Ey = E + n, *Ez = E + 2*n;
  const unsigned col = (x0 + (act ? tx : 0)) + nx*(y0 + (act ? ty : 0));
  // halo task of this lane: 1 = y-halo row (waves 0-2: y0-2, y0-1, y0+th),
  // 2 = x-halo cell (wave 3, lane 3r+s: row y0+r, column x0-2, x0-1, x0+tw),
  // 3 = the (x0-1, y0-1) corner (wave 3, lane 12)
  int task = 0, hr = 0, hc = 0;
  bool hf = false;                 // the task also carries face records
  unsigned ho = 0;
  if (ty < 3) {
    if (tx < tw) {
      task = 1; hr = ty == 2 ? th + 2 : ty; hc = tx; hf = ty == 1;
      ho = (x0 + tx) + nx*wrapm(ty == 2 ? y0 + th : y0 - 2 + ty, ny);
    }
  } else if (tx < 3*th) {
    const int r = tx/3, sl = tx - 3*r;
    task = 2; hr = r; hc = sl == 2 ? tw + 2 : sl; hf = sl == 1;
    ho = wrapm(sl == 2 ? x0 + tw : x0 - 2 + sl, nx) + nx*(y0 + r);
  } else if (tx == 12) {
    task = 3;
    ho = wrapm(x0 - 1, nx) + nx*wrapm(y0 - 1, ny);
  }
  R zbx[4], zex[4], zby[4], zey[4];   // z-stencils of X3a (Bx, e[J][I]) and X3b (By, e[I][J])
  R zn[4] = {0, 0, 0, 0};             // their next plane
  R cxm[3], cym[3];                   // x- and y-face weights of plane k-1
  R o[11];                            // own cell of the staged plane
  R ha[4], hq[4];                     // halo stencil values and face records
  auto load_plane = [&](int k) {
    const unsigned b = (unsigned)k*pl;
    if (act) {
      const unsigned c = col + b;
      o[0] = ldo(ey1, c); o[1] = (R)ldo(Bz, c); o[2] = ldo(ez0, c);
      o[3] = ldo(ex0, c); o[4] = ldo(ez1, c);
      o[5] = ldo(fx2, c); o[6] = ldo(fx3, c);
      o[7] = ldo(fy2, c); o[8] = ldo(fy3, c);
      o[9] = ldo(fz2, c); o[10] = ldo(fz3, c);
    }
    const unsigned c = ho + b;
    if (task == 1) {
      ha[0] = (R)ldo(Bx, c); ha[1] = ldo(ex0, c); ha[2] = (R)ldo(Bz, c); ha[3] = ldo(ez1, c);
      if (hf) {
        hq[0] = ldo(fx2, c); hq[1] = ldo(fx3, c);
        hq[2] = ldo(fz2, c); hq[3] = ldo(fz3, c);
      }
    } else if (task == 2) {
      ha[0] = (R)ldo(By, c); ha[1] = ldo(ey1, c); ha[2] = (R)ldo(Bz, c); ha[3] = ldo(ez0, c);
      if (hf) {
        hq[0] = ldo(fy2, c); hq[1] = ldo(fy3, c);
        hq[2] = ldo(fz2, c); hq[3] = ldo(fz3, c);
      }
    } else if (task == 3) {
      hq[2] = ldo(fz2, c); hq[3] = ldo(fz3, c);
    }
  };
  if (act) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const unsigned c = col + wrapm(k0 + q - 2, nz)*pl;
      zbx[q] = (R)ldo(Bx, c); zex[q] = ldo(ex1, c);
      zby[q] = (R)ldo(By, c); zey[q] = ldo(ey0, c);
    }
    const unsigned c = col + wrapm(k0 - 1, nz)*pl;
    uct_hll_coef<R>(ldo(fx2, c), ldo(fx3, c), cxm[0], cxm[1], cxm[2]);
    uct_hll_coef<R>(ldo(fy2, c), ldo(fy3, c), cym[0], cym[1], cym[2]);
  }
  load_plane(k0);
  for (int k = k0; k < k1; k++) {
    // stage plane k
    if (act) {
      xs[0][ty][tx + 2] = zby[2];  xs[1][ty][tx + 2] = o[0];
      xs[2][ty][tx + 2] = o[1];    xs[3][ty][tx + 2] = o[2];
      ys[0][ty + 2][tx] = zbx[2];  ys[1][ty + 2][tx] = o[3];
      ys[2][ty + 2][tx] = o[1];    ys[3][ty + 2][tx] = o[4];
      R w[3];
      uct_hll_coef<R>(o[5], o[6], w[0], w[1], w[2]);
#pragma unroll
      for (int q = 0; q < 3; q++) cxs[q][ty + 1][tx] = w[q];
      uct_hll_coef<R>(o[7], o[8], w[0], w[1], w[2]);
#pragma unroll
      for (int q = 0; q < 3; q++) cys[q][ty][tx + 1] = w[q];
      uct_hll_coef<R>(o[9], o[10], w[0], w[1], w[2]);
#pragma unroll
      for (int q = 0; q < 3; q++) czs[q][ty + 1][tx + 1] = w[q];
    }
    if (task == 1) {
#pragma unroll
      for (int a = 0; a < 4; a++) ys[a][hr][hc] = ha[a];
      if (hf) {
        R wa[3], wz[3];
        uct_hll_coef<R>(hq[0], hq[1], wa[0], wa[1], wa[2]);
        uct_hll_coef<R>(hq[2], hq[3], wz[0], wz[1], wz[2]);
#pragma unroll
        for (int q = 0; q < 3; q++) { cxs[q][0][hc] = wa[q]; czs[q][0][hc + 1] = wz[q]; }
      }
    } else if (task == 2) {
#pragma unroll
      for (int a = 0; a < 4; a++) xs[a][hr][hc] = ha[a];
      if (hf) {
        R wa[3], wz[3];
        uct_hll_coef<R>(hq[0], hq[1], wa[0], wa[1], wa[2]);
        uct_hll_coef<R>(hq[2], hq[3], wz[0], wz[1], wz[2]);
#pragma unroll
        for (int q = 0; q < 3; q++) { cys[q][hr][0] = wa[q]; czs[q][hr + 1][0] = wz[q]; }
      }
    } else if (task == 3) {
      R wz[3];
      uct_hll_coef<R>(hq[2], hq[3], wz[0], wz[1], wz[2]);
#pragma unroll
      for (int q = 0; q < 3; q++) czs[q][0][0] = wz[q];
    }
    __syncthreads();
    // loads of plane k+1 (and the z-stencil's plane k+2) fly during plane k
    if (k + 1 < k1) {
      load_plane(k + 1);
      if (act) {
        const unsigned c = col + wrapm(k + 2, nz)*pl;
        zn[0] = (R)ldo(Bx, c); zn[1] = ldo(ex1, c);
        zn[2] = (R)ldo(By, c); zn[3] = ldo(ey0, c);
      }
    }
    if (act) {
      R cx[3], cy[3], cz[3], cxj[3], cyi[3], czi[3], czj[3];
#pragma unroll
      for (int q = 0; q < 3; q++) {
        cx[q] = cxs[q][ty + 1][tx];  cxj[q] = cxs[q][ty][tx];
        cy[q] = cys[q][ty][tx + 1];  cyi[q] = cys[q][ty][tx];
        cz[q] = czs[q][ty + 1][tx + 1];
        czi[q] = czs[q][ty + 1][tx];
        czj[q] = czs[q][ty][tx + 1];
      }
      R q4[4], e4[4];
#pragma unroll
      for (int a = 0; a < 4; a++) { q4[a] = xs[0][ty][tx + a]; e4[a] = xs[1][ty][tx + a]; }
      R ez = uct_val<R, LIMV>(q4, e4, cxj, cx, +1);                      // X1a
#pragma unroll
      for (int a = 0; a < 4; a++) { q4[a] = ys[0][ty + a][tx]; e4[a] = ys[1][ty + a][tx]; }
      ez += uct_val<R, LIMV>(q4, e4, cyi, cy, -1);                       // X2b
#pragma unroll
      for (int a = 0; a < 4; a++) { q4[a] = xs[2][ty][tx + a]; e4[a] = xs[3][ty][tx + a]; }
      R ey = uct_val<R, LIMV>(q4, e4, cxm, cx, -1);                      // X1b
      ey += uct_val<R, LIMV>(zbx, zex, czi, cz, +1);                     // X3a
#pragma unroll
      for (int a = 0; a < 4; a++) { q4[a] = ys[2][ty + a][tx]; e4[a] = ys[3][ty + a][tx]; }
      R ex = uct_val<R, LIMV>(q4, e4, cym, cy, +1);                      // X2a
      ex += uct_val<R, LIMV>(zby, zey, czj, cz, -1);                     // X3b
      const unsigned c = col + (unsigned)k*pl;
      sto(Ex, c, ex);
      sto(Ey, c, ey);
      sto(Ez, c, ez);
#pragma unroll
      for (int a = 0; a < 3; a++) { zbx[a] = zbx[a + 1]; zex[a] = zex[a + 1]; zby[a] = zby[a + 1]; zey[a] = zey[a + 1]; }
      zbx[3] = zn[0]; zex[3] = zn[1]; zby[3] = zn[2]; zey[3] = zn[3];
#pragma unroll
      for (int q = 0; q < 3; q++) { cxm[q] = cx[q]; cym[q] = cy[q]; }
    }
    __syncthreads();                 // the next plane overwrites the tile
  }
}

// The same edge EMFs with one thread per edge and every stencil value loaded
// from memory (2D thread blocks over x-y; blockIdx.z is the plane).
template <typename R, int LIMV, int NQ = 3>
__global__ __launch_bounds__(256)
void k_emf_edge(Box3 g, const double *__restrict__ Vs, const R *__restrict__ FD, R *__restrict__ E)
{
  constexpr int NR = NQ == 4 ? 6 : 4;   // records per face: et, eb, then the weights' data
  const int i = blockIdx.x*64 + (threadIdx.x & 63), j = blockIdx.y*4 + (threadIdx.x >> 6);
  const int k = blockIdx.z;
  if (i >= g.nx || j >= g.ny) return;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const size_t n = (size_t)g.ncell;
  const unsigned pl = nx*ny, base = k*pl, c = i + nx*j + base;
  // uniform array bases (scalar registers) + 32-bit cell offsets
  const double *Bx = Vs, *By = Vs + n, *Bz = Vs + 2*n;
  const R *ex0 = FD, *ex1 = FD + n, *ey0 = FD + NR*n, *ey1 = FD + (NR + 1)*n;
  const R *ez0 = FD + 2*NR*n, *ez1 = FD + (2*NR + 1)*n;
  const R *fdx = FD, *fdy = FD + NR*n, *fdz = FD + 2*NR*n;
  unsigned xo[4], yo[4], zo[4];
#pragma unroll
  for (int o = 0; o < 4; o++) {
    xo[o] = wrapm(i + o - 2, nx) + nx*j + base;
    yo[o] = i + nx*wrapm(j + o - 2, ny) + base;
    zo[o] = i + nx*j + wrapm(k + o - 2, nz)*pl;
  }
  R cx[NQ], cy[NQ], cz[NQ], cxj[NQ], cyi[NQ], czi[NQ], czj[NQ], cxk[NQ], cyk[NQ];
  uct_weights<R, NQ>(fdx, c, n, cx);  uct_weights<R, NQ>(fdy, c, n, cy);  uct_weights<R, NQ>(fdz, c, n, cz);
  uct_weights<R, NQ>(fdx, yo[1], n, cxj);
  uct_weights<R, NQ>(fdy, xo[1], n, cyi);
  uct_weights<R, NQ>(fdz, xo[1], n, czi);
  uct_weights<R, NQ>(fdz, yo[1], n, czj);
  uct_weights<R, NQ>(fdx, zo[1], n, cxk);
  uct_weights<R, NQ>(fdy, zo[1], n, cyk);
  R q4[4], e4[4];
#pragma unroll
  for (int o = 0; o < 4; o++) { q4[o] = (R)ldo(By, xo[o]); e4[o] = ldo(ey1, xo[o]); }
  R ez = uct_val<R, LIMV, NQ>(q4, e4, cxj, cx, +1);                      // X1a
#pragma unroll
  for (int o = 0; o < 4; o++) { q4[o] = (R)ldo(Bx, yo[o]); e4[o] = ldo(ex0, yo[o]); }
  ez += uct_val<R, LIMV, NQ>(q4, e4, cyi, cy, -1);                       // X2b
#pragma unroll
  for (int o = 0; o < 4; o++) { q4[o] = (R)ldo(Bz, xo[o]); e4[o] = ldo(ez0, xo[o]); }
  R ey = uct_val<R, LIMV, NQ>(q4, e4, cxk, cx, -1);                      // X1b
#pragma unroll
  for (int o = 0; o < 4; o++) { q4[o] = (R)ldo(Bx, zo[o]); e4[o] = ldo(ex1, zo[o]); }
  ey += uct_val<R, LIMV, NQ>(q4, e4, czi, cz, +1);                       // X3a
#pragma unroll
  for (int o = 0; o < 4; o++) { q4[o] = (R)ldo(Bz, yo[o]); e4[o] = ldo(ez1, yo[o]); }
  R ex = uct_val<R, LIMV, NQ>(q4, e4, cyk, cy, +1);                      // X2a
#pragma unroll
  for (int o = 0; o < 4; o++) { q4[o] = (R)ldo(By, zo[o]); e4[o] = ldo(ey0, zo[o]); }
  ex += uct_val<R, LIMV, NQ>(q4, e4, czj, cz, -1);                       // X3b
  sto(E, c, ex);
  sto(E + n, c, ey);
  sto(E + 2*n, c, ez);
}

// Rs = -dt curl E on the three lower faces of cell (i,j,k), Cartesian
// CT_Update (ct_update.cpp:181-258, 266-339, 348-422).
struct FaceR { double bx, by, bz; };

template <typename R>
__device__ __forceinline__ FaceR ct_rs(int i, int j, int k, const Box3 &g, const R *__restrict__ E,
                                       double dt)
{
  const size_t n = (size_t)g.ncell;
  const int ip = wrapm(i + 1, g.nx), jp = wrapm(j + 1, g.ny), kp = wrapm(k + 1, g.nz);
  const R *Ex = E, *Ey = E + n, *Ez = E + 2*n;
  const double ex0 = ldo(Ex, cid(i, j, k, g)), exj = ldo(Ex, cid(i, jp, k, g)), exk = ldo(Ex, cid(i, j, kp, g));
  const double ey0 = ldo(Ey, cid(i, j, k, g)), eyi = ldo(Ey, cid(ip, j, k, g)), eyk = ldo(Ey, cid(i, j, kp, g));
  const double ez0 = ldo(Ez, cid(i, j, k, g)), ezi = ldo(Ez, cid(ip, j, k, g)), ezj = ldo(Ez, cid(i, jp, k, g));
  const double dtdx1 = dt/g.dx, dtdx2 = dt/g.dy, dtdx3 = dt/g.dz;
  FaceR r;
  double curlE = 0.0 + dtdx2*(ezj - ez0);
  curlE = curlE - dtdx3*(eyk - ey0);
  r.bx = -curlE;
  curlE = - dtdx1*(ezi - ez0);
  curlE = curlE + dtdx3*(exk - ex0);
  r.by = -curlE;
  curlE = dtdx1*(eyi - ey0) - dtdx2*(exj - ex0);
  r.bz = -curlE;
  return r;
}

// ---------------------------------------------------------------------------
// Face update + cell-centred field + ConsToPrim.
// Staggered RK combination (rk_step.cpp:378-385,540-552), field average
// (ct_field_average.cpp:182-184) and ConsToPrim3D (mappers3D.cpp:104-194).
// Each thread owns the three lower faces of its cell; the three upper faces
// it needs for the average are recomputed from the same inputs (bitwise the
// value their owner writes).  Vs is double-buffered (Vin -> Vout).
// ---------------------------------------------------------------------------
template <typename R, int PH, bool NET = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PH == 1 && !NET ? 6 : 1)))
void k_finish(Box3 g, const R *__restrict__ E, const double *__restrict__ Vin,
              double *__restrict__ Vout, const double *__restrict__ Bs0,
              double *__restrict__ U, R *__restrict__ V,
              const StepCtl *__restrict__ ctl, RKW w, int first_stage,
              double gmm1, double small_rho, double small_prs, uint32_t *__restrict__ err,
              const NetPack *__restrict__ net, ZoneRep zr)
{
  // no early exit: the learned con2prim guess is evaluated by whole wavefronts
  const int c_beg = g.kb*g.nx*g.ny, c_end = g.ke*g.nx*g.ny;   // planes kb .. ke-1
  const int c0 = c_beg + xcd_runs<8>(blockIdx.x, gridDim.x)*blockDim.x + threadIdx.x;
  const bool live = c0 < c_end;
  const int c = live ? c0 : c_end - 1;
  const size_t n = (size_t)g.ncell;
  const Cell p = cell_of(c, g);
  const double dt = ctl->dt;
  const int ci = shift<0>(p, 1, g), cj = shift<1>(p, 1, g), ck = shift<2>(p, 1, g);
  const FaceR r0 = ct_rs<R>(p.i, p.j, p.k, g, E, dt);
  const FaceR rx = ct_rs<R>(wrapm(p.i + 1, g.nx), p.j, p.k, g, E, dt);
  const FaceR ry = ct_rs<R>(p.i, wrapm(p.j + 1, g.ny), p.k, g, E, dt);
  const FaceR rz = ct_rs<R>(p.i, p.j, wrapm(p.k + 1, g.nz), g, E, dt);
  double b_lo[3], b_hi[3];
  const double rl[3] = {r0.bx, r0.by, r0.bz};
  const double rh[3] = {rx.bx, ry.by, rz.bz};
  const int up[3] = {ci, cj, ck};
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const double *vin = Vin + d*n;
    if (first_stage) {
      b_lo[d] = ldo(vin, c) + w.wR*rl[d];
      b_hi[d] = ldo(vin, up[d]) + w.wR*rh[d];
    } else {
      const double *b0 = Bs0 + d*n;
      b_lo[d] = w.w0*ldo(b0, c) + w.w1*ldo(vin, c) + w.wR*rl[d];
      b_hi[d] = w.w0*ldo(b0, up[d]) + w.w1*ldo(vin, up[d]) + w.wR*rh[d];
    }
    if (live) sto(Vout + d*n, c, b_lo[d]);
  }
  double b[3], u[5], v[8];
#pragma unroll
  for (int d = 0; d < 3; d++) b[d] = 0.5*(b_lo[d] + b_hi[d]);
#pragma unroll
  for (int nv = 0; nv < 5; nv++) u[nv] = ldo((const double *)U + nv*n, c);
  if constexpr (PH == 1) {
    // ConsToPrim RMHD (mappers.cpp:33-157); written back when the mapper reports
    // anything (mappers3D.cpp:182-190).  gmm1 carries gamma here.
    double u8[8] = {u[0], u[1], u[2], u[3], b[0], b[1], b[2], u[4]};
    double scale = 1.0;
    if constexpr (NET) {           // a separate kernel: the drop-in keeps its registers
      float x[NET_F];
      net_features(u8, rmhd_guess_a27(u8), x);
      scale = (double)__expf(net_eval64(net, x));
    }
    uint32_t local = 0;
    const uint32_t e = rmhd_cons_to_prim(u8, v, gmm1, small_rho, small_prs, local, NET, scale);
    if (!live) return;
    if (local) {
      sto(U, c, u8[0]); sto(U + n, c, u8[1]); sto(U + 2*n, c, u8[2]); sto(U + 3*n, c, u8[3]);
      sto(U + 4*n, c, u8[7]);
    }
    if (is_active(g, p.i, p.j, p.k)) {
      if (e) atomicOr(err, e);
      zone_report(zr, c, e);
    }
  } else {
    const uint32_t e = mhd_cons_to_prim(u, b, gmm1, small_rho, small_prs, v);
    if (!live) return;
    if (e) {
#pragma unroll
      for (int nv = 0; nv < 5; nv++) sto(U + nv*n, c, u[nv]);
      if (is_active(g, p.i, p.j, p.k)) atomicOr(err, e);
    }
    if (is_active(g, p.i, p.j, p.k)) zone_report(zr, c, e);
  }
#pragma unroll
  for (int nv = 0; nv < 8; nv++) sto(V + nv*n, c, (R)v[nv]);
}

// ---------------------------------------------------------------------------
// Host-side launch helpers: one entry per kernel family, so that the host
// orchestration can bracket each launch with timing events.
// ---------------------------------------------------------------------------
template <typename R, int LIM, int SOLVER, int L, int PH>
static void launch_t(int which, const Mhd3dBufs<R> &b, const Box3 &g, const StepCtl *ctl, RKW w,
                     int first_stage, int want_dt, double gamma, double small_rho,
                     double small_prs, hipStream_t s)
{
  const int T = 256;
  const int nb = (g.ncell + T - 1)/T;
  const int nk = g.ke - g.kb;          // planes of the fused kernels' launch
  switch (which) {
  case K_SWEEP_X:
    hipLaunchKernelGGL((k_sweep<R, 0, LIM, SOLVER>), dim3(nb), dim3(T), 0, s,
                       g, b.V, b.Vin, b.F, b.CM, b.FD, (R)gamma, b.err);
    break;
  case K_SWEEP_Y:
    hipLaunchKernelGGL((k_sweep<R, 1, LIM, SOLVER>), dim3(nb), dim3(T), 0, s,
                       g, b.V, b.Vin, b.F, b.CM, b.FD, (R)gamma, b.err);
    break;
  case K_SWEEP_Z:
    hipLaunchKernelGGL((k_sweep<R, 2, LIM, SOLVER>), dim3(nb), dim3(T), 0, s,
                       g, b.V, b.Vin, b.F, b.CM, b.FD, (R)gamma, b.err);
    break;
  case K_RHS:
    hipLaunchKernelGGL((k_rhs_update<R>), dim3(nb), dim3(T), 0, s,
                       g, b.F, b.CM, b.Uin, b.U, b.U0, ctl, w, first_stage,
                       want_dt ? b.cdt_part : nullptr);
    break;
  case K_SWEEPX_F: {
    const size_t shm = 14*(size_t)g.nx*sizeof(R);
    hipLaunchKernelGGL((k_sweepx_fused<R, LIM, SOLVER, PH>), dim3(g.ny*nk), dim3(g.nx), shm, s,
                       g, b.V, b.Vin, b.FD, b.dU, b.Cdt, ctl, (R)gamma, want_dt, b.err);
    break;
  }
  case K_SWEEPXY_F: {
    // ideal MHD only: the RMHD x+y kernel would hold more registers than a
    // wavefront has (the engine never selects it for RMHD)
    if constexpr (PH == 0) {
      const size_t shm = 20*(size_t)g.nx*sizeof(R);
      const int nb = g.nz*((g.ny + L - 1)/L);
      hipLaunchKernelGGL((k_sweep_xy<R, LIM, SOLVER, L, PH>), dim3(nb), dim3(g.nx), shm, s,
                         g, b.V, b.Vin, b.FD, b.dU, b.Cdt, ctl, (R)gamma, want_dt, b.err);
    }
    break;
  }
  case K_SWEEPY_F: {
    const int nt = g.nx*nk*((g.ny + L - 1)/L);
    hipLaunchKernelGGL((k_sweep_march<R, 1, LIM, SOLVER, L, PH>), dim3((nt + T - 1)/T), dim3(T), 0, s,
                       g, b.V, b.Vin, b.FD, b.dU, b.Cdt, b.Uin, b.U, b.U0, ctl, w, first_stage, want_dt,
                       b.cdt_part, (R)gamma, b.err);
    break;
  }
  case K_SWEEPZ_F: {
    const int nt = g.nx*g.ny*((nk + L - 1)/L);
    hipLaunchKernelGGL((k_sweep_march<R, 2, LIM, SOLVER, L, PH>), dim3((nt + T - 1)/T), dim3(T), 0, s,
                       g, b.V, b.Vin, b.FD, b.dU, b.Cdt, b.Uin, b.U, b.U0, ctl, w, first_stage, want_dt,
                       b.cdt_part, (R)gamma, b.err);
    break;
  }
  case K_EMF: {
    if (SOLVER == 4) {                   // UCT_HLLD: six records per face, one edge per thread
      hipLaunchKernelGGL((k_emf_edge<R, LIM, 4>), dim3((g.nx + 63)/64, (g.ny + 3)/4, g.nz), dim3(256), 0, s,
                         g, b.Vin, b.FD, b.E);
    } else if (b.emf_march == 0) {
      hipLaunchKernelGGL((k_emf_edge<R, LIM>), dim3((g.nx + 63)/64, (g.ny + 3)/4, g.nz), dim3(256), 0, s,
                         g, b.Vin, b.FD, b.E);
    } else if (b.emf_march == 4) {
      const int nt = ((g.nx + 63)/64)*((g.ny + 3)/4)*((g.nz + 3)/4);
      hipLaunchKernelGGL((k_emf_march<R, LIM, 4>), dim3(nt), dim3(256), 0, s, g, b.Vin, b.FD, b.E);
    } else if (b.emf_march < 0) {        // LDS-staged tiles marching -emf_march planes
      const int lt = -b.emf_march;
      const int nt = ((g.nx + 63)/64)*((g.ny + 3)/4)*((nk + lt - 1)/lt);
      if (lt == 4)
        hipLaunchKernelGGL((k_emf_tile<R, LIM, 4>), dim3(nt), dim3(256), 0, s, g, b.Vin, b.FD, b.E);
      else if (lt == 8)
        hipLaunchKernelGGL((k_emf_tile<R, LIM, 8>), dim3(nt), dim3(256), 0, s, g, b.Vin, b.FD, b.E);
      else if (lt == 16)
        hipLaunchKernelGGL((k_emf_tile<R, LIM, 16>), dim3(nt), dim3(256), 0, s, g, b.Vin, b.FD, b.E);
      else
        hipLaunchKernelGGL((k_emf_tile<R, LIM, 32>), dim3(nt), dim3(256), 0, s, g, b.Vin, b.FD, b.E);
    } else if (b.emf_march == 8) {
      const int nt = ((g.nx + 63)/64)*((g.ny + 3)/4)*((g.nz + 7)/8);
      hipLaunchKernelGGL((k_emf_march<R, LIM, 8>), dim3(nt), dim3(256), 0, s, g, b.Vin, b.FD, b.E);
    } else if (b.emf_march == 32) {
      const int nt = ((g.nx + 63)/64)*((g.ny + 3)/4)*((g.nz + 31)/32);
      hipLaunchKernelGGL((k_emf_march<R, LIM, 32>), dim3(nt), dim3(256), 0, s, g, b.Vin, b.FD, b.E);
    } else {
      const int nt = ((g.nx + 63)/64)*((g.ny + 3)/4)*((g.nz + 15)/16);
      hipLaunchKernelGGL((k_emf_march<R, LIM, 16>), dim3(nt), dim3(256), 0, s, g, b.Vin, b.FD, b.E);
    }
    break;
  }
  case K_FINISH:
    // the RMHD mapper takes gamma, the MHD one gamma - 1
    if (PH == 1 && b.net)
      hipLaunchKernelGGL((k_finish<R, PH, PH == 1>), dim3((nk*g.nx*g.ny + T - 1)/T), dim3(T), 0, s,
                         g, b.E, b.Vin, b.Vout, b.Bs0, b.U, b.V, ctl, w, first_stage, gamma,
                         small_rho, small_prs, b.err, b.net, b.zr);
    else
      hipLaunchKernelGGL((k_finish<R, PH, false>), dim3((nk*g.nx*g.ny + T - 1)/T), dim3(T), 0, s,
                         g, b.E, b.Vin, b.Vout, b.Bs0, b.U, b.V, ctl, w, first_stage,
                         PH == 1 ? gamma : gamma - 1.0, small_rho, small_prs, b.err, nullptr, b.zr);
    break;
  default:
    break;
  }
}

template <typename R, int LIM, int SOLVER, int PH>
static void launch_l(int which, const Mhd3dBufs<R> &b, const Box3 &g, const StepCtl *ctl, RKW w,
                     int first_stage, int want_dt, int march, double gamma, double small_rho,
                     double small_prs, hipStream_t s)
{
  if (march == 32) launch_t<R, LIM, SOLVER, 32, PH>(which, b, g, ctl, w, first_stage, want_dt, gamma, small_rho, small_prs, s);
  else if (march == 8) launch_t<R, LIM, SOLVER, 8, PH>(which, b, g, ctl, w, first_stage, want_dt, gamma, small_rho, small_prs, s);
  else launch_t<R, LIM, SOLVER, 16, PH>(which, b, g, ctl, w, first_stage, want_dt, gamma, small_rho, small_prs, s);
}

template <typename R>
void mhd3d_launch(int which, const Mhd3dBufs<R> &b, const Box3 &g, const StepCtl *ctl, RKW w,
                  int first_stage, int want_dt, int limiter, int solver, int march, double gamma,
                  double small_rho, double small_prs, hipStream_t s)
{
  if constexpr (std::is_same<R, double>::value) {
    if (b.phys == 1) {            // RMHD: HLL, HLLC or HLLD, FP64 only
      if (solver == 3) {
        if (limiter == MC_LIM) launch_l<R, MC_LIM, 3, 1>(which, b, g, ctl, w, first_stage, want_dt, march, gamma, small_rho, small_prs, s);
        else                   launch_l<R, VANLEER_LIM, 3, 1>(which, b, g, ctl, w, first_stage, want_dt, march, gamma, small_rho, small_prs, s);
      } else if (solver == 2) {
        if (limiter == MC_LIM) launch_l<R, MC_LIM, 2, 1>(which, b, g, ctl, w, first_stage, want_dt, march, gamma, small_rho, small_prs, s);
        else                   launch_l<R, VANLEER_LIM, 2, 1>(which, b, g, ctl, w, first_stage, want_dt, march, gamma, small_rho, small_prs, s);
      } else {
        if (limiter == MC_LIM) launch_l<R, MC_LIM, 1, 1>(which, b, g, ctl, w, first_stage, want_dt, march, gamma, small_rho, small_prs, s);
        else                   launch_l<R, VANLEER_LIM, 1, 1>(which, b, g, ctl, w, first_stage, want_dt, march, gamma, small_rho, small_prs, s);
      }
      return;
    }
  }
  if (solver == 4) {               // HLLD with the UCT_HLLD EMF
    if (limiter == MC_LIM) launch_l<R, MC_LIM, 4, 0>(which, b, g, ctl, w, first_stage, want_dt, march, gamma, small_rho, small_prs, s);
    else                   launch_l<R, VANLEER_LIM, 4, 0>(which, b, g, ctl, w, first_stage, want_dt, march, gamma, small_rho, small_prs, s);
    return;
  }
  if (limiter == MC_LIM) {
    if (solver == 3)      launch_l<R, MC_LIM, 3, 0>(which, b, g, ctl, w, first_stage, want_dt, march, gamma, small_rho, small_prs, s);
    else if (solver == 2) launch_l<R, MC_LIM, 2, 0>(which, b, g, ctl, w, first_stage, want_dt, march, gamma, small_rho, small_prs, s);
    else                  launch_l<R, MC_LIM, 1, 0>(which, b, g, ctl, w, first_stage, want_dt, march, gamma, small_rho, small_prs, s);
  } else {
    if (solver == 3)      launch_l<R, VANLEER_LIM, 3, 0>(which, b, g, ctl, w, first_stage, want_dt, march, gamma, small_rho, small_prs, s);
    else if (solver == 2) launch_l<R, VANLEER_LIM, 2, 0>(which, b, g, ctl, w, first_stage, want_dt, march, gamma, small_rho, small_prs, s);
    else                  launch_l<R, VANLEER_LIM, 1, 0>(which, b, g, ctl, w, first_stage, want_dt, march, gamma, small_rho, small_prs, s);
  }
}

}  // namespace gpl
