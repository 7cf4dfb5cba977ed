// C-ABI of the MI355X stage update (include/gpluto_amd.h).
//
// Host orchestration of AdvanceStep (Src/rk_step.cpp:88-720) over the HIP
// kernels of mhd3d.hip: the device owns all state; the time step lives in
// a device-side control block so a run of many steps is replayed from one
// captured HIP graph without host round trips.
#include "../../include/gpluto_amd.h"
#include "mhd3d.hpp"
#include "physics.hpp"

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

using namespace gpl;

namespace {

const char *kKernelNames[K_COUNT] = {"k_sweep_x", "k_sweep_y", "k_sweep_z", "k_rhs_update",
                                     "k_sweepx_fused", "k_sweepy_fused", "k_sweepz_fused",
                                     "k_emf", "k_finish", "k_dt"};

struct Timing {
  bool on = false;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
  long long launches[K_COUNT] = {0};
  double total_ms[K_COUNT] = {0};
  std::vector<hipEvent_t> pool;
  hipEvent_t get()
  {
    if (!pool.empty()) { hipEvent_t e = pool.back(); pool.pop_back(); return e; }
    hipEvent_t e;
    hipEventCreate(&e);
    return e;
  }
};

}  // namespace

struct gpl_ctx {
  gpl_config cfg;
  Box3 g;
  int dev = 0;
  hipStream_t stream = nullptr;
  bool fp64 = true;
  size_t n = 0;
  // device buffers
  void *V = nullptr;          // [8][n] R
  double *U = nullptr, *U0 = nullptr;
  double *VsA = nullptr, *VsB = nullptr, *Bs0 = nullptr;
  void *F = nullptr, *CM = nullptr, *FD = nullptr, *E = nullptr;
  double *dU = nullptr, *Cdt = nullptr;
  bool fused = true;          // fused sweep kernels (nx <= 512) or the split path
  int march = 16;             // cells per thread in the y/z marching sweeps
  double *cdt_part = nullptr;
  int nparts = 0;
  uint32_t *err = nullptr;
  StepCtl *ctl = nullptr;
  double *cur_vs = nullptr;   // VsA or VsB: faces of the current state
  size_t bytes = 0;
  // graph
  hipGraphExec_t graph_exec = nullptr;
  double *graph_vs = nullptr; // face buffer the captured graph starts from
  int graph_steps = 0;
  Timing timing;
  hipEvent_t run_beg = nullptr, run_end = nullptr;
  double last_run_ms = 0.0;
};

#define HIPCHK(x)                                                        \
  do {                                                                   \
    hipError_t e_ = (x);                                                 \
    if (e_ != hipSuccess) {                                              \
      fprintf(stderr, "gpluto_amd: %s failed: %s (%s:%d)\n", #x,         \
              hipGetErrorString(e_), __FILE__, __LINE__);                \
      return GPL_EHIP;                                                   \
    }                                                                    \
  } while (0)

static int dev_alloc(gpl_ctx *c, void **p, size_t bytes)
{
  if (hipMalloc(p, bytes) != hipSuccess) { *p = nullptr; return GPL_ENOMEM; }
  c->bytes += bytes;
  return 0;
}

extern "C" void gpl_default_config(gpl_config *cfg)
{
  memset(cfg, 0, sizeof(*cfg));
  cfg->physics = GPL_MHD;
  cfg->dimensions = 3;
  cfg->nx = cfg->ny = cfg->nz = 64;
  for (int d = 0; d < 3; d++) { cfg->xbeg[d] = 0.0; cfg->xend[d] = 1.0; }
  cfg->boundary = GPL_PERIODIC;
  cfg->time_stepping = GPL_RK2;
  cfg->limiter = GPL_MC_LIM;
  cfg->solver = GPL_SOLVER_HLLD;
  cfg->ct_emf_average = GPL_UCT_HLL;
  cfg->precision = GPL_FP64;
  cfg->nghost = 2;
  cfg->device = 0;
  cfg->gamma = 5.0/3.0;
  cfg->cfl = 0.4;
  cfg->cfl_max_var = 1.1;
  cfg->small_rho = 1.e-12;
  cfg->small_prs = 1.e-12;
}

static void free_all(gpl_ctx *c)
{
  void *ptrs[] = {c->V, c->U, c->U0, c->VsA, c->VsB, c->Bs0, c->F, c->CM, c->FD, c->E,
                  c->dU, c->Cdt, c->cdt_part, c->err, c->ctl};
  for (void *p : ptrs) if (p) hipFree(p);
  if (c->graph_exec) hipGraphExecDestroy(c->graph_exec);
  for (auto &pe : c->timing.pending) { hipEventDestroy(pe.second.first); hipEventDestroy(pe.second.second); }
  for (auto e : c->timing.pool) hipEventDestroy(e);
  if (c->run_beg) hipEventDestroy(c->run_beg);
  if (c->run_end) hipEventDestroy(c->run_end);
  if (c->stream) hipStreamDestroy(c->stream);
}

extern "C" int gpl_create(const gpl_config *cfg, gpl_ctx **out)
{
  if (!cfg || !out) return GPL_EINVAL;
  *out = nullptr;
  if (cfg->physics != GPL_MHD || cfg->dimensions != 3) return GPL_EUNSUPPORTED;
  if (cfg->boundary != GPL_PERIODIC || cfg->ct_emf_average != GPL_UCT_HLL) return GPL_EUNSUPPORTED;
  if (cfg->time_stepping != GPL_RK2 && cfg->time_stepping != GPL_RK3) return GPL_EINVAL;
  if (cfg->limiter != GPL_MC_LIM && cfg->limiter != GPL_VANLEER_LIM) return GPL_EINVAL;
  if (cfg->solver != GPL_SOLVER_HLLD && cfg->solver != GPL_SOLVER_HLL) return GPL_EUNSUPPORTED;
  if (cfg->precision != GPL_FP64 && cfg->precision != GPL_MIXED) return GPL_EINVAL;
  // the PLM stencil wraps at most two cells: every direction needs >= 3 zones
  if (cfg->nx < 3 || cfg->ny < 3 || cfg->nz < 3) return GPL_EINVAL;
  const long long ncell = (long long)cfg->nx*cfg->ny*cfg->nz;
  if (ncell > (1LL << 30)) return GPL_EINVAL;   // 32-bit cell index, 5*n fits size_t math
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || cfg->device < 0 || cfg->device >= ndev)
    return GPL_EHIP;

  gpl_ctx *c = new (std::nothrow) gpl_ctx();
  if (!c) return GPL_ENOMEM;
  c->cfg = *cfg;
  c->dev = cfg->device;
  c->fp64 = cfg->precision == GPL_FP64;
  c->n = (size_t)ncell;
  c->g.nx = cfg->nx; c->g.ny = cfg->ny; c->g.nz = cfg->nz;
  c->g.ncell = (int)ncell;
  // dx = (xend - xbeg)/N for every cell (Src/set_grid.cpp:196)
  c->g.dx = (cfg->xend[0] - cfg->xbeg[0])/cfg->nx;
  c->g.dy = (cfg->xend[1] - cfg->xbeg[1])/cfg->ny;
  c->g.dz = (cfg->xend[2] - cfg->xbeg[2])/cfg->nz;
  if (hipSetDevice(c->dev) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return GPL_EHIP;
  }
  const size_t n = c->n, rs = c->fp64 ? sizeof(double) : sizeof(float);
  // the fused x sweep holds one periodic x-row per block
  c->fused = cfg->nx <= 512 && getenv("GPL_SPLIT_KERNELS") == nullptr;
  // GPL_MARCH = 8/16/32: cells per thread in the y/z marching sweeps
  if (const char *m = getenv("GPL_MARCH")) c->march = atoi(m);
  if (c->march != 8 && c->march != 16 && c->march != 32) c->march = 16;
  c->nparts = c->fused ? mhd3d_march_blocks(c->g, c->march) : mhd3d_blocks(c->g);
  int rc = 0;
  rc |= dev_alloc(c, &c->V, 8*n*rs);
  rc |= dev_alloc(c, (void **)&c->U, 5*n*sizeof(double));
  rc |= dev_alloc(c, (void **)&c->U0, 5*n*sizeof(double));
  rc |= dev_alloc(c, (void **)&c->VsA, 3*n*sizeof(double));
  rc |= dev_alloc(c, (void **)&c->VsB, 3*n*sizeof(double));
  rc |= dev_alloc(c, (void **)&c->Bs0, 3*n*sizeof(double));
  if (c->fused) {
    rc |= dev_alloc(c, (void **)&c->dU, 5*n*sizeof(double));
    rc |= dev_alloc(c, (void **)&c->Cdt, n*sizeof(double));
  } else {
    rc |= dev_alloc(c, &c->F, 15*n*rs);
    rc |= dev_alloc(c, &c->CM, 3*n*rs);
  }
  rc |= dev_alloc(c, &c->FD, 15*n*rs);
  rc |= dev_alloc(c, &c->E, 3*n*rs);
  rc |= dev_alloc(c, (void **)&c->cdt_part, (size_t)c->nparts*sizeof(double));
  rc |= dev_alloc(c, (void **)&c->err, sizeof(uint32_t));
  rc |= dev_alloc(c, (void **)&c->ctl, sizeof(StepCtl));
  if (rc) {
    free_all(c);
    delete c;
    return GPL_ENOMEM;
  }
  c->cur_vs = c->VsA;
  hipEventCreate(&c->run_beg);
  hipEventCreate(&c->run_end);
  StepCtl z = {0.0, 0.0, 0.0, 0};
  if (hipMemcpy(c->ctl, &z, sizeof(z), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(c->err, 0, sizeof(uint32_t)) != hipSuccess) {
    free_all(c);
    delete c;
    return GPL_EHIP;
  }
  *out = c;
  return 0;
}

extern "C" int gpl_destroy(gpl_ctx *c)
{
  if (!c) return GPL_EINVAL;
  hipSetDevice(c->dev);
  hipStreamSynchronize(c->stream);
  free_all(c);
  delete c;
  return 0;
}

// ---------------------------------------------------------------------------
// layout conversion
// ---------------------------------------------------------------------------
namespace {

struct Layout {
  int nx, ny, nz, ng;
  int tx, ty, tz;          // totals of the ARY4D layout
  bool ary;
  size_t cell(int nv, int i, int j, int k) const
  {
    if (!ary) return (size_t)nv*nx*ny*nz + ((size_t)k*ny + j)*nx + i;
    return (size_t)nv*tx*ty*tz + ((size_t)(k + ng)*ty + (j + ng))*tx + (i + ng);
  }
  // staggered component d: one extra face along d (Src/initialize.cpp:424-431)
  size_t face(int d, int i, int j, int k) const
  {
    if (!ary) return ((size_t)k*ny + j)*nx + i;
    const int sx = tx + (d == 0), sy = ty + (d == 1);
    return ((size_t)(k + ng)*sy + (j + ng))*sx + (i + ng);
  }
};

Layout make_layout(const gpl_ctx *c, int layout)
{
  Layout L;
  L.nx = c->cfg.nx; L.ny = c->cfg.ny; L.nz = c->cfg.nz;
  L.ng = c->cfg.nghost;
  L.ary = layout == GPL_LAYOUT_ARY4D;
  L.tx = L.nx + 2*L.ng; L.ty = L.ny + 2*L.ng; L.tz = L.nz + 2*L.ng;
  return L;
}

template <typename T>
int to_device(gpl_ctx *c, void *dst, const std::vector<double> &src)
{
  if (sizeof(T) == sizeof(double)) {
    if (hipMemcpy(dst, src.data(), src.size()*sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return GPL_EHIP;
    return 0;
  }
  std::vector<T> tmp(src.size());
  for (size_t q = 0; q < src.size(); q++) tmp[q] = (T)src[q];
  if (hipMemcpy(dst, tmp.data(), tmp.size()*sizeof(T), hipMemcpyHostToDevice) != hipSuccess) return GPL_EHIP;
  return 0;
}

}  // namespace

extern "C" int gpl_upload(gpl_ctx *c, const double *vc, const double *const *vs, int layout)
{
  if (!c || !vc || !vs || !vs[0] || !vs[1] || !vs[2]) return GPL_EINVAL;
  if (layout != GPL_LAYOUT_ACTIVE && layout != GPL_LAYOUT_ARY4D) return GPL_EINVAL;
  hipSetDevice(c->dev);
  if (hipStreamSynchronize(c->stream) != hipSuccess) return GPL_EHIP;
  const Layout L = make_layout(c, layout);
  const int nx = L.nx, ny = L.ny, nz = L.nz;
  const size_t n = c->n;
  std::vector<double> hv(8*n), hu(5*n), hs(3*n);
  for (int d = 0; d < 3; d++)
    for (int k = 0; k < nz; k++)
      for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++)
          hs[d*n + ((size_t)k*ny + j)*nx + i] = vs[d][L.face(d, i, j, k)];
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        const size_t o = ((size_t)k*ny + j)*nx + i;
        double v[8], u[8];
        for (int nv = 0; nv < 8; nv++) v[nv] = vc[L.cell(nv, i, j, k)];
        // cell-centred field = face average (Src/startup.cpp:388 -> ct_field_average.cpp:182-184)
        const int ip = i + 1 == nx ? 0 : i + 1, jp = j + 1 == ny ? 0 : j + 1, kp = k + 1 == nz ? 0 : k + 1;
        v[4] = 0.5*(hs[0*n + o] + hs[0*n + ((size_t)k*ny + j)*nx + ip]);
        v[5] = 0.5*(hs[1*n + o] + hs[1*n + ((size_t)k*ny + jp)*nx + i]);
        v[6] = 0.5*(hs[2*n + o] + hs[2*n + ((size_t)kp*ny + j)*nx + i]);
        // Uc = PrimToCons(Vc) (Src/startup.cpp:514-561)
        mhd_prim_to_cons<double>(v, u, c->cfg.gamma - 1.0);
        for (int nv = 0; nv < 8; nv++) hv[nv*n + o] = v[nv];
        hu[0*n + o] = u[0]; hu[1*n + o] = u[1]; hu[2*n + o] = u[2]; hu[3*n + o] = u[3];
        hu[4*n + o] = u[7];
      }
  int rc = c->fp64 ? to_device<double>(c, c->V, hv) : to_device<float>(c, c->V, hv);
  if (rc) return rc;
  HIPCHK(hipMemcpy(c->U, hu.data(), 5*n*sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->VsA, hs.data(), 3*n*sizeof(double), hipMemcpyHostToDevice));
  c->cur_vs = c->VsA;
  StepCtl z = {0.0, 0.0, 0.0, 0};
  HIPCHK(hipMemcpy(c->ctl, &z, sizeof(z), hipMemcpyHostToDevice));
  return 0;
}

extern "C" int gpl_download(gpl_ctx *c, double *vc, double *const *vs, double *uc, int layout)
{
  if (!c) return GPL_EINVAL;
  if (layout != GPL_LAYOUT_ACTIVE && layout != GPL_LAYOUT_ARY4D) return GPL_EINVAL;
  hipSetDevice(c->dev);
  HIPCHK(hipStreamSynchronize(c->stream));
  const Layout L = make_layout(c, layout);
  const int nx = L.nx, ny = L.ny, nz = L.nz;
  const size_t n = c->n;
  if (vc) {
    std::vector<double> hv(8*n);
    if (c->fp64) {
      HIPCHK(hipMemcpy(hv.data(), c->V, 8*n*sizeof(double), hipMemcpyDeviceToHost));
    } else {
      std::vector<float> tmp(8*n);
      HIPCHK(hipMemcpy(tmp.data(), c->V, 8*n*sizeof(float), hipMemcpyDeviceToHost));
      for (size_t q = 0; q < 8*n; q++) hv[q] = tmp[q];
    }
    for (int nv = 0; nv < 8; nv++)
      for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
          for (int i = 0; i < nx; i++)
            vc[L.cell(nv, i, j, k)] = hv[nv*n + ((size_t)k*ny + j)*nx + i];
  }
  if (vs) {
    std::vector<double> hs(3*n);
    HIPCHK(hipMemcpy(hs.data(), c->cur_vs, 3*n*sizeof(double), hipMemcpyDeviceToHost));
    for (int d = 0; d < 3; d++) {
      if (!vs[d]) continue;
      for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
          for (int i = 0; i < nx; i++)
            vs[d][L.face(d, i, j, k)] = hs[d*n + ((size_t)k*ny + j)*nx + i];
      if (L.ary) {   // the duplicate upper face equals the lower one (PeriodicBoundary)
        for (int k = 0; k < nz; k++)
          for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
              if (d == 0 && i == 0) vs[0][L.face(0, nx, j, k)] = hs[((size_t)k*ny + j)*nx];
              if (d == 1 && j == 0) vs[1][L.face(1, i, ny, k)] = hs[n + ((size_t)k*ny)*nx + i];
              if (d == 2 && k == 0) vs[2][L.face(2, i, j, nz)] = hs[2*n + (size_t)j*nx + i];
            }
      }
    }
  }
  if (uc) {
    std::vector<double> hu(5*n), hv(8*n);
    HIPCHK(hipMemcpy(hu.data(), c->U, 5*n*sizeof(double), hipMemcpyDeviceToHost));
    // the cell-centred field of Uc is the face average (= Vc's B1..3)
    if (c->fp64) {
      HIPCHK(hipMemcpy(hv.data(), c->V, 8*n*sizeof(double), hipMemcpyDeviceToHost));
    } else {
      std::vector<float> tmp(8*n);
      HIPCHK(hipMemcpy(tmp.data(), c->V, 8*n*sizeof(float), hipMemcpyDeviceToHost));
      for (size_t q = 0; q < 8*n; q++) hv[q] = tmp[q];
    }
    const int map[8] = {0, 1, 2, 3, -1, -1, -1, 4};
    for (int nv = 0; nv < 8; nv++)
      for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
          for (int i = 0; i < nx; i++) {
            const size_t o = ((size_t)k*ny + j)*nx + i;
            uc[L.cell(nv, i, j, k)] = map[nv] >= 0 ? hu[map[nv]*n + o] : hv[nv*n + o];
          }
  }
  return 0;
}

// ---------------------------------------------------------------------------
// stepping
// ---------------------------------------------------------------------------
static void record_launch(gpl_ctx *c, int which, bool begin, hipEvent_t *slot)
{
  if (!c->timing.on) return;
  if (begin) { *slot = c->timing.get(); hipEventRecord(*slot, c->stream); }
  else {
    hipEvent_t e = c->timing.get();
    hipEventRecord(e, c->stream);
    c->timing.pending.push_back({which, {*slot, e}});
  }
}

template <typename R>
static void enqueue_stage(gpl_ctx *c, double *vin, double *vout, RKW w, int first_stage,
                          int want_dt, bool timed)
{
  Mhd3dBufs<R> b;
  b.V = (R *)c->V; b.U = c->U; b.U0 = c->U0; b.Vin = vin; b.Vout = vout; b.Bs0 = c->Bs0;
  b.F = (R *)c->F; b.CM = (R *)c->CM; b.FD = (R *)c->FD; b.E = (R *)c->E;
  b.dU = c->dU; b.Cdt = c->Cdt;
  b.cdt_part = c->cdt_part; b.err = c->err;
  static const int split_seq[] = {K_SWEEP_X, K_SWEEP_Y, K_SWEEP_Z, K_RHS, K_EMF, K_FINISH};
  static const int fused_seq[] = {K_SWEEPX_F, K_SWEEPY_F, K_SWEEPZ_F, K_EMF, K_FINISH};
  const int *seq = c->fused ? fused_seq : split_seq;
  const int nk = c->fused ? 5 : 6;
  for (int q = 0; q < nk; q++) {
    const int k = seq[q];
    hipEvent_t ev = nullptr;
    if (timed) record_launch(c, k, true, &ev);
    mhd3d_launch<R>(k, b, c->g, c->ctl, w, first_stage, want_dt, c->cfg.limiter, c->cfg.solver,
                    c->march, c->cfg.gamma, c->cfg.small_rho, c->cfg.small_prs, c->stream);
    if (timed) record_launch(c, k, false, &ev);
  }
}

// Enqueue one full AdvanceStep starting from faces *vs; returns the buffer
// that holds the faces afterwards.  advance_dt != 0 applies NextTimeStep on
// the device at the end of the step.
static double *enqueue_step(gpl_ctx *c, double *vs, int advance_dt, bool timed)
{
  // RK-SSP weights (Src/rk_step.cpp:20-51)
  RKW w1 = {0.0, 0.0, 1.0};
  RKW w2 = c->cfg.time_stepping == GPL_RK3 ? RKW{0.75, 0.25, 0.25} : RKW{0.5, 0.5, 0.5};
  RKW w3 = {1.0/3.0, 2.0/3.0, 2.0/3.0};
  const int nst = c->cfg.time_stepping == GPL_RK3 ? 3 : 2;
  double *in = vs, *out = vs == c->VsA ? c->VsB : c->VsA;
  for (int st = 1; st <= nst; st++) {
    RKW w = st == 1 ? w1 : (st == 2 ? w2 : w3);
    if (c->fp64) enqueue_stage<double>(c, in, out, w, st == 1, st == 1, timed);
    else         enqueue_stage<float>(c, in, out, w, st == 1, st == 1, timed);
    double *t = in; in = out; out = t;
  }
  hipEvent_t ev = nullptr;
  if (timed) record_launch(c, K_DT, true, &ev);
  launch_dt(c->cdt_part, c->nparts, 3.0, c->ctl, c->cfg.cfl, c->cfg.cfl_max_var, advance_dt, c->stream);
  if (timed) record_launch(c, K_DT, false, &ev);
  return in;
}

static void harvest_timing(gpl_ctx *c)
{
  for (auto &pe : c->timing.pending) {
    float ms = 0.f;
    hipEventSynchronize(pe.second.second);
    if (hipEventElapsedTime(&ms, pe.second.first, pe.second.second) == hipSuccess) {
      c->timing.launches[pe.first] += 1;
      c->timing.total_ms[pe.first] += ms;
    }
    c->timing.pool.push_back(pe.second.first);
    c->timing.pool.push_back(pe.second.second);
  }
  c->timing.pending.clear();
}

extern "C" uint32_t gpl_advance_step(gpl_ctx *c, double dt, double *inv_dt_hyp)
{
  if (!c) return GPL_STEP_API_ERROR;
  hipSetDevice(c->dev);
  StepCtl h_in, h_out;
  if (hipMemcpy(&h_in, c->ctl, sizeof(h_in), hipMemcpyDeviceToHost) != hipSuccess)
    return GPL_STEP_API_ERROR;
  h_in.dt = dt;
  const uint32_t zero = 0;
  if (hipStreamSynchronize(c->stream) != hipSuccess ||
      hipMemcpy(c->ctl, &h_in, sizeof(h_in), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->err, &zero, sizeof(zero), hipMemcpyHostToDevice) != hipSuccess)
    return GPL_STEP_API_ERROR;
  c->cur_vs = enqueue_step(c, c->cur_vs, 0, c->timing.on);
  uint32_t err = 0;
  if (hipStreamSynchronize(c->stream) != hipSuccess ||
      hipMemcpy(&h_out, c->ctl, sizeof(h_out), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(&err, c->err, sizeof(err), hipMemcpyDeviceToHost) != hipSuccess)
    return GPL_STEP_API_ERROR;
  harvest_timing(c);
  if (inv_dt_hyp) *inv_dt_hyp = h_out.inv_dt;
  return err;
}

extern "C" double gpl_next_time_step(const gpl_ctx *c, double dt, double inv_dt_hyp)
{
  const double dt_new = c->cfg.cfl/inv_dt_hyp;                       // next_time_step.cpp:80
  return dt_new <= c->cfg.cfl_max_var*dt ? dt_new : c->cfg.cfl_max_var*dt;   // :125 MIN()
}

static int build_graph(gpl_ctx *c)
{
  if (c->graph_exec && c->graph_vs == c->cur_vs) return 0;
  if (c->graph_exec) { hipGraphExecDestroy(c->graph_exec); c->graph_exec = nullptr; }
  // RK2 maps the face buffers A -> B -> A within a step; RK3 flips them, so
  // two steps are captured to return to the starting buffer.
  c->graph_steps = c->cfg.time_stepping == GPL_RK3 ? 2 : 1;
  hipGraph_t graph;
  if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return GPL_EHIP;
  double *vs = c->cur_vs;
  for (int s = 0; s < c->graph_steps; s++) vs = enqueue_step(c, vs, 1, false);
  if (hipStreamEndCapture(c->stream, &graph) != hipSuccess) return GPL_EHIP;
  if (vs != c->cur_vs) { hipGraphDestroy(graph); return GPL_EHIP; }
  hipError_t e = hipGraphInstantiate(&c->graph_exec, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  if (e != hipSuccess) { c->graph_exec = nullptr; return GPL_EHIP; }
  c->graph_vs = c->cur_vs;
  return 0;
}

extern "C" uint32_t gpl_run(gpl_ctx *c, int nsteps, double first_dt, double *time_out, double *dt_out)
{
  if (!c || nsteps < 0) return GPL_STEP_API_ERROR;
  hipSetDevice(c->dev);
  StepCtl h_in, h_out;
  if (hipStreamSynchronize(c->stream) != hipSuccess ||
      hipMemcpy(&h_in, c->ctl, sizeof(h_in), hipMemcpyDeviceToHost) != hipSuccess)
    return GPL_STEP_API_ERROR;
  if (first_dt > 0.0) { h_in.dt = first_dt; h_in.time = 0.0; h_in.nstep = 0; }
  const uint32_t zero = 0;
  if (hipMemcpy(c->ctl, &h_in, sizeof(h_in), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->err, &zero, sizeof(zero), hipMemcpyHostToDevice) != hipSuccess)
    return GPL_STEP_API_ERROR;
  const bool use_graph = !c->timing.on && getenv("GPL_NO_GRAPH") == nullptr;
  int done = 0;
  hipEventRecord(c->run_beg, c->stream);
  if (use_graph && nsteps > 0) {
    if (build_graph(c) != 0) return GPL_STEP_API_ERROR;
    while (nsteps - done >= c->graph_steps) {
      if (hipGraphLaunch(c->graph_exec, c->stream) != hipSuccess) return GPL_STEP_API_ERROR;
      done += c->graph_steps;
    }
  }
  while (done < nsteps) {
    c->cur_vs = enqueue_step(c, c->cur_vs, 1, c->timing.on);
    done++;
  }
  hipEventRecord(c->run_end, c->stream);
  uint32_t err = 0;
  if (hipStreamSynchronize(c->stream) != hipSuccess ||
      hipMemcpy(&h_out, c->ctl, sizeof(h_out), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(&err, c->err, sizeof(err), hipMemcpyDeviceToHost) != hipSuccess)
    return GPL_STEP_API_ERROR;
  float ms = 0.f;
  hipEventElapsedTime(&ms, c->run_beg, c->run_end);
  c->last_run_ms = ms;
  harvest_timing(c);
  if (time_out) *time_out = h_out.time;
  if (dt_out) *dt_out = h_out.dt;
  return err;
}

extern "C" int gpl_set_timing(gpl_ctx *c, int enable)
{
  if (!c) return GPL_EINVAL;
  c->timing.on = enable != 0;
  for (int k = 0; k < K_COUNT; k++) { c->timing.launches[k] = 0; c->timing.total_ms[k] = 0.0; }
  return 0;
}

extern "C" int gpl_kernel_count(const gpl_ctx *) { return K_COUNT; }

extern "C" int gpl_kernel_stats(gpl_ctx *c, int idx, const char **name, long long *launches,
                                double *total_ms)
{
  if (!c || idx < 0 || idx >= K_COUNT) return GPL_EINVAL;
  harvest_timing(c);
  if (name) *name = kKernelNames[idx];
  if (launches) *launches = c->timing.launches[idx];
  if (total_ms) *total_ms = c->timing.total_ms[idx];
  return 0;
}

extern "C" double gpl_last_run_ms(const gpl_ctx *c) { return c ? c->last_run_ms : 0.0; }

extern "C" size_t gpl_device_bytes(const gpl_ctx *c) { return c ? c->bytes : 0; }

extern "C" const char *gpl_strerror(int code)
{
  switch (code) {
  case 0: return "ok";
  case GPL_EINVAL: return "invalid argument";
  case GPL_EHIP: return "HIP runtime error";
  case GPL_ENOMEM: return "out of device memory";
  case GPL_EUNSUPPORTED: return "configuration not supported";
  default: return "unknown error";
  }
}
