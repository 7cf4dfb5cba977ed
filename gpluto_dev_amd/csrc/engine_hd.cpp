// HD engine (1D/2D/3D, periodic, outflow or reflective sides, HLLC/HLL, RK2/RK3, FP64):
// device state and AdvanceStep orchestration (Src/rk_step.cpp:88-720) over
// the kernels of hd_kernels.hip.
#include "c_guard.hpp"
#include "diagnostics.hpp"
#include "engine.hpp"
#include "hd.hpp"
#include "physics.hpp"

#include <new>
#include <vector>

namespace gpl {

namespace {

class HdEngine : public Engine {
public:
  gpl_config cfg;
  EngineEnv env;
  HdBox g;
  size_t n = 0;
  double *VA = nullptr, *VB = nullptr;   // [5][n] primitives, ping-pong across stages
  double *U = nullptr, *U0 = nullptr, *dU = nullptr, *Cdt = nullptr, *cdt_part = nullptr;
  int nparts = 0;
  double *cur_v = nullptr;
  double *diag = nullptr;     // diagnostics partials (DIAG_BLOCKS rows) and result row

  ~HdEngine() override
  {
    void *ptrs[] = {VA, VB, U, U0, dU, Cdt, cdt_part, diag};
    for (void *p : ptrs) if (p) (void)hipFree(p);
  }

  int init(const gpl_config &c, const EngineEnv &e)
  {
    cfg = c; env = e;
    n = (size_t)c.nx*c.ny*c.nz;
    g.nx = c.nx; g.ny = c.ny; g.nz = c.nz;
    g.ncell = (int)n;
    g.dims = c.dimensions;
    for (int d = 0; d < 3; d++) {
      g.lo[d] = d < c.dimensions ? side_type(c, 2*d) : GPL_PERIODIC;
      g.hi[d] = d < c.dimensions ? side_type(c, 2*d + 1) : GPL_PERIODIC;
    }
    // dx = (xend - xbeg)/N (Src/set_grid.cpp:196); unused directions are never swept
    g.dx = (c.xend[0] - c.xbeg[0])/c.nx;
    g.dy = (c.xend[1] - c.xbeg[1])/c.ny;
    g.dz = (c.xend[2] - c.xbeg[2])/c.nz;
    nparts = hd_blocks(g);
    int rc = 0;
    rc |= dev_alloc(bytes, (void **)&VA, 5*n*sizeof(double));
    rc |= dev_alloc(bytes, (void **)&VB, 5*n*sizeof(double));
    rc |= dev_alloc(bytes, (void **)&U, 5*n*sizeof(double));
    rc |= dev_alloc(bytes, (void **)&U0, 5*n*sizeof(double));
    if (g.dims > 1) {
      rc |= dev_alloc(bytes, (void **)&dU, 5*n*sizeof(double));
      rc |= dev_alloc(bytes, (void **)&Cdt, n*sizeof(double));
    }
    rc |= dev_alloc(bytes, (void **)&cdt_part, (size_t)nparts*sizeof(double));
    rc |= dev_alloc(bytes, (void **)&diag, (size_t)(DIAG_BLOCKS + 1)*DIAG_MAXQ*sizeof(double));
    cur_v = VA;
    return rc ? GPL_ENOMEM : 0;
  }

  int upload(const double *vc, const double *const *, int layout) override
  {
    const Layout L = make_layout(cfg, layout);
    std::vector<double> hv(5*n), hu(5*n);
    for (int k = 0; k < L.nz; k++)
      for (int j = 0; j < L.ny; j++)
        for (int i = 0; i < L.nx; i++) {
          const size_t o = ((size_t)k*L.ny + j)*L.nx + i;
          double v[5], u[5];
          for (int nv = 0; nv < 5; nv++) v[nv] = vc[L.cell(nv, i, j, k)];
          // Uc = PrimToCons(Vc) (Src/startup.cpp:514-561)
          hd_prim_to_cons<double>(v, u, cfg.gamma - 1.0);
          for (int nv = 0; nv < 5; nv++) { hv[nv*n + o] = v[nv]; hu[nv*n + o] = u[nv]; }
        }
    if (hipMemcpy(VA, hv.data(), 5*n*sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(U, hu.data(), 5*n*sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      return GPL_EHIP;
    cur_v = VA;
    return 0;
  }

  // the primitives are stored active-only, i fastest: the dump image itself
  int dump_gather(double *dst, hipStream_t s) override
  {
    return hipMemcpyAsync(dst, cur_v, 5*n*sizeof(double), hipMemcpyDeviceToDevice, s) == hipSuccess
               ? 0 : GPL_EHIP;
  }

  int download(double *vc, double *const *, double *uc, int layout) override
  {
    const Layout L = make_layout(cfg, layout);
    for (int which = 0; which < 2; which++) {
      double *dst = which == 0 ? vc : uc;
      if (!dst) continue;
      std::vector<double> h(5*n);
      if (hipMemcpy(h.data(), which == 0 ? cur_v : U, 5*n*sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        return GPL_EHIP;
      for (int nv = 0; nv < 5; nv++)
        for (int k = 0; k < L.nz; k++)
          for (int j = 0; j < L.ny; j++)
            for (int i = 0; i < L.nx; i++)
              dst[L.cell(nv, i, j, k)] = h[nv*n + ((size_t)k*L.ny + j)*L.nx + i];
    }
    return 0;
  }

  void enqueue_step(int advance_dt, LaunchTimer *t) override
  {
    // RK-SSP weights (Src/rk_step.cpp:20-51)
    const RKW w1 = {0.0, 0.0, 1.0};
    const RKW w2 = cfg.time_stepping == GPL_RK3 ? RKW{0.75, 0.25, 0.25} : RKW{0.5, 0.5, 0.5};
    const RKW w3 = {1.0/3.0, 2.0/3.0, 2.0/3.0};
    const int nst = cfg.time_stepping == GPL_RK3 ? 3 : 2;
    static const int kid[3] = {K_HD_X, K_HD_Y, K_HD_Z};
    double *in = cur_v, *out = cur_v == VA ? VB : VA;
    if (zcnt) launch_zone_reset(zcnt, env.stream);
    for (int st = 1; st <= nst; st++) {
      const RKW w = st == 1 ? w1 : (st == 2 ? w2 : w3);
      HdBufs b = {in, out, U, U0, dU, Cdt, cdt_part, env.err, zone_rep(st - 1)};
      for (int d = 0; d < g.dims; d++) {
        if (t) t->begin(kid[d]);
        hd_launch(d, b, g, env.ctl, w, st == 1, st == 1, cfg.limiter, cfg.solver, cfg.gamma,
                  cfg.small_rho, cfg.small_prs, env.stream);
        if (t) t->end(kid[d]);
      }
      double *tmp = in; in = out; out = tmp;
    }
    if (t) t->begin(K_DT);
    launch_dt(cdt_part, nparts, (double)g.dims, env.ctl, cfg.cfl, cfg.cfl_max_var, advance_dt,
              env.stream);
    if (t) t->end(K_DT);
    cur_v = in;
  }

  int graph_period() const override { return cfg.time_stepping == GPL_RK3 ? 2 : 1; }
  const void *state_key() const override { return cur_v; }

  int nvar() const override { return 5; }
  int nstages() const override { return cfg.time_stepping == GPL_RK3 ? 3 : 2; }
  // the device cells are the active zones, i fastest
  size_t zone_cells() const override { return n; }
  size_t zone_active() const override { return n; }
  void zone_extract(const uint32_t *src, uint32_t *dst) const override
  {
    for (size_t c = 0; c < n; c++) dst[c] = src[c];
  }

  int diagnostics(double *out) override
  {
    DiagBox d;
    d.nx = g.nx; d.ny = g.ny; d.nz = g.nz; d.ncell = g.ncell; d.nv = 5; d.nb = 0;
    d.lo[0] = d.lo[1] = d.lo[2] = 0;
    d.hi[0] = g.nx - 1; d.hi[1] = g.ny - 1; d.hi[2] = g.nz - 1;
    d.ax = d.ay = d.az = 0.0;
    double *res = diag + (size_t)DIAG_BLOCKS*DIAG_MAXQ;
    launch_diagnostics<double>(d, cur_v, nullptr, diag, res, env.stream);
    if (hipMemcpyAsync(out, res, 12*sizeof(double), hipMemcpyDeviceToHost, env.stream) != hipSuccess ||
        hipStreamSynchronize(env.stream) != hipSuccess)
      return GPL_EHIP;
    return 0;
  }
};

}  // namespace

int make_hd_engine(const gpl_config &cfg, const EngineEnv &env, Engine **out)
{
  *out = nullptr;
  if (cfg.dimensions < 1 || cfg.dimensions > 3) return GPL_EINVAL;
  for (int d = 0; d < cfg.dimensions; d++) {
    const int lo = side_type(cfg, 2*d), hi = side_type(cfg, 2*d + 1);
    if (lo < GPL_PERIODIC || lo > GPL_REFLECTIVE || hi < GPL_PERIODIC || hi > GPL_REFLECTIVE)
      return GPL_EUNSUPPORTED;
    if ((lo == GPL_PERIODIC) != (hi == GPL_PERIODIC)) return GPL_EINVAL;
  }
  if (cfg.solver != GPL_SOLVER_HLLC && cfg.solver != GPL_SOLVER_HLL) return GPL_EUNSUPPORTED;
  if (cfg.precision != GPL_FP64) return GPL_EUNSUPPORTED;
  const int nn[3] = {cfg.nx, cfg.ny, cfg.nz};
  for (int d = 0; d < 3; d++) {
    if (d < cfg.dimensions && nn[d] < 3) return GPL_EINVAL;   // the PLM stencil spans 5 cells
    if (d >= cfg.dimensions && nn[d] != 1) return GPL_EINVAL; // unused directions hold 1 zone
  }
  HdEngine *e = new (std::nothrow) HdEngine();
  if (!e) return GPL_ENOMEM;
  const int rc = c_guard(GPL_ENOMEM, [&] { return e->init(cfg, env); });
  if (rc) { delete e; return rc; }
  *out = e;
  return 0;
}

}  // namespace gpl
