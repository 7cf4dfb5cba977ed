// 3D MHD engine (ideal MHD, or relativistic MHD with the same CT machinery):
// device state and AdvanceStep orchestration (Src/rk_step.cpp:88-720) over
// the kernels of mhd3d_kernels.hpp.
#include "c_guard.hpp"
#include "engine.hpp"
#include "mhd3d.hpp"
#include "c2p_net.hpp"
#include "diagnostics.hpp"
#include "physics.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

namespace gpl {

namespace {

template <typename T>
int to_device(void *dst, const std::vector<double> &src)
{
  if (sizeof(T) == sizeof(double))
    return hipMemcpy(dst, src.data(), src.size()*sizeof(double), hipMemcpyHostToDevice) == hipSuccess
           ? 0 : GPL_EHIP;
  std::vector<T> tmp(src.size());
  for (size_t q = 0; q < src.size(); q++) tmp[q] = (T)src[q];
  return hipMemcpy(dst, tmp.data(), tmp.size()*sizeof(T), hipMemcpyHostToDevice) == hipSuccess
         ? 0 : GPL_EHIP;
}

int from_device_v(void *src, bool fp64, std::vector<double> &dst)
{
  if (fp64)
    return hipMemcpy(dst.data(), src, dst.size()*sizeof(double), hipMemcpyDeviceToHost) == hipSuccess
           ? 0 : GPL_EHIP;
  std::vector<float> tmp(dst.size());
  if (hipMemcpy(tmp.data(), src, tmp.size()*sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
    return GPL_EHIP;
  for (size_t q = 0; q < tmp.size(); q++) dst[q] = tmp[q];
  return 0;
}

class Mhd3dEngine : public Engine {
public:
  gpl_config cfg;
  EngineEnv env;
  Box3 g;
  bool fp64 = true;
  bool rmhd = false;
  size_t n = 0;
  void *V = nullptr;          // [8][n] R
  // Conserved state: U holds U^n between steps; the first stage writes its
  // state into Us, later stages read Us and U^n and the last writes U.
  double *U = nullptr, *Us = nullptr;
  // Staggered faces rotate over three buffers: the step starts from vs[cur],
  // stage 1 writes vs[cur+1], stage 2 vs[cur+2] and RK3's stage 3 vs[cur+1].
  double *vs[3] = {nullptr, nullptr, nullptr};
  int cur = 0;
  void *F = nullptr, *CM = nullptr, *FD = nullptr, *E = nullptr;
  void *dU = nullptr, *Cdt = nullptr;   // R: the within-stage sums (FP64 in the drop-in)
  bool fused = true;          // fused sweep kernels (nx <= 512) or the split path
  int march = 16;             // cells per thread in the y/z marching sweeps
  // x and y sweeps in one row-marching kernel: the mixed form, where it saves
  // HBM traffic (its three LDS rows must fit in 64 KB); the FP64 sweeps are
  // VALU-bound and run faster as two kernels
  bool fused_xy = false;
  // EMF kernel: n > 0 register z-march of n edges per thread, -n LDS-staged
  // 64 x 4 tiles marching n planes (the default, 32), 0 one edge per thread
  int emf_march = -32;
  double *cdt_part = nullptr;
  int nparts = 0;
  NetPack *net = nullptr;     // RMHD learned con2prim guess (matrix-core form), when enabled
  float *net_w = nullptr;     // its weight vector
  bool net_on = false;
  double *diag = nullptr;     // diagnostics partials (DIAG_BLOCKS rows) and result row
  // A box with an outflow or reflective side runs padded: the device box is the reference's
  // array, two ghost layers + the active cells + two ghost layers + one spare
  // layer (the outermost faces), refilled at every stage by mhd3d_boundary;
  // the kernels treat it as periodic, and the cells within two of its edges,
  // whose wrapped values never reach an active result, are refilled
  // before they are read again.
  bool padded = false;
  int side[6] = {GPL_PERIODIC, GPL_PERIODIC, GPL_PERIODIC, GPL_PERIODIC, GPL_PERIODIC, GPL_PERIODIC};
  int na[3] = {0, 0, 0};      // active cells per direction
  // Slab pipeline (FP64 fused kernels with the EMF tiles): a stage runs the
  // three sweeps slab by slab of z-planes on the context stream while a
  // second stream runs the EMF and the finish of the slabs whose inputs are
  // complete, so the VALU-bound sweeps and the HBM-bound EMF and finish
  // overlap.  The finish then writes the primitives into a second buffer
  // (Valt), since later slabs' sweeps still read the stage's input; the
  // buffers swap after every stage.  Each cell's arithmetic is unchanged.
  bool pipelined = false;
  int nslab = 0, slab_planes = 0;
  void *Valt = nullptr;       // the other primitive buffer (pipelined)
  void *V0 = nullptr;         // the buffer V held after upload (graph key)
  hipStream_t s2 = nullptr;   // the EMF/finish stream (default priority)
  static constexpr int kMaxSlab = 64;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_slab[kMaxSlab] = {};
  std::vector<int> slab_part0;   // first C_dt partial slot of each slab's z march
  char key_tab[3][2] = {};       // graph keys: (face buffer, primitive buffer)

  ~Mhd3dEngine() override
  {
    void *ptrs[] = {V, Valt, U, Us, vs[0], vs[1], vs[2], F, CM, FD, E, dU, Cdt, cdt_part, net, net_w, diag};
    for (void *p : ptrs) if (p) (void)hipFree(p);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    for (hipEvent_t e : ev_slab) if (e) (void)hipEventDestroy(e);
    if (s2) (void)hipStreamDestroy(s2);
  }

  int init(const gpl_config &c, const EngineEnv &e)
  {
    cfg = c; env = e;
    fp64 = c.precision == GPL_FP64;
    rmhd = c.physics == GPL_RMHD;
    na[0] = c.nx; na[1] = c.ny; na[2] = c.nz;
    padded = !all_periodic(c);
    for (int q = 0; q < 6; q++) side[q] = side_type(c, q);
    const int pad = padded ? 5 : 0, off = padded ? 2 : 0;
    g.nx = c.nx + pad; g.ny = c.ny + pad; g.nz = c.nz + pad;
    for (int d = 0; d < 3; d++) { g.lo[d] = off; g.hi[d] = off + na[d] - 1; }
    n = (size_t)g.nx*g.ny*g.nz;
    g.ncell = (int)n;
    g.kb = 0; g.ke = g.nz; g.part0 = 0;
    // dx = (xend - xbeg)/N for every cell (Src/set_grid.cpp:196)
    g.small_rho = c.small_rho;
    g.small_prs = c.small_prs;
    g.dx = (c.xend[0] - c.xbeg[0])/c.nx;
    g.dy = (c.xend[1] - c.xbeg[1])/c.ny;
    g.dz = (c.xend[2] - c.xbeg[2])/c.nz;
    const size_t rs = fp64 ? sizeof(double) : sizeof(float);
    // the fused x sweep holds one periodic x-row per block
    fused = g.nx <= 512 && getenv("GPL_SPLIT_KERNELS") == nullptr;
    if (padded && !fused) return GPL_EUNSUPPORTED;
    // GPL_MARCH = 8/16/32: cells per thread in the y/z marching sweeps
    if (const char *m = getenv("GPL_MARCH")) march = atoi(m);
    if (march != 8 && march != 16 && march != 32) march = 16;
    emf_march = -32;
    if (const char *m = getenv("GPL_EMF_MARCH")) emf_march = atoi(m);
    {
      const int l = emf_march < 0 ? -emf_march : emf_march;
      if (emf_march != 0 && l != 4 && l != 8 && l != 16 && l != 32) emf_march = -32;
    }
    fused_xy = !fp64;
    if (const char *m = getenv("GPL_FUSED_XY")) fused_xy = atoi(m) != 0;
    fused_xy = fused_xy && fused && !rmhd && 20*(size_t)g.nx*rs <= 65536;
    nparts = fused ? mhd3d_march_blocks(g, march) : mhd3d_blocks(g);
    // the slab pipeline: GPL_PIPE_SLABS slabs (0 or 1: off); every slab but
    // the last starts and ends at a multiple of the y/z march and of the EMF
    // tiles' march length
    {
      int want = 0;
      if (const char *m = getenv("GPL_PIPE_SLABS")) want = atoi(m);
      // a slab's EMF launch covers only its planes: tiles marching 8 planes
      // keep four times as many workgroups in flight as the default 32
      if (want >= 3 && !getenv("GPL_EMF_MARCH")) emf_march = -8;
      const int lt = emf_march < 0 ? -emf_march : 0;
      if (want >= 3 && fp64 && fused && !fused_xy && lt > 0 &&
          c.ct_emf_average == GPL_UCT_HLL && want <= kMaxSlab) {
        const int unit = march > lt ? march : lt;
        int p = (g.nz + want - 1)/want;
        p = (p + unit - 1)/unit*unit;
        if (p > 0 && (g.nz + p - 1)/p >= 3) {
          pipelined = true;
          slab_planes = p;
          nslab = (g.nz + p - 1)/p;
        }
      }
      if (!pipelined && want >= 3 && !getenv("GPL_EMF_MARCH")) emf_march = -32;
    }
    if (pipelined) {
      // C_dt partial slots: each slab's z march writes its own range
      slab_part0.resize(nslab);
      int total = 0;
      for (int s = 0; s < nslab; s++) {
        slab_part0[s] = total;
        const int nk = (s + 1)*slab_planes < g.nz ? slab_planes : g.nz - s*slab_planes;
        const long long nt = (long long)g.nx*g.ny*((nk + march - 1)/march);
        total += (int)((nt + 255)/256);
      }
      nparts = total;
      if (hipStreamCreateWithFlags(&s2, hipStreamNonBlocking) != hipSuccess ||
          hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming) != hipSuccess ||
          hipEventCreateWithFlags(&ev_join, hipEventDisableTiming) != hipSuccess)
        return GPL_EHIP;
      for (int s = 0; s < nslab; s++)
        if (hipEventCreateWithFlags(&ev_slab[s], hipEventDisableTiming) != hipSuccess) return GPL_EHIP;
    }
    int rc = 0;
    rc |= dev_alloc(bytes, &V, 8*n*rs);
    if (pipelined) rc |= dev_alloc(bytes, &Valt, 8*n*rs);
    V0 = V;
    rc |= dev_alloc(bytes, (void **)&U, 5*n*sizeof(double));
    rc |= dev_alloc(bytes, (void **)&Us, 5*n*sizeof(double));
    for (int q = 0; q < 3; q++) rc |= dev_alloc(bytes, (void **)&vs[q], 3*n*sizeof(double));
    if (fused) {
      rc |= dev_alloc(bytes, &dU, 5*n*rs);
      rc |= dev_alloc(bytes, &Cdt, n*rs);
    } else {
      rc |= dev_alloc(bytes, &F, 15*n*rs);
      rc |= dev_alloc(bytes, &CM, 3*n*rs);
    }
    // UCT face records: 4 per direction (et, eb, aL_1, aR_1), 6 with UCT_HLLD
    rc |= dev_alloc(bytes, &FD, (c.ct_emf_average == GPL_UCT_HLLD ? 18 : 12)*n*rs);
    rc |= dev_alloc(bytes, &E, 3*n*rs);
    rc |= dev_alloc(bytes, (void **)&cdt_part, (size_t)nparts*sizeof(double));
    rc |= dev_alloc(bytes, (void **)&diag, (size_t)(DIAG_BLOCKS + 1)*DIAG_MAXQ*sizeof(double));
    cur = 0;
    return rc ? GPL_ENOMEM : 0;
  }

  // device index of active cell (i, j, k) and of the device cell that an
  // index of the padded box reads at upload (image or clamped edge)
  size_t dev_cell(int i, int j, int k) const
  {
    return (size_t)(i + g.lo[0]) + (size_t)g.nx*((size_t)(j + g.lo[1]) + (size_t)g.ny*(k + g.lo[2]));
  }
  int src_index(int d, int p, bool wrap) const
  {
    const int a = p - g.lo[d], m = na[d];
    if (wrap || side[2*d] == GPL_PERIODIC) return ((a % m) + m) % m;
    return a < 0 ? 0 : (a >= m ? m - 1 : a);
  }

  // host index along dd of the face that device face (component d, index p
  // along dd) takes at upload: the periodic image, except the upper
  // boundary face of a non-periodic direction, which the reference's array
  // layout carries (a restart dump's face n, Src/restart.cpp:145-172)
  int face_src(const Layout &L, int d, int dd, int p) const
  {
    if (L.ary && dd == d && d < cfg.dimensions && side[2*d] != GPL_PERIODIC && p - g.lo[d] == na[d])
      return na[d];
    return src_index(dd, p, true);
  }

  int upload(const double *vc, const double *const *vsh, int layout) override
  {
    if (!vsh || !vsh[0] || !vsh[1] || !vsh[2]) return GPL_EINVAL;
    const Layout L = make_layout(cfg, layout);
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    std::vector<double> hv(8*n), hu(5*n), hs(3*n);
    // the whole device box: active values, and in the ghost layers of a
    // padded box the periodic image or the clamped edge value of a cell, and
    // the periodic image of a face (the seeded IC's face field is the curl of
    // a periodic potential; with the active layout an outflow box's upper
    // boundary face starts as the image of its lower one)
    for (int d = 0; d < 3; d++)
      for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
          for (int i = 0; i < nx; i++)
            hs[d*n + ((size_t)k*ny + j)*nx + i] =
                vsh[d][L.face(d, face_src(L, d, 0, i), face_src(L, d, 1, j), face_src(L, d, 2, k))];
    for (int k = 0; k < nz; k++)
      for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
          const size_t o = ((size_t)k*ny + j)*nx + i;
          double v[8], u[8];
          for (int nv = 0; nv < 8; nv++)
            v[nv] = vc[L.cell(nv, src_index(0, i, false), src_index(1, j, false), src_index(2, k, false))];
          // cell-centred field = face average (Src/startup.cpp:388 -> ct_field_average.cpp:182-184)
          const int ip = i + 1 == nx ? 0 : i + 1, jp = j + 1 == ny ? 0 : j + 1, kp = k + 1 == nz ? 0 : k + 1;
          v[4] = 0.5*(hs[0*n + o] + hs[0*n + ((size_t)k*ny + j)*nx + ip]);
          v[5] = 0.5*(hs[1*n + o] + hs[1*n + ((size_t)k*ny + jp)*nx + i]);
          v[6] = 0.5*(hs[2*n + o] + hs[2*n + ((size_t)kp*ny + j)*nx + i]);
          // Uc = PrimToCons(Vc) (Src/startup.cpp:514-561)
          if (rmhd) rmhd_prim_to_cons<double>(v, u, cfg.gamma);
          else      mhd_prim_to_cons<double>(v, u, cfg.gamma - 1.0);
          for (int nv = 0; nv < 8; nv++) hv[nv*n + o] = v[nv];
          hu[0*n + o] = u[0]; hu[1*n + o] = u[1]; hu[2*n + o] = u[2]; hu[3*n + o] = u[3];
          hu[4*n + o] = u[7];
        }
    int rc = fp64 ? to_device<double>(V, hv) : to_device<float>(V, hv);
    if (rc) return rc;
    if (hipMemcpy(U, hu.data(), 5*n*sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(vs[0], hs.data(), 3*n*sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      return GPL_EHIP;
    cur = 0;
    return 0;
  }

  int dump_gather(double *dst, hipStream_t s) override
  {
    int per[3];
    for (int d = 0; d < 3; d++) per[d] = side[2*d] == GPL_PERIODIC;
    if (fp64) mhd3d_dump_gather<double>(g, na, per, (const double *)V, vs[cur], dst, s);
    else      mhd3d_dump_gather<float>(g, na, per, (const float *)V, vs[cur], dst, s);
    return hipGetLastError() == hipSuccess ? 0 : GPL_EHIP;
  }

  int download(double *vc, double *const *vsh, double *uc, int layout) override
  {
    const Layout L = make_layout(cfg, layout);
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    std::vector<double> hv(8*n);
    if ((vc || uc) && from_device_v(V, fp64, hv)) return GPL_EHIP;
    if (vc) {
      for (int nv = 0; nv < 8; nv++)
        for (int k = 0; k < nz; k++)
          for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++)
              vc[L.cell(nv, i, j, k)] = hv[nv*n + dev_cell(i, j, k)];
    }
    if (vsh) {
      std::vector<double> hs(3*n);
      if (hipMemcpy(hs.data(), vs[cur], 3*n*sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        return GPL_EHIP;
      for (int d = 0; d < 3; d++) {
        if (!vsh[d]) continue;
        for (int k = 0; k < nz; k++)
          for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++)
              vsh[d][L.face(d, i, j, k)] = hs[d*n + dev_cell(i, j, k)];
        if (L.ary) {
          // the upper boundary face: the periodic image of the lower one
          // (PeriodicBoundary), or its own value on an outflow side
          for (int k = 0; k < nz; k++)
            for (int j = 0; j < ny; j++)
              for (int i = 0; i < nx; i++) {
                const bool per = side[2*d] == GPL_PERIODIC;
                if (d == 0 && i == 0) vsh[0][L.face(0, nx, j, k)] = hs[dev_cell(per ? 0 : nx, j, k)];
                if (d == 1 && j == 0) vsh[1][L.face(1, i, ny, k)] = hs[n + dev_cell(i, per ? 0 : ny, k)];
                if (d == 2 && k == 0) vsh[2][L.face(2, i, j, nz)] = hs[2*n + dev_cell(i, j, per ? 0 : nz)];
              }
        }
      }
    }
    if (uc) {
      std::vector<double> hu(5*n);
      if (hipMemcpy(hu.data(), U, 5*n*sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        return GPL_EHIP;
      // the cell-centred field of Uc is the face average (= Vc's B1..3)
      const int map[8] = {0, 1, 2, 3, -1, -1, -1, 4};
      for (int nv = 0; nv < 8; nv++)
        for (int k = 0; k < nz; k++)
          for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
              const size_t o = dev_cell(i, j, k);
              uc[L.cell(nv, i, j, k)] = map[nv] >= 0 ? hu[map[nv]*n + o] : hv[nv*n + o];
            }
    }
    return 0;
  }

  template <typename R>
  void enqueue_stage(const double *uin, double *uout, double *vin, double *vout,
                     const double *vbase, RKW w, int first_stage, int want_dt, int stage_idx,
                     LaunchTimer *t)
  {
    // Boundary (Src/rk_step.cpp: at the start of every stage) on a padded box
    if (padded) mhd3d_boundary<R>(g, side, (R *)V, vin, env.stream);
    Mhd3dBufs<R> b;
    b.V = (R *)V; b.Uin = uin; b.U = uout; b.U0 = U;
    b.Vin = vin; b.Vout = vout; b.Bs0 = vbase;
    b.F = (R *)F; b.CM = (R *)CM; b.FD = (R *)FD; b.E = (R *)E;
    b.dU = (R *)dU; b.Cdt = (R *)Cdt;
    b.cdt_part = cdt_part; b.err = env.err;
    b.emf_march = emf_march;
    b.phys = rmhd ? 1 : 0;
    b.net = net_on ? net : nullptr;
    b.zr = zone_rep(stage_idx);
    static const int split_seq[] = {K_SWEEP_X, K_SWEEP_Y, K_SWEEP_Z, K_RHS, K_EMF, K_FINISH};
    static const int fused_seq[] = {K_SWEEPX_F, K_SWEEPY_F, K_SWEEPZ_F, K_EMF, K_FINISH};
    static const int xy_seq[] = {K_SWEEPXY_F, K_SWEEPZ_F, K_EMF, K_FINISH};
    if (pipelined && !t) {
      enqueue_stage_pipe<R>(b, w, first_stage, want_dt);
      return;
    }
    const int *seq = fused ? (fused_xy ? xy_seq : fused_seq) : split_seq;
    const int nk = fused ? (fused_xy ? 4 : 5) : 6;
    for (int q = 0; q < nk; q++) {
      const int k = seq[q];
      if (t) t->begin(k);
      // solver 4: HLLD with the UCT_HLLD edge EMF (CT_Store_UCTHlld)
      const int solver = cfg.ct_emf_average == GPL_UCT_HLLD ? 4 : cfg.solver;
      mhd3d_launch<R>(k, b, g, env.ctl, w, first_stage, want_dt, cfg.limiter, solver,
                      march, cfg.gamma, cfg.small_rho, cfg.small_prs, env.stream);
      if (t) t->end(k);
    }
  }

  // One stage as a slab pipeline.  Slab s covers planes [s P, (s+1) P).  The
  // EMF of slab s reads face records and faces of planes s P - 2 .. (s+1) P
  // (its z stencils and the next plane's records), so it waits for the
  // sweeps of slabs s-1, s, s+1 (slab 0 and the last slab, which wrap, for
  // every slab); the finish of slab s reads the edge EMFs of planes s P ..
  // (s+1) P, so it follows the EMF of slab s+1 (of slab 0 for the last).  It
  // writes the primitives into Valt, which no sweep of the stage reads.  The
  // stage ends when both streams have drained (the next stage's z march
  // reads every plane).
  template <typename R>
  void enqueue_stage_pipe(const Mhd3dBufs<R> &b, RKW w, int first_stage, int want_dt)
  {
    const int solver = cfg.solver;
    const hipStream_t sa = env.stream;
    Mhd3dBufs<R> bf = b;
    bf.V = (R *)Valt;                       // the finish writes the other buffer
    auto slab_box = [&](int s) {
      Box3 q = g;
      q.kb = s*slab_planes;
      q.ke = (s + 1)*slab_planes < g.nz ? (s + 1)*slab_planes : g.nz;
      q.part0 = slab_part0[s];
      return q;
    };
    auto launch = [&](int k, const Mhd3dBufs<R> &bb, int s, hipStream_t st) {
      mhd3d_launch<R>(k, bb, slab_box(s), env.ctl, w, first_stage, want_dt, cfg.limiter, solver,
                      march, cfg.gamma, cfg.small_rho, cfg.small_prs, st);
    };
    (void)hipEventRecord(ev_fork, sa);
    (void)hipStreamWaitEvent(s2, ev_fork, 0);
    for (int s = 0; s < nslab; s++) {
      launch(K_SWEEPX_F, b, s, sa);
      launch(K_SWEEPY_F, b, s, sa);
      launch(K_SWEEPZ_F, b, s, sa);
      (void)hipEventRecord(ev_slab[s], sa);
    }
    const int S = nslab;
    for (int s = 1; s <= S - 2; s++) {
      (void)hipStreamWaitEvent(s2, ev_slab[s + 1], 0);
      launch(K_EMF, b, s, s2);
      if (s >= 2) launch(K_FINISH, bf, s - 1, s2);
    }
    (void)hipStreamWaitEvent(s2, ev_slab[S - 1], 0);
    launch(K_EMF, b, S - 1, s2);
    launch(K_EMF, b, 0, s2);
    launch(K_FINISH, bf, S - 2, s2);
    launch(K_FINISH, bf, S - 1, s2);
    launch(K_FINISH, bf, 0, s2);
    (void)hipEventRecord(ev_join, s2);
    (void)hipStreamWaitEvent(sa, ev_join, 0);
    std::swap(V, Valt);
  }

  void enqueue_step(int advance_dt, LaunchTimer *t) override
  {
    // RK-SSP weights (Src/rk_step.cpp:20-51)
    const RKW w1 = {0.0, 0.0, 1.0};
    const RKW w2 = cfg.time_stepping == GPL_RK3 ? RKW{0.75, 0.25, 0.25} : RKW{0.5, 0.5, 0.5};
    const RKW w3 = {1.0/3.0, 2.0/3.0, 2.0/3.0};
    const bool rk3 = cfg.time_stepping == GPL_RK3;
    double *f0 = vs[cur], *f1 = vs[(cur + 1) % 3], *f2 = vs[(cur + 2) % 3];
    if (zcnt) launch_zone_reset(zcnt, env.stream);
    // stage 1: U^n, faces f0 -> Us, f1
    stage(U, Us, f0, f1, f0, w1, 1, 0, t);
    if (!rk3) {
      // stage 2: 1/2 U^n + 1/2 (Us + dt L(Us)) -> U, f2
      stage(Us, U, f1, f2, f0, w2, 0, 1, t);
      cur = (cur + 2) % 3;
    } else {
      stage(Us, Us, f1, f2, f0, w2, 0, 1, t);     // -> Us, f2
      stage(Us, U, f2, f1, f0, w3, 0, 2, t);      // -> U, f1
      cur = (cur + 1) % 3;
    }
    if (t) t->begin(K_DT);
    launch_dt(cdt_part, nparts, 3.0, env.ctl, cfg.cfl, cfg.cfl_max_var, advance_dt, env.stream);
    if (t) t->end(K_DT);
  }

  void stage(const double *uin, double *uout, double *vin, double *vout, const double *vbase,
             RKW w, int first, int stage_idx, LaunchTimer *t)
  {
    if (fp64) enqueue_stage<double>(uin, uout, vin, vout, vbase, w, first, first, stage_idx, t);
    else      enqueue_stage<float>(uin, uout, vin, vout, vbase, w, first, first, stage_idx, t);
  }

  // the face buffers rotate with period 3 (RK2: cur += 2, RK3: cur += 1)
  // faces: period 3; the primitive buffers of the pipeline swap every stage,
  // so RK3's three stages need six steps to return both
  int graph_period() const override { return pipelined && cfg.time_stepping == GPL_RK3 ? 6 : 3; }
  const void *state_key() const override { return &key_tab[cur][V == V0 ? 0 : 1]; }

  int nvar() const override { return 8; }
  int nstages() const override { return cfg.time_stepping == GPL_RK3 ? 3 : 2; }
  size_t zone_cells() const override { return (size_t)g.ncell; }
  size_t zone_active() const override { return (size_t)na[0]*na[1]*na[2]; }
  void zone_extract(const uint32_t *src, uint32_t *dst) const override
  {
    for (int k = 0; k < na[2]; k++)
      for (int j = 0; j < na[1]; j++)
        for (int i = 0; i < na[0]; i++)
          dst[((size_t)k*na[1] + j)*na[0] + i] = src[dev_cell(i, j, k)];
  }

  int diagnostics(double *out) override
  {
    DiagBox d;
    d.nx = g.nx; d.ny = g.ny; d.nz = g.nz; d.ncell = g.ncell; d.nv = 8; d.nb = 3;
    for (int q = 0; q < 3; q++) { d.lo[q] = g.lo[q]; d.hi[q] = g.hi[q]; }
    // face areas of the uniform Cartesian grid (Src/set_geometry.cpp:124,147,166)
    d.ax = 1.0*g.dy*g.dz; d.ay = g.dx*1.0*g.dz; d.az = g.dx*g.dy*1.0;
    double *res = diag + (size_t)DIAG_BLOCKS*DIAG_MAXQ;
    if (fp64) launch_diagnostics<double>(d, (const double *)V, vs[cur], diag, res, env.stream);
    else      launch_diagnostics<float>(d, (const float *)V, vs[cur], diag, res, env.stream);
    if (hipMemcpyAsync(out, res, 18*sizeof(double), hipMemcpyDeviceToHost, env.stream) != hipSuccess ||
        hipStreamSynchronize(env.stream) != hipSuccess)
      return GPL_EHIP;
    return 0;
  }

  int set_c2p_network(const float *w) override
  {
    if (!rmhd) return GPL_EUNSUPPORTED;
    if (!w) { net_on = false; return 0; }
    if (!net && dev_alloc(bytes, (void **)&net, sizeof(NetPack))) return GPL_ENOMEM;
    if (!net_w && dev_alloc(bytes, (void **)&net_w, NET_SIZE*sizeof(float))) return GPL_ENOMEM;
    if (net_pack_upload(w, net_w, net, env.stream) ||
        hipStreamSynchronize(env.stream) != hipSuccess)
      return GPL_EHIP;
    net_on = true;
    return 0;
  }
};

}  // namespace

int make_mhd3d_engine(const gpl_config &cfg, const EngineEnv &env, Engine **out)
{
  *out = nullptr;
  if (cfg.dimensions != 3) return GPL_EUNSUPPORTED;
  // UCT_HLL, or UCT_HLLD with the MHD HLLD solver (the reference's HLL and
  // HLLC quit with it: Src/MHD/hll.cpp:75-76, Src/MHD/hllc.cpp:81-83)
  if (cfg.ct_emf_average != GPL_UCT_HLL && cfg.ct_emf_average != GPL_UCT_HLLD) return GPL_EUNSUPPORTED;
  if (cfg.ct_emf_average == GPL_UCT_HLLD && (cfg.physics != GPL_MHD || cfg.solver != GPL_SOLVER_HLLD))
    return GPL_EUNSUPPORTED;
  // periodic, outflow or reflective sides (the last two: the padded box); a
  // direction is periodic on both sides or on neither
  for (int d = 0; d < 3; d++) {
    const int lo = side_type(cfg, 2*d), hi = side_type(cfg, 2*d + 1);
    if (lo < GPL_PERIODIC || lo > GPL_REFLECTIVE || hi < GPL_PERIODIC || hi > GPL_REFLECTIVE)
      return GPL_EUNSUPPORTED;
    if ((lo == GPL_PERIODIC) != (hi == GPL_PERIODIC)) return GPL_EINVAL;
  }
  const int pad = all_periodic(cfg) ? 0 : 5;
  if (cfg.solver != GPL_SOLVER_HLLD && cfg.solver != GPL_SOLVER_HLLC && cfg.solver != GPL_SOLVER_HLL)
    return GPL_EUNSUPPORTED;
  if (cfg.precision != GPL_FP64 && cfg.precision != GPL_MIXED) return GPL_EINVAL;
  // RMHD: HLL, HLLC or HLLD, FP64 and the fused kernels (rows of at most 512 zones)
  if (cfg.physics == GPL_RMHD &&
      (cfg.precision != GPL_FP64 ||
       cfg.nx + pad > 512 ||
       getenv("GPL_SPLIT_KERNELS") != nullptr))
    return GPL_EUNSUPPORTED;
  // the PLM stencil wraps at most two cells: every direction needs >= 3 zones
  if (cfg.nx < 3 || cfg.ny < 3 || cfg.nz < 3) return GPL_EINVAL;
  // the kernels address every array slot with 32-bit byte offsets (< 4 GiB)
  if ((long long)(cfg.nx + pad)*(cfg.ny + pad)*(cfg.nz + pad) >= (1LL << 29)) return GPL_EINVAL;
  Mhd3dEngine *e = new (std::nothrow) Mhd3dEngine();
  if (!e) return GPL_ENOMEM;
  const int rc = c_guard(GPL_ENOMEM, [&] { return e->init(cfg, env); });
  if (rc) { delete e; return rc; }
  *out = e;
  return 0;
}

}  // namespace gpl
