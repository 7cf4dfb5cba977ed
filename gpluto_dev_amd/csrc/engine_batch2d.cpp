// Batched 2D ideal MHD: many independent boxes of uneven sizes (periodic, or
// stored with the reference's ghost layers when a side is outflow or
// reflective) in one device state, advanced together (one launch per phase
// over each group of boxes, the groups on streams of their own, one HIP graph
// per step), each box with its own dt.  Launch shapes (the
// y-march length and the persistent grid size) are chosen by an autotuner
// that times the candidates on the run's own first steps: every candidate
// computes the same bits, so tuning never changes a result.
#include "../../include/gpluto_amd.h"
#include "c_guard.hpp"
#include "batch2d.hpp"
#include "diagnostics.hpp"
#include "engine.hpp"
#include "physics.hpp"

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

using namespace gpl;

namespace {

constexpr int kMarch[3] = {8, 16, 32};
constexpr int kBpc[4] = {2, 4, 8, 0};      // workgroups per CU; 0 = one per work group

template <typename T>
int upload_vec(T **dst, const std::vector<T> &v, size_t &bytes)
{
  if (dev_alloc(bytes, (void **)dst, (v.empty() ? 1 : v.size())*sizeof(T))) return GPL_ENOMEM;
  if (!v.empty() && hipMemcpy(*dst, v.data(), v.size()*sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
    return GPL_EHIP;
  return 0;
}

}  // namespace

struct gpl_batch {
  gpl_config cfg;
  int dev = 0;
  hipStream_t stream = nullptr;
  bool fp64 = true;
  int nbox = 0, ncell = 0, ncu = 0;
  int ksolver = 0;      // kernel solver code: the Riemann solver, 4 = HLLD with UCT_HLLD
  // sides X1-beg, X1-end, X2-beg, X2-end; with an outflow or reflective side
  // every box is stored padded (active + 5 per direction, active cells from 2)
  int side[4] = {GPL_PERIODIC, GPL_PERIODIC, GPL_PERIODIC, GPL_PERIODIC};
  bool padded = false;
  int maxp = 0;         // largest stored box side
  // a batch behind a gpl_ctx (one box) borrows the context's stream,
  // step-control block and error bits instead of owning them
  bool owns_env = true;
  std::vector<BoxDesc> hbox;
  size_t bytes = 0;
  // device state
  void *V = nullptr, *FD = nullptr, *E = nullptr, *dU = nullptr, *Cdt = nullptr;
  // U holds U^n between steps; stage 1 writes its state to Us, later stages
  // read Us and U^n and the last writes U.  The faces rotate over vs[0..2]:
  // a step starts from vs[cur] (see enqueue_step).
  double *U = nullptr, *Us = nullptr, *vs[3] = {nullptr, nullptr, nullptr};
  int cur = 0;
  double *cdt_part = nullptr;
  uint32_t *err = nullptr;
  StepCtl *ctl = nullptr;
  // zone-by-zone reports (a batch behind a gpl_ctx only): [stage][ncell] and
  // the counters [stage 0..2, total], owned by the engine; null while off
  uint32_t *zerr = nullptr;
  unsigned long long *zcnt = nullptr;
  // work lists
  BoxDesc *box = nullptr;
  Item4 *xitems = nullptr;
  Item2 *tiles = nullptr;
  int nx_items = 0, ntiles = 0;
  Item4 *yitems[3] = {nullptr, nullptr, nullptr};   // per march length
  int *part_beg[3] = {nullptr, nullptr, nullptr};
  int ny_items[3] = {0, 0, 0};
  // Box groups on streams of their own: group g holds boxes [gbox[g],
  // gbox[g+1]) with about the same number of cells.  The groups share no
  // cell, so their phases run side by side and fill each other's
  // latency-bound stretches and tails.  Group g's items are [beg[g],
  // beg[g+1]) of each work list; group 0 runs on the batch's stream.
  static constexpr int kMaxGroups = 4;
  int ng = 1;
  int gbox[kMaxGroups + 1] = {};
  hipStream_t gs[kMaxGroups] = {};
  hipEvent_t ev_fork = nullptr, ev_join[kMaxGroups] = {};
  int xbeg[kMaxGroups + 1] = {}, tbeg[kMaxGroups + 1] = {}, ybeg[3][kMaxGroups + 1] = {};
  // launch shape
  int march_idx = 1, bpc_idx = 2;
  bool tune_pending = true;
  double tuned_ms = 0.0;
  // graph
  hipGraphExec_t graph_exec = nullptr;
  const void *graph_key = nullptr;
  int graph_shape = -1;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_run_ms = 0.0;
};

static void batch_free(gpl_batch *b)
{
  void *ptrs[] = {b->V, b->FD, b->E, b->dU, b->Cdt, b->U, b->Us, b->vs[0], b->vs[1], b->vs[2],
                  b->cdt_part, b->box, b->xitems, b->tiles,
                  b->yitems[0], b->yitems[1], b->yitems[2], b->part_beg[0], b->part_beg[1],
                  b->part_beg[2]};
  for (void *p : ptrs) if (p) (void)hipFree(p);
  if (b->graph_exec) hipGraphExecDestroy(b->graph_exec);
  if (b->ev_fork) hipEventDestroy(b->ev_fork);
  for (int g = 1; g < gpl_batch::kMaxGroups; g++) {
    if (b->ev_join[g]) hipEventDestroy(b->ev_join[g]);
    if (b->gs[g]) hipStreamDestroy(b->gs[g]);
  }
  if (b->ev0) hipEventDestroy(b->ev0);
  if (b->ev1) hipEventDestroy(b->ev1);
  if (b->owns_env) {
    if (b->err) hipFree(b->err);
    if (b->ctl) hipFree(b->ctl);
    if (b->stream) hipStreamDestroy(b->stream);
  }
}

// Everything gpl_batch_create sets up after the validation; on failure the
// caller frees what was allocated (batch_free takes a partly built batch).
// env: the stream, step-control block and error bits of a gpl_ctx to borrow
// (one box), or null for a batch of its own.
static int batch_init(gpl_batch *b, const gpl_config *cfg, int nbox, const int *nx, const int *ny,
                      long long total, const EngineEnv *env)
{
  b->cfg = *cfg;
  b->dev = cfg->device;
  b->fp64 = cfg->precision == GPL_FP64;
  b->nbox = nbox;
  b->ncell = (int)total;
  b->ksolver = cfg->ct_emf_average == GPL_UCT_HLLD ? 4 : cfg->solver;
  if (env) {
    b->owns_env = false;
    b->stream = env->stream;
    b->ctl = env->ctl;
    b->err = env->err;
  }
  if (hipSetDevice(b->dev) != hipSuccess ||
      (!env && hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) ||
      hipDeviceGetAttribute(&b->ncu, hipDeviceAttributeMultiprocessorCount, b->dev) != hipSuccess) {
    return GPL_EHIP;
  }
  hipEventCreate(&b->ev0);
  hipEventCreate(&b->ev1);
  // boxes and work lists
  std::vector<Item4> xi;
  std::vector<Item2> ti;
  std::vector<int> xstart, tstart;   // first item of every box
  int off = 0;
  for (int q = 0; q < 4; q++) b->side[q] = side_type(*cfg, q);
  b->padded = !all_periodic(*cfg);
  const int pad = b->padded ? 5 : 0, lo = b->padded ? 2 : 0;
  for (int q = 0; q < nbox; q++) {
    BoxDesc d;
    d.nx = nx[q] + pad; d.ny = ny[q] + pad; d.off = off;
    d.dx = (cfg->xend[0] - cfg->xbeg[0])/nx[q];   // Src/set_grid.cpp:196
    d.dy = (cfg->xend[1] - cfg->xbeg[1])/ny[q];
    d.lo[0] = d.lo[1] = lo;
    d.hi[0] = lo + nx[q] - 1;
    d.hi[1] = lo + ny[q] - 1;
    b->hbox.push_back(d);
    if (d.nx > b->maxp) b->maxp = d.nx;
    if (d.ny > b->maxp) b->maxp = d.ny;
    xstart.push_back((int)xi.size());
    tstart.push_back((int)ti.size());
    for (int j = 0; j < d.ny; j++)
      for (int i0 = 0; i0 < d.nx; i0 += 63) xi.push_back({q, j, i0, d.nx - i0 < 63 ? d.nx - i0 : 63});
    for (int s = 0; s < d.nx*d.ny; s += 256) ti.push_back({q, s});
    off += d.nx*d.ny;
  }
  const size_t nc = (size_t)b->ncell, rs = b->fp64 ? sizeof(double) : sizeof(float);
  int rc = 0;
  rc |= upload_vec(&b->box, b->hbox, b->bytes);
  rc |= upload_vec(&b->xitems, xi, b->bytes);
  rc |= upload_vec(&b->tiles, ti, b->bytes);
  b->nx_items = (int)xi.size();
  b->ntiles = (int)ti.size();
  int max_y = 0;
  std::vector<int> yb[3];
  for (int m = 0; m < 3; m++) {
    const int L = kMarch[m];
    std::vector<Item4> yi;
    std::vector<int> pb;
    for (int q = 0; q < nbox; q++) {
      pb.push_back((int)yi.size());
      const int sx = b->hbox[q].nx, sy = b->hbox[q].ny;
      const int nseg = (sy + L - 1)/L;
      for (int s = 0; s < nseg; s++)
        for (int i0 = 0; i0 < sx; i0 += 64) yi.push_back({q, i0, s, 0});
    }
    pb.push_back((int)yi.size());
    yb[m] = pb;
    rc |= upload_vec(&b->yitems[m], yi, b->bytes);
    rc |= upload_vec(&b->part_beg[m], pb, b->bytes);
    b->ny_items[m] = (int)yi.size();
    if ((int)yi.size() > max_y) max_y = (int)yi.size();
  }
  // the box groups (a batch of its own: GPL_BATCH_STREAMS groups, 2 by
  // default, at most one per box): group g starts at the first box whose
  // middle lies past g/ng of the cells
  {
    int ng = 2;
    if (const char *e = getenv("GPL_BATCH_STREAMS")) ng = atoi(e);
    if (ng < 1) ng = 1;
    if (ng > gpl_batch::kMaxGroups) ng = gpl_batch::kMaxGroups;
    if (ng > nbox) ng = nbox;
    if (env) ng = 1;
    b->ng = ng;
    b->gbox[0] = 0;
    long long acc = 0;
    int q = 0;
    for (int g = 1; g < ng; g++) {
      // at least one box per group, and room for one per later group
      while (q < nbox - (ng - g) &&
             (q < b->gbox[g - 1] + 1 ||
              (long long)ng*(2*acc + (long long)b->hbox[q].nx*b->hbox[q].ny) <= 2LL*g*total)) {
        acc += (long long)b->hbox[q].nx*b->hbox[q].ny;
        q++;
      }
      b->gbox[g] = q;
    }
    b->gbox[ng] = nbox;
    xstart.push_back(b->nx_items);
    tstart.push_back(b->ntiles);
    for (int g = 0; g <= ng; g++) {
      b->xbeg[g] = xstart[b->gbox[g]];
      b->tbeg[g] = tstart[b->gbox[g]];
      for (int m = 0; m < 3; m++) b->ybeg[m][g] = yb[m][b->gbox[g]];
    }
    b->gs[0] = b->stream;
    if (ng > 1 && hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) != hipSuccess) return GPL_EHIP;
    for (int g = 1; g < ng; g++)
      if (hipStreamCreateWithFlags(&b->gs[g], hipStreamNonBlocking) != hipSuccess ||
          hipEventCreateWithFlags(&b->ev_join[g], hipEventDisableTiming) != hipSuccess)
        return GPL_EHIP;
  }
  rc |= dev_alloc(b->bytes, &b->V, 8*nc*rs);
  rc |= dev_alloc(b->bytes, (void **)&b->U, 6*nc*sizeof(double));
  rc |= dev_alloc(b->bytes, (void **)&b->Us, 6*nc*sizeof(double));
  for (int q = 0; q < 3; q++) rc |= dev_alloc(b->bytes, (void **)&b->vs[q], 2*nc*sizeof(double));
  rc |= dev_alloc(b->bytes, &b->FD, 2*(b->ksolver == 4 ? 6 : 4)*nc*rs);
  rc |= dev_alloc(b->bytes, &b->E, nc*rs);
  rc |= dev_alloc(b->bytes, &b->dU, 6*nc*rs);
  rc |= dev_alloc(b->bytes, &b->Cdt, nc*rs);
  rc |= dev_alloc(b->bytes, (void **)&b->cdt_part, (size_t)max_y*sizeof(double));
  if (!env) {
    rc |= dev_alloc(b->bytes, (void **)&b->err, sizeof(uint32_t));
    rc |= dev_alloc(b->bytes, (void **)&b->ctl, (size_t)nbox*sizeof(StepCtl));
  }
  if (rc) return rc == GPL_EHIP ? GPL_EHIP : GPL_ENOMEM;
  std::vector<StepCtl> z(nbox, StepCtl{0.0, 0.0, 0.0, 0});
  if (hipMemcpy(b->ctl, z.data(), nbox*sizeof(StepCtl), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(b->err, 0, sizeof(uint32_t)) != hipSuccess)
    return GPL_EHIP;
  b->cur = 0;
  return 0;
}

// The configurations the batched kernels take; *total: stored cells.
static int batch_validate(const gpl_config *cfg, int nbox, const int *nx, const int *ny, long long *total)
{
  if (cfg->physics != GPL_MHD || cfg->dimensions != 2) return GPL_EUNSUPPORTED;
  // periodic, outflow or reflective sides (outflow and reflective: padded
  // boxes); a direction is periodic on both sides or on neither
  for (int d = 0; d < 2; d++) {
    const int lo = side_type(*cfg, 2*d), hi = side_type(*cfg, 2*d + 1);
    if (lo < GPL_PERIODIC || lo > GPL_REFLECTIVE || hi < GPL_PERIODIC || hi > GPL_REFLECTIVE)
      return GPL_EUNSUPPORTED;
    if ((lo == GPL_PERIODIC) != (hi == GPL_PERIODIC)) return GPL_EINVAL;
  }
  const int pad = all_periodic(*cfg) ? 0 : 5;
  // UCT_HLL, or UCT_HLLD with the HLLD solver (the reference's HLL and HLLC
  // quit with it: Src/MHD/hll.cpp:75-76, Src/MHD/hllc.cpp:81-83)
  if (cfg->ct_emf_average != GPL_UCT_HLL && cfg->ct_emf_average != GPL_UCT_HLLD) return GPL_EUNSUPPORTED;
  if (cfg->ct_emf_average == GPL_UCT_HLLD && cfg->solver != GPL_SOLVER_HLLD) return GPL_EUNSUPPORTED;
  if (cfg->solver != GPL_SOLVER_HLLD && cfg->solver != GPL_SOLVER_HLLC && cfg->solver != GPL_SOLVER_HLL)
    return GPL_EUNSUPPORTED;
  if (cfg->time_stepping != GPL_RK2 && cfg->time_stepping != GPL_RK3) return GPL_EINVAL;
  if (cfg->limiter != GPL_MC_LIM && cfg->limiter != GPL_VANLEER_LIM) return GPL_EINVAL;
  if (cfg->precision != GPL_FP64 && cfg->precision != GPL_MIXED) return GPL_EINVAL;
  *total = 0;
  for (int q = 0; q < nbox; q++) {
    if (nx[q] < 3 || nx[q] > 8192 || ny[q] < 3 || ny[q] > 8192) return GPL_EINVAL;
    *total += (long long)(nx[q] + pad)*(ny[q] + pad);
  }
  if (*total > (1LL << 30)) return GPL_EINVAL;
  return 0;
}

extern "C" int gpl_batch_create(const gpl_config *cfg, int nbox, const int *nx, const int *ny,
                                gpl_batch **out)
{
  if (!cfg || !nx || !out || nbox < 1) return GPL_EINVAL;
  if (!ny) ny = nx;
  *out = nullptr;
  long long total = 0;
  if (int rc = batch_validate(cfg, nbox, nx, ny, &total)) return rc;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || cfg->device < 0 || cfg->device >= ndev) return GPL_EHIP;

  gpl_batch *b = new (std::nothrow) gpl_batch();
  if (!b) return GPL_ENOMEM;
  const int rc = c_guard(GPL_ENOMEM, [&] { return batch_init(b, cfg, nbox, nx, ny, total, nullptr); });
  if (rc) {
    batch_free(b);
    delete b;
    return rc;
  }
  if (const char *e = getenv("GPL_BATCH_SHAPE")) {   // "march,bpc" fixes the launch shape
    int m = 16, p = 8;
    if (sscanf(e, "%d,%d", &m, &p) == 2) gpl_batch_set_tuning(b, m, p);
  }
  *out = b;
  return 0;
}

extern "C" int gpl_batch_destroy(gpl_batch *b)
{
  if (!b) return GPL_EINVAL;
  hipSetDevice(b->dev);
  hipStreamSynchronize(b->stream);
  batch_free(b);
  delete b;
  return 0;
}

static int active_n(const BoxDesc &d, int dir) { return d.hi[dir] - d.lo[dir] + 1; }

// the caller's layout of one box's active zones
static Layout box_layout(const gpl_batch *b, int box, int layout)
{
  gpl_config c = b->cfg;
  c.nx = active_n(b->hbox[box], 0);
  c.ny = active_n(b->hbox[box], 1);
  c.nz = 1;
  c.dimensions = 2;
  return make_layout(c, layout);
}

// Host index along dir that stored cell index p reads at upload: the active
// cell, or in a ghost layer its periodic image or the clamped edge cell (the
// boundary fill rewrites the ghosts before every stage).
static int src_cell(const gpl_batch *b, const BoxDesc &d, int dir, int p)
{
  const int a = p - d.lo[dir], m = active_n(d, dir);
  if (b->side[2*dir] == GPL_PERIODIC) return ((a % m) + m) % m;
  return a < 0 ? 0 : (a >= m ? m - 1 : a);
}

// Host index along dir of the face that stored face (component comp, index p
// along dir) takes at upload: the periodic image, except the upper boundary
// face of a non-periodic direction, which the reference's array layout
// carries (face n); with the active layout it starts as the image of face 0.
static int src_face(const gpl_batch *b, const BoxDesc &d, const Layout &L, int comp, int dir, int p)
{
  const int a = p - d.lo[dir], m = active_n(d, dir);
  if (L.ary && comp == dir && b->side[2*dir] != GPL_PERIODIC && a == m) return m;
  return ((a % m) + m) % m;
}

static int gpl_batch_upload_impl(gpl_batch *b, int box, const double *vc, const double *const *vs,
                                 int layout)
{
  if (!b || box < 0 || box >= b->nbox || !vc || !vs || !vs[0] || !vs[1]) return GPL_EINVAL;
  if (layout != GPL_LAYOUT_ACTIVE && layout != GPL_LAYOUT_ARY4D) return GPL_EINVAL;
  hipSetDevice(b->dev);
  if (hipStreamSynchronize(b->stream) != hipSuccess) return GPL_EHIP;
  const Layout L = box_layout(b, box, layout);
  const BoxDesc d = b->hbox[box];
  const int NX = d.nx, NY = d.ny;
  const size_t m = (size_t)NX*NY, nc = (size_t)b->ncell;
  std::vector<double> hv(8*m), hu(6*m), hs(2*m);
  for (int c = 0; c < 2; c++)
    for (int j = 0; j < NY; j++)
      for (int i = 0; i < NX; i++)
        hs[c*m + (size_t)j*NX + i] = vs[c][L.face(c, src_face(b, d, L, c, 0, i), src_face(b, d, L, c, 1, j), 0)];
  for (int j = 0; j < NY; j++)
    for (int i = 0; i < NX; i++) {
      const size_t o = (size_t)j*NX + i;
      double v[8], u[8];
      for (int s = 0; s < 8; s++) v[s] = vc[L.cell(s, src_cell(b, d, 0, i), src_cell(b, d, 1, j), 0)];
      // cell-centred B1, B2 = face averages (Src/startup.cpp:388, ct_field_average.cpp:182-184)
      const int ip = i + 1 == NX ? 0 : i + 1, jp = j + 1 == NY ? 0 : j + 1;
      v[4] = 0.5*(hs[o] + hs[(size_t)j*NX + ip]);
      v[5] = 0.5*(hs[m + o] + hs[m + (size_t)jp*NX + i]);
      mhd_prim_to_cons<double>(v, u, b->cfg.gamma - 1.0);   // Src/startup.cpp:514-561
      for (int s = 0; s < 8; s++) hv[s*m + o] = v[s];
      const int us[6] = {0, 1, 2, 3, 6, 7};
      for (int s = 0; s < 6; s++) hu[s*m + o] = u[us[s]];
    }
  const size_t rs = b->fp64 ? sizeof(double) : sizeof(float);
  for (int s = 0; s < 8; s++) {
    char *dst = (char *)b->V + (s*nc + d.off)*rs;
    if (b->fp64) {
      if (hipMemcpy(dst, &hv[s*m], m*rs, hipMemcpyHostToDevice) != hipSuccess) return GPL_EHIP;
    } else {
      std::vector<float> t(m);
      for (size_t q = 0; q < m; q++) t[q] = (float)hv[s*m + q];
      if (hipMemcpy(dst, t.data(), m*rs, hipMemcpyHostToDevice) != hipSuccess) return GPL_EHIP;
    }
  }
  for (int s = 0; s < 6; s++)
    if (hipMemcpy(b->U + s*nc + d.off, &hu[s*m], m*sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      return GPL_EHIP;
  for (int c = 0; c < 2; c++)
    if (hipMemcpy(b->vs[b->cur] + c*nc + d.off, &hs[c*m], m*sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      return GPL_EHIP;
  const StepCtl z = {0.0, 0.0, 0.0, 0};
  if (hipMemcpy(b->ctl + box, &z, sizeof(z), hipMemcpyHostToDevice) != hipSuccess) return GPL_EHIP;
  return 0;
}

extern "C" int gpl_batch_upload(gpl_batch *b, int box, const double *vc, const double *const *vs,
                                int layout)
{
  return c_guard<int>(GPL_ENOMEM, [&] { return gpl_batch_upload_impl(b, box, vc, vs, layout); });
}

static int gpl_batch_download_impl(gpl_batch *b, int box, double *vc, double *const *vs, double *uc,
                                   int layout)
{
  if (!b || box < 0 || box >= b->nbox) return GPL_EINVAL;
  if (layout != GPL_LAYOUT_ACTIVE && layout != GPL_LAYOUT_ARY4D) return GPL_EINVAL;
  hipSetDevice(b->dev);
  if (hipStreamSynchronize(b->stream) != hipSuccess) return GPL_EHIP;
  const Layout L = box_layout(b, box, layout);
  const BoxDesc d = b->hbox[box];
  const int NX = d.nx, NY = d.ny;
  const size_t m = (size_t)NX*NY, nc = (size_t)b->ncell;
  const size_t rs = b->fp64 ? sizeof(double) : sizeof(float);
  std::vector<double> hv(8*m), hu(6*m), hs(2*m);
  for (int s = 0; s < 8; s++) {
    const char *src = (const char *)b->V + (s*nc + d.off)*rs;
    if (b->fp64) {
      if (hipMemcpy(&hv[s*m], src, m*rs, hipMemcpyDeviceToHost) != hipSuccess) return GPL_EHIP;
    } else {
      std::vector<float> t(m);
      if (hipMemcpy(t.data(), src, m*rs, hipMemcpyDeviceToHost) != hipSuccess) return GPL_EHIP;
      for (size_t q = 0; q < m; q++) hv[s*m + q] = t[q];
    }
  }
  for (int s = 0; s < 6; s++)
    if (hipMemcpy(&hu[s*m], b->U + s*nc + d.off, m*sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
      return GPL_EHIP;
  for (int c = 0; c < 2; c++)
    if (hipMemcpy(&hs[c*m], b->vs[b->cur] + c*nc + d.off, m*sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
      return GPL_EHIP;
  // the active zones of the stored box
  for (int j = 0; j < L.ny; j++)
    for (int i = 0; i < L.nx; i++) {
      const size_t o = (size_t)(j + d.lo[1])*NX + (i + d.lo[0]);
      if (vc) for (int s = 0; s < 8; s++) vc[L.cell(s, i, j, 0)] = hv[s*m + o];
      if (uc) {
        // B1, B2 of Uc are the face averages (= Vc's)
        const double u8[8] = {hu[o], hu[m + o], hu[2*m + o], hu[3*m + o], hv[4*m + o], hv[5*m + o],
                              hu[4*m + o], hu[5*m + o]};
        for (int s = 0; s < 8; s++) uc[L.cell(s, i, j, 0)] = u8[s];
      }
      if (vs) for (int c = 0; c < 2; c++) if (vs[c]) vs[c][L.face(c, i, j, 0)] = hs[c*m + o];
    }
  if (vs && L.ary) {
    // the upper boundary face: the periodic image of the lower one
    // (PeriodicBoundary), or its own value on an outflow or reflective side
    const bool px = b->side[0] == GPL_PERIODIC, py = b->side[2] == GPL_PERIODIC;
    for (int j = 0; j < L.ny; j++)
      if (vs[0]) vs[0][L.face(0, L.nx, j, 0)] = hs[(size_t)(j + d.lo[1])*NX + d.lo[0] + (px ? 0 : L.nx)];
    for (int i = 0; i < L.nx; i++)
      if (vs[1]) vs[1][L.face(1, i, L.ny, 0)] = hs[m + (size_t)(d.lo[1] + (py ? 0 : L.ny))*NX + d.lo[0] + i];
  }
  return 0;
}

extern "C" int gpl_batch_download(gpl_batch *b, int box, double *vc, double *const *vs, double *uc,
                                  int layout)
{
  return c_guard<int>(GPL_ENOMEM, [&] { return gpl_batch_download_impl(b, box, vc, vs, uc, layout); });
}

// ---------------------------------------------------------------------------
// stepping
// ---------------------------------------------------------------------------
static int grid_for(const gpl_batch *b, int groups)
{
  const int bpc = kBpc[b->bpc_idx];
  if (bpc == 0 || groups <= bpc*b->ncu) return groups > 0 ? groups : 1;
  return bpc*b->ncu;
}

template <typename R>
static void enqueue_stage(gpl_batch *b, const double *uin, double *uout, double *vin,
                          double *vout, const double *vbase, RKW w, int first_stage, int stage_idx)
{
  Batch2dBufs<R> q;
  q.zr = b->zerr ? ZoneRep{b->zerr + (size_t)stage_idx*b->ncell, b->zcnt + stage_idx, b->zcnt + 3}
                 : ZoneRep{nullptr, nullptr, nullptr};
  q.V = (R *)b->V; q.Uin = uin; q.U = uout; q.U0 = b->U;
  q.Vin = vin; q.Vout = vout; q.Bs0 = vbase;
  q.FD = (R *)b->FD; q.E = (R *)b->E; q.dU = (R *)b->dU; q.Cdt = (R *)b->Cdt;
  q.err = b->err; q.ctl = b->ctl; q.box = b->box;
  q.ncell = b->ncell;
  const int m = b->march_idx, march = kMarch[m];
  const gpl_config &c = b->cfg;
  // the box groups, each on its own stream; the kernels index the items of
  // their group and the C_dt partial slots that go with them
  for (int gi = 0; gi < b->ng; gi++) {
    const hipStream_t st = b->gs[gi];
    const int box0 = b->gbox[gi], nbg = b->gbox[gi + 1] - b->gbox[gi];
    q.xitems = b->xitems + b->xbeg[gi]; q.nx_items = b->xbeg[gi + 1] - b->xbeg[gi];
    q.yitems = b->yitems[m] + b->ybeg[m][gi]; q.ny_items = b->ybeg[m][gi + 1] - b->ybeg[m][gi];
    q.cdt_part = b->cdt_part + b->ybeg[m][gi];
    q.tiles = b->tiles + b->tbeg[gi]; q.ntiles = b->tbeg[gi + 1] - b->tbeg[gi];
    const int gx = grid_for(b, (q.nx_items + 3)/4), gy = grid_for(b, (q.ny_items + 3)/4);
    const int gt = grid_for(b, q.ntiles);
    // Boundary (Src/rk_step.cpp: at the start of every stage) on padded boxes
    if (b->padded)
      batch2d_boundary<R>(b->box + box0, nbg, b->maxp, b->side, (R *)b->V, vin, (size_t)b->ncell, st);
    batch2d_launch<R>(B2_SWEEP_X, q, w, first_stage, first_stage, c.limiter, b->ksolver, march, gx,
                      c.gamma, c.small_rho, c.small_prs, st);
    batch2d_launch<R>(B2_SWEEP_Y, q, w, first_stage, first_stage, c.limiter, b->ksolver, march, gy,
                      c.gamma, c.small_rho, c.small_prs, st);
    batch2d_launch<R>(B2_EMF, q, w, first_stage, first_stage, c.limiter, b->ksolver, march, gt,
                      c.gamma, c.small_rho, c.small_prs, st);
    batch2d_launch<R>(B2_FINISH, q, w, first_stage, first_stage, c.limiter, b->ksolver, march, gt,
                      c.gamma, c.small_rho, c.small_prs, st);
  }
}

// one AdvanceStep of every box (Src/rk_step.cpp:88-720), then per-box dt
static void enqueue_stage_r(gpl_batch *b, const double *uin, double *uout, double *vin,
                            double *vout, const double *vbase, RKW w, int first, int stage_idx)
{
  if (b->fp64) enqueue_stage<double>(b, uin, uout, vin, vout, vbase, w, first, stage_idx);
  else         enqueue_stage<float>(b, uin, uout, vin, vout, vbase, w, first, stage_idx);
}

static void enqueue_step(gpl_batch *b, int advance_dt)
{
  const RKW w1 = {0.0, 0.0, 1.0};
  const RKW w2 = b->cfg.time_stepping == GPL_RK3 ? RKW{0.75, 0.25, 0.25} : RKW{0.5, 0.5, 0.5};
  const RKW w3 = {1.0/3.0, 2.0/3.0, 2.0/3.0};
  double *f0 = b->vs[b->cur], *f1 = b->vs[(b->cur + 1) % 3], *f2 = b->vs[(b->cur + 2) % 3];
  if (b->zcnt) launch_zone_reset(b->zcnt, b->stream);
  if (b->ng > 1) {                     // the other groups' streams join here
    (void)hipEventRecord(b->ev_fork, b->stream);
    for (int g = 1; g < b->ng; g++) (void)hipStreamWaitEvent(b->gs[g], b->ev_fork, 0);
  }
  enqueue_stage_r(b, b->U, b->Us, f0, f1, f0, w1, 1, 0);            // -> Us, f1
  if (b->cfg.time_stepping != GPL_RK3) {
    enqueue_stage_r(b, b->Us, b->U, f1, f2, f0, w2, 0, 1);          // -> U, f2
    b->cur = (b->cur + 2) % 3;
  } else {
    enqueue_stage_r(b, b->Us, b->Us, f1, f2, f0, w2, 0, 1);         // -> Us, f2
    enqueue_stage_r(b, b->Us, b->U, f2, f1, f0, w3, 0, 2);          // -> U, f1
    b->cur = (b->cur + 1) % 3;
  }
  // per group: each box's dt reads only its own partial slots
  for (int g = 0; g < b->ng; g++)
    batch2d_launch_dt(b->cdt_part, b->part_beg[b->march_idx] + b->gbox[g], b->gbox[g + 1] - b->gbox[g],
                      b->ctl + b->gbox[g], b->cfg.cfl, b->cfg.cfl_max_var, advance_dt, b->gs[g]);
  for (int g = 1; g < b->ng; g++) {
    (void)hipEventRecord(b->ev_join[g], b->gs[g]);
    (void)hipStreamWaitEvent(b->stream, b->ev_join[g], 0);
  }
}

// the face buffers rotate with period 3 (RK2: cur += 2, RK3: cur += 1)
static int graph_period(const gpl_batch *) { return 3; }

static int build_graph(gpl_batch *b)
{
  const int shape = b->march_idx*8 + b->bpc_idx;
  if (b->graph_exec && b->graph_key == b->vs[b->cur] && b->graph_shape == shape) return 0;
  if (b->graph_exec) { hipGraphExecDestroy(b->graph_exec); b->graph_exec = nullptr; }
  const void *key = b->vs[b->cur];
  hipGraph_t graph;
  if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return GPL_EHIP;
  for (int s = 0; s < graph_period(b); s++) enqueue_step(b, 1);
  if (hipStreamEndCapture(b->stream, &graph) != hipSuccess) return GPL_EHIP;
  if (b->vs[b->cur] != key) { hipGraphDestroy(graph); return GPL_EHIP; }
  const hipError_t e = hipGraphInstantiate(&b->graph_exec, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  if (e != hipSuccess) { b->graph_exec = nullptr; return GPL_EHIP; }
  b->graph_key = key;
  b->graph_shape = shape;
  return 0;
}

// Time every launch shape on one step each and keep the fastest.  Each
// trial step is captured as a graph of its own and launched once, so it is
// timed as the replayed steps run, without the host's launch gaps; it is a
// step of the run like any other.
static int autotune(gpl_batch *b, int *done)
{
  double best = 1e30;
  int bm = b->march_idx, bp = b->bpc_idx;
  for (int m = 0; m < 3; m++)
    for (int p = 0; p < 4; p++) {
      b->march_idx = m; b->bpc_idx = p;
      hipGraph_t graph;
      hipGraphExec_t exec = nullptr;
      if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return GPL_EHIP;
      enqueue_step(b, 1);
      if (hipStreamEndCapture(b->stream, &graph) != hipSuccess) return GPL_EHIP;
      const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      hipGraphDestroy(graph);
      if (e != hipSuccess) return GPL_EHIP;
      hipEventRecord(b->ev0, b->stream);
      const hipError_t l = hipGraphLaunch(exec, b->stream);
      hipEventRecord(b->ev1, b->stream);
      const hipError_t sy = hipEventSynchronize(b->ev1);
      hipGraphExecDestroy(exec);
      if (l != hipSuccess || sy != hipSuccess) return GPL_EHIP;
      float ms = 0.f;
      hipEventElapsedTime(&ms, b->ev0, b->ev1);
      if (ms < best) { best = ms; bm = m; bp = p; }
      (*done)++;
    }
  b->march_idx = bm; b->bpc_idx = bp;
  b->tuned_ms = best;
  b->tune_pending = false;
  return 0;
}

static constexpr int kTuneSteps = 12;

static uint32_t gpl_batch_run_impl(gpl_batch *b, int nsteps, const double *first_dt, double *time_out,
                                   double *dt_out)
{
  if (!b || nsteps < 0) return GPL_STEP_API_ERROR;
  hipSetDevice(b->dev);
  std::vector<StepCtl> h(b->nbox);
  if (hipStreamSynchronize(b->stream) != hipSuccess ||
      hipMemcpy(h.data(), b->ctl, b->nbox*sizeof(StepCtl), hipMemcpyDeviceToHost) != hipSuccess)
    return GPL_STEP_API_ERROR;
  if (first_dt)
    for (int q = 0; q < b->nbox; q++) { h[q].dt = first_dt[q]; h[q].time = 0.0; h[q].nstep = 0; }
  const uint32_t zero = 0;
  if (hipMemcpy(b->ctl, h.data(), b->nbox*sizeof(StepCtl), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->err, &zero, sizeof(zero), hipMemcpyHostToDevice) != hipSuccess)
    return GPL_STEP_API_ERROR;
  int done = 0;
  if (b->tune_pending && nsteps >= 2*kTuneSteps && getenv("GPL_NO_GRAPH") == nullptr)
    if (autotune(b, &done) != 0) return GPL_STEP_API_ERROR;
  hipEventRecord(b->ev0, b->stream);
  const int period = graph_period(b);
  if (getenv("GPL_NO_GRAPH") == nullptr && nsteps - done >= period) {
    if (build_graph(b) != 0) return GPL_STEP_API_ERROR;
    while (nsteps - done >= period) {
      if (hipGraphLaunch(b->graph_exec, b->stream) != hipSuccess) return GPL_STEP_API_ERROR;
      done += period;
    }
  }
  while (done < nsteps) { enqueue_step(b, 1); done++; }
  hipEventRecord(b->ev1, b->stream);
  uint32_t err = 0;
  if (hipStreamSynchronize(b->stream) != hipSuccess ||
      hipMemcpy(h.data(), b->ctl, b->nbox*sizeof(StepCtl), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(&err, b->err, sizeof(err), hipMemcpyDeviceToHost) != hipSuccess)
    return GPL_STEP_API_ERROR;
  float ms = 0.f;
  hipEventElapsedTime(&ms, b->ev0, b->ev1);
  b->last_run_ms = ms;
  for (int q = 0; q < b->nbox; q++) {
    if (time_out) time_out[q] = h[q].time;
    if (dt_out) dt_out[q] = h[q].dt;
  }
  return err;
}

extern "C" uint32_t gpl_batch_run(gpl_batch *b, int nsteps, const double *first_dt, double *time_out,
                                  double *dt_out)
{
  return c_guard<uint32_t>(GPL_STEP_API_ERROR, 
                  [&] { return gpl_batch_run_impl(b, nsteps, first_dt, time_out, dt_out); });
}

static int gpl_batch_set_tuning_impl(gpl_batch *b, int march, int blocks_per_cu)
{
  if (!b) return GPL_EINVAL;
  if (march == 0) { b->tune_pending = true; return 0; }
  int m = -1, p = -1;
  for (int q = 0; q < 3; q++) if (kMarch[q] == march) m = q;
  for (int q = 0; q < 4; q++) if (kBpc[q] == blocks_per_cu) p = q;
  if (m < 0 || p < 0) return GPL_EINVAL;
  b->march_idx = m; b->bpc_idx = p;
  b->tune_pending = false;
  return 0;
}

extern "C" int gpl_batch_set_tuning(gpl_batch *b, int march, int blocks_per_cu)
{
  return c_guard<int>(GPL_ENOMEM, [&] { return gpl_batch_set_tuning_impl(b, march, blocks_per_cu); });
}

extern "C" int gpl_batch_get_tuning(const gpl_batch *b, int *march, int *blocks_per_cu, double *step_ms)
{
  if (!b) return GPL_EINVAL;
  if (march) *march = kMarch[b->march_idx];
  if (blocks_per_cu) *blocks_per_cu = kBpc[b->bpc_idx];
  if (step_ms) *step_ms = b->tuned_ms;
  return 0;
}

extern "C" double gpl_batch_last_run_ms(const gpl_batch *b) { return b ? b->last_run_ms : 0.0; }

extern "C" size_t gpl_batch_device_bytes(const gpl_batch *b) { return b ? b->bytes : 0; }

// ---------------------------------------------------------------------------
// 2D MHD behind gpl_create: one box of the batched kernels on the context's
// stream, step-control block and error bits, so gpl_run, the dbl dumps and
// restarts and the diagnostics work on a 2D domain as on a 3D one.
// ---------------------------------------------------------------------------
namespace gpl {

namespace {

class Batch2dEngine : public Engine {
public:
  gpl_batch *b = nullptr;
  double *diag = nullptr;   // diagnostics scratch: partial results, then the result

  ~Batch2dEngine() override
  {
    if (diag) hipFree(diag);
    if (b) { batch_free(b); delete b; }
  }

  int upload(const double *vc, const double *const *vs, int layout) override
  {
    return gpl_batch_upload_impl(b, 0, vc, vs, layout);
  }
  int download(double *vc, double *const *vs, double *uc, int layout) override
  {
    return gpl_batch_download_impl(b, 0, vc, vs, uc, layout);
  }
  void enqueue_step(int advance_dt, LaunchTimer *) override { ::enqueue_step(b, advance_dt); }
  int graph_period() const override { return ::graph_period(b); }
  const void *state_key() const override { return b->vs[b->cur]; }
  int nvar() const override { return 8; }
  int nstages() const override { return b->cfg.time_stepping == GPL_RK3 ? 3 : 2; }
  size_t zone_cells() const override { return (size_t)b->ncell; }
  size_t zone_active() const override
  {
    const BoxDesc &d = b->hbox[0];
    return (size_t)(d.hi[0] - d.lo[0] + 1)*(d.hi[1] - d.lo[1] + 1);
  }
  void zone_extract(const uint32_t *src, uint32_t *dst) const override
  {
    const BoxDesc &d = b->hbox[0];
    const int nx = d.hi[0] - d.lo[0] + 1, ny = d.hi[1] - d.lo[1] + 1;
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++)
        dst[(size_t)j*nx + i] = src[d.off + (size_t)(j + d.lo[1])*d.nx + (i + d.lo[0])];
  }
  void zone_changed() override
  {
    b->zerr = zerr;
    b->zcnt = zcnt;
    if (b->graph_exec) { hipGraphExecDestroy(b->graph_exec); b->graph_exec = nullptr; }
  }

  int diagnostics(double *out) override
  {
    const BoxDesc &x = b->hbox[0];
    DiagBox d;
    d.nx = x.nx; d.ny = x.ny; d.nz = 1; d.ncell = b->ncell; d.nv = 8; d.nb = 2;
    d.lo[0] = x.lo[0]; d.hi[0] = x.hi[0];
    d.lo[1] = x.lo[1]; d.hi[1] = x.hi[1];
    d.lo[2] = d.hi[2] = 0;
    // face areas of the uniform Cartesian grid in 2D (Src/set_geometry.cpp:124,147)
    d.ax = 1.0*x.dy; d.ay = x.dx*1.0; d.az = 0.0;
    double *res = diag + (size_t)DIAG_BLOCKS*DIAG_MAXQ;
    if (b->fp64) launch_diagnostics<double>(d, (const double *)b->V, b->vs[b->cur], diag, res, b->stream);
    else         launch_diagnostics<float>(d, (const float *)b->V, b->vs[b->cur], diag, res, b->stream);
    if (hipMemcpyAsync(out, res, 18*sizeof(double), hipMemcpyDeviceToHost, b->stream) != hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess)
      return GPL_EHIP;
    return 0;
  }

  int dump_gather(double *dst, hipStream_t s) override
  {
    const int per[2] = {b->side[0] == GPL_PERIODIC, b->side[2] == GPL_PERIODIC};
    if (b->fp64) batch2d_dump_gather<double>(b->hbox[0], per, (const double *)b->V, b->vs[b->cur],
                                             (size_t)b->ncell, dst, s);
    else         batch2d_dump_gather<float>(b->hbox[0], per, (const float *)b->V, b->vs[b->cur],
                                            (size_t)b->ncell, dst, s);
    return hipGetLastError() == hipSuccess ? 0 : GPL_EHIP;
  }
};

}  // namespace

int make_batch2d_engine(const gpl_config &cfg, const EngineEnv &env, Engine **out)
{
  *out = nullptr;
  if (cfg.nz != 1) return GPL_EINVAL;              // unused directions hold 1 zone
  long long total = 0;
  if (int rc = batch_validate(&cfg, 1, &cfg.nx, &cfg.ny, &total)) return rc;
  Batch2dEngine *e = new (std::nothrow) Batch2dEngine();
  if (!e) return GPL_ENOMEM;
  e->b = new (std::nothrow) gpl_batch();
  if (!e->b) { delete e; return GPL_ENOMEM; }
  int rc = c_guard(GPL_ENOMEM, [&] { return batch_init(e->b, &cfg, 1, &cfg.nx, &cfg.ny, total, &env); });
  // a fixed launch shape (the autotuner belongs to gpl_batch_run)
  e->b->tune_pending = false;
  if (!rc) rc = dev_alloc(e->b->bytes, (void **)&e->diag, (size_t)(DIAG_BLOCKS + 1)*DIAG_MAXQ*sizeof(double));
  if (rc) { delete e; return rc; }
  e->bytes = e->b->bytes;
  *out = e;
  return 0;
}

}  // namespace gpl
