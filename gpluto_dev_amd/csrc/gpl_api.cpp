// C-ABI of the MI355X stage update (include/gpluto_amd.h).
//
// A context owns one engine (3D MHD or HD) on one device, the stream all of
// its work runs on, the device-side step-control block (dt lives on the
// device, so a run of many steps replays one captured HIP graph without
// host round trips), the error bits, and the per-kernel timing.
#include "../../include/gpluto_amd.h"
#include "c_guard.hpp"
#include "dbl_io.hpp"
#include "diagnostics.hpp"
#include "engine.hpp"

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <thread>
#include <vector>

using namespace gpl;

namespace {

const char *kKernelNames[K_COUNT] = {"k_sweep_x", "k_sweep_y", "k_sweep_z", "k_rhs_update",
                                     "k_sweepx_fused", "k_sweepy_fused", "k_sweepz_fused",
                                     "k_emf", "k_finish", "k_dt",
                                     "k_hd_sweep_x", "k_hd_sweep_y", "k_hd_sweep_z",
                                     "k_sweepxy_fused"};

// HIP events around each launch on the context stream; harvested after the
// stream drains.
struct Timing : LaunchTimer {
  hipStream_t stream = nullptr;
  bool on = false;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
  hipEvent_t open = nullptr;
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
  void begin(int) override { open = get(); hipEventRecord(open, stream); }
  void end(int k) override
  {
    hipEvent_t e = get();
    hipEventRecord(e, stream);
    pending.push_back({k, {open, e}});
    open = nullptr;
  }
  void harvest()
  {
    for (auto &pe : pending) {
      float ms = 0.f;
      hipEventSynchronize(pe.second.second);
      if (hipEventElapsedTime(&ms, pe.second.first, pe.second.second) == hipSuccess) {
        launches[pe.first] += 1;
        total_ms[pe.first] += ms;
      }
      pool.push_back(pe.second.first);
      pool.push_back(pe.second.second);
    }
    pending.clear();
  }
  void release()
  {
    for (auto &pe : pending) { hipEventDestroy(pe.second.first); hipEventDestroy(pe.second.second); }
    pending.clear();
    for (auto e : pool) hipEventDestroy(e);
    pool.clear();
  }
};

}  // namespace

struct gpl_ctx {
  gpl_config cfg;
  int dev = 0;
  hipStream_t stream = nullptr;
  StepCtl *ctl = nullptr;
  uint32_t *err = nullptr;
  Engine *engine = nullptr;
  size_t bytes = 0;
  hipGraphExec_t graph_exec = nullptr;
  const void *graph_key = nullptr;   // engine state the captured graph starts from
  Timing timing;
  hipEvent_t run_beg = nullptr, run_end = nullptr;
  double last_run_ms = 0.0;
  // dbl dumps: the state's file image is gathered on the context stream into
  // dump_dev (followed by the step-control block), copied to pinned host
  // memory on dump_stream and written to disk by dump_thread
  hipStream_t dump_stream = nullptr;
  hipEvent_t dump_snap = nullptr, dump_copied = nullptr;
  double *dump_dev = nullptr, *dump_host = nullptr;
  size_t dump_cap = 0;              // doubles
  std::thread dump_thread;
  bool dump_inflight = false;
  int dump_rc = 0;
};

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
  if (c->dump_inflight) { c->dump_thread.join(); c->dump_inflight = false; }
  if (c->dump_dev) hipFree(c->dump_dev);
  if (c->dump_host) hipHostFree(c->dump_host);
  if (c->dump_snap) hipEventDestroy(c->dump_snap);
  if (c->dump_copied) hipEventDestroy(c->dump_copied);
  if (c->dump_stream) hipStreamDestroy(c->dump_stream);
  delete c->engine;
  c->engine = nullptr;
  if (c->err) hipFree(c->err);
  if (c->ctl) hipFree(c->ctl);
  if (c->graph_exec) hipGraphExecDestroy(c->graph_exec);
  c->timing.release();
  if (c->run_beg) hipEventDestroy(c->run_beg);
  if (c->run_end) hipEventDestroy(c->run_end);
  if (c->stream) hipStreamDestroy(c->stream);
}

extern "C" int gpl_create(const gpl_config *cfg, gpl_ctx **out)
{
  if (!cfg || !out) return GPL_EINVAL;
  *out = nullptr;
  if (cfg->physics != GPL_MHD && cfg->physics != GPL_HD && cfg->physics != GPL_RMHD)
    return GPL_EUNSUPPORTED;
  if (cfg->time_stepping != GPL_RK2 && cfg->time_stepping != GPL_RK3) return GPL_EINVAL;
  if (cfg->limiter != GPL_MC_LIM && cfg->limiter != GPL_VANLEER_LIM) return GPL_EINVAL;
  if (cfg->nx < 1 || cfg->ny < 1 || cfg->nz < 1) return GPL_EINVAL;
  const long long ncell = (long long)cfg->nx*cfg->ny*cfg->nz;
  if (ncell > (1LL << 30)) return GPL_EINVAL;   // 32-bit cell index, 5*n fits size_t math
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || cfg->device < 0 || cfg->device >= ndev)
    return GPL_EHIP;

  gpl_ctx *c = new (std::nothrow) gpl_ctx();
  if (!c) return GPL_ENOMEM;
  c->cfg = *cfg;
  c->dev = cfg->device;
  if (hipSetDevice(c->dev) != hipSuccess ||
      hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return GPL_EHIP;
  }
  c->timing.stream = c->stream;
  int rc = dev_alloc(c->bytes, (void **)&c->err, sizeof(uint32_t));
  rc |= dev_alloc(c->bytes, (void **)&c->ctl, sizeof(StepCtl));
  if (rc) {
    free_all(c);
    delete c;
    return GPL_ENOMEM;
  }
  const EngineEnv env = {c->dev, c->stream, c->ctl, c->err};
  rc = cfg->physics == GPL_HD ? make_hd_engine(*cfg, env, &c->engine)
     : cfg->physics == GPL_MHD && cfg->dimensions == 2 ? make_batch2d_engine(*cfg, env, &c->engine)
                                                      : make_mhd3d_engine(*cfg, env, &c->engine);
  if (rc) {
    free_all(c);
    delete c;
    return rc;
  }
  c->bytes += c->engine->bytes;
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

static int gpl_upload_impl(gpl_ctx *c, const double *vc, const double *const *vs, int layout)
{
  if (!c || !vc) return GPL_EINVAL;
  if (layout != GPL_LAYOUT_ACTIVE && layout != GPL_LAYOUT_ARY4D) return GPL_EINVAL;
  hipSetDevice(c->dev);
  if (hipStreamSynchronize(c->stream) != hipSuccess) return GPL_EHIP;
  const int rc = c->engine->upload(vc, vs, layout);
  if (rc) return rc;
  StepCtl z = {0.0, 0.0, 0.0, 0};
  if (hipMemcpy(c->ctl, &z, sizeof(z), hipMemcpyHostToDevice) != hipSuccess) return GPL_EHIP;
  return 0;
}

extern "C" int gpl_upload(gpl_ctx *c, const double *vc, const double *const *vs, int layout)
{
  return c_guard<int>(GPL_ENOMEM, [&] { return gpl_upload_impl(c, vc, vs, layout); });
}

static int gpl_download_impl(gpl_ctx *c, double *vc, double *const *vs, double *uc, int layout)
{
  if (!c) return GPL_EINVAL;
  if (layout != GPL_LAYOUT_ACTIVE && layout != GPL_LAYOUT_ARY4D) return GPL_EINVAL;
  hipSetDevice(c->dev);
  if (hipStreamSynchronize(c->stream) != hipSuccess) return GPL_EHIP;
  return c->engine->download(vc, vs, uc, layout);
}

extern "C" int gpl_download(gpl_ctx *c, double *vc, double *const *vs, double *uc, int layout)
{
  return c_guard<int>(GPL_ENOMEM, [&] { return gpl_download_impl(c, vc, vs, uc, layout); });
}

extern "C" int gpl_set_step(gpl_ctx *c, double time, double dt, long long nstep)
{
  if (!c) return GPL_EINVAL;
  hipSetDevice(c->dev);
  StepCtl h;
  if (hipStreamSynchronize(c->stream) != hipSuccess ||
      hipMemcpy(&h, c->ctl, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess)
    return GPL_EHIP;
  h.time = time;
  h.dt = dt;
  h.nstep = nstep;
  return hipMemcpy(c->ctl, &h, sizeof(h), hipMemcpyHostToDevice) == hipSuccess ? 0 : GPL_EHIP;
}

extern "C" int gpl_get_step(gpl_ctx *c, double *time, double *dt, long long *nstep)
{
  if (!c) return GPL_EINVAL;
  hipSetDevice(c->dev);
  StepCtl h;
  if (hipStreamSynchronize(c->stream) != hipSuccess ||
      hipMemcpy(&h, c->ctl, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess)
    return GPL_EHIP;
  if (time) *time = h.time;
  if (dt) *dt = h.dt;
  if (nstep) *nstep = h.nstep;
  return 0;
}

// ---------------------------------------------------------------------------
// dbl dumps and restarts
// ---------------------------------------------------------------------------
namespace {

// dumped variables in file order (Src/set_output.cpp:129-193: NVAR, then the
// staggered components of the DIMENSIONS directions)
std::vector<std::string> dump_names(const gpl_config &g)
{
  std::vector<std::string> v = {"rho", "vx1", "vx2", "vx3"};
  if (g.physics != GPL_HD) { v.push_back("Bx1"); v.push_back("Bx2"); v.push_back("Bx3"); }
  v.push_back("prs");
  if (g.physics != GPL_HD)
    for (int d = 0; d < g.dimensions; d++) v.push_back("Bx" + std::to_string(d + 1) + "s");
  return v;
}

// the dump's staggered component d: n + 1 faces along d
void face_dims(const gpl_config &g, int d, int e[3])
{
  e[0] = g.nx + (d == 0); e[1] = g.ny + (d == 1); e[2] = g.nz + (d == 2);
}

size_t face_array_size(const Layout &A, int d)
{
  return (size_t)(A.tx + (d == 0))*(A.ty + (d == 1))*(A.tz + (d == 2));
}

// doubles in a dump of this configuration
size_t dump_count(const gpl_config &g, int nv)
{
  size_t count = (size_t)nv*g.nx*g.ny*g.nz;
  if (g.physics != GPL_HD)
    for (int d = 0; d < g.dimensions; d++) {
      int e[3];
      face_dims(g, d, e);
      count += (size_t)e[0]*e[1]*e[2];
    }
  return count;
}

constexpr size_t kCtlDoubles = sizeof(StepCtl)/sizeof(double);
static_assert(sizeof(StepCtl) % sizeof(double) == 0, "StepCtl packs into doubles");

}  // namespace

extern "C" int gpl_wait_dbl(gpl_ctx *c)
{
  if (!c) return GPL_EINVAL;
  if (!c->dump_inflight) return 0;
  c->dump_thread.join();
  c->dump_inflight = false;
  const int rc = c->dump_rc;
  c->dump_rc = 0;
  return rc;
}

static int gpl_write_dbl_async_impl(gpl_ctx *c, const char *dir, int nfile, int nstep)
{
  if (!c || !dir || nfile < 0) return GPL_EINVAL;
  hipSetDevice(c->dev);
  // one dump in flight: the previous one finishes (and reports) first
  if (int rc = gpl_wait_dbl(c)) return rc;
  const size_t count = dump_count(c->cfg, c->engine->nvar());
  const size_t need = count + kCtlDoubles;
  if (!c->dump_stream) {
    if (hipStreamCreateWithFlags(&c->dump_stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->dump_snap, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->dump_copied, hipEventDisableTiming) != hipSuccess)
      return GPL_EHIP;
  }
  if (need > c->dump_cap) {
    if (c->dump_dev) { hipFree(c->dump_dev); c->bytes -= c->dump_cap*sizeof(double); }
    if (c->dump_host) hipHostFree(c->dump_host);
    c->dump_dev = c->dump_host = nullptr;
    c->dump_cap = 0;
    if (hipMalloc((void **)&c->dump_dev, need*sizeof(double)) != hipSuccess) {
      c->dump_dev = nullptr;
      return GPL_ENOMEM;
    }
    if (hipHostMalloc((void **)&c->dump_host, need*sizeof(double), hipHostMallocDefault) != hipSuccess) {
      c->dump_host = nullptr;
      hipFree(c->dump_dev);
      c->dump_dev = nullptr;
      return GPL_ENOMEM;
    }
    c->dump_cap = need;
    c->bytes += need*sizeof(double);
  }
  // snapshot on the context stream, behind the steps already enqueued: the
  // file image, then the step-control block (time, dt, step counter)
  if (int rc = c->engine->dump_gather(c->dump_dev, c->stream)) return rc;
  if (hipMemcpyAsync(c->dump_dev + count, c->ctl, sizeof(StepCtl), hipMemcpyDeviceToDevice,
                     c->stream) != hipSuccess ||
      hipEventRecord(c->dump_snap, c->stream) != hipSuccess ||
      hipStreamWaitEvent(c->dump_stream, c->dump_snap, 0) != hipSuccess ||
      hipMemcpyAsync(c->dump_host, c->dump_dev, need*sizeof(double), hipMemcpyDeviceToHost,
                     c->dump_stream) != hipSuccess ||
      hipEventRecord(c->dump_copied, c->dump_stream) != hipSuccess)
    return GPL_EHIP;
  const std::string d(dir);
  std::vector<std::string> names = dump_names(c->cfg);
  const int dev = c->dev;
  c->dump_rc = 0;
  c->dump_inflight = true;
  c->dump_thread = std::thread([c, d, nfile, nstep, count, dev, names]() {
    hipSetDevice(dev);
    if (hipEventSynchronize(c->dump_copied) != hipSuccess) { c->dump_rc = GPL_EHIP; return; }
    StepCtl h;
    memcpy(&h, c->dump_host + count, sizeof(h));
    const DumpStep st = {h.time, h.dt, nstep >= 0 ? nstep : (int)h.nstep};
    c->dump_rc = dbl_write(d.c_str(), nfile, st, names, c->dump_host, count);
  });
  return 0;
}

extern "C" int gpl_write_dbl_async(gpl_ctx *c, const char *dir, int nfile, int nstep)
{
  return c_guard<int>(GPL_ENOMEM, [&] { return gpl_write_dbl_async_impl(c, dir, nfile, nstep); });
}

extern "C" int gpl_write_dbl(gpl_ctx *c, const char *dir, int nfile, int nstep)
{
  if (int rc = gpl_write_dbl_async(c, dir, nfile, nstep)) return rc;
  return gpl_wait_dbl(c);
}

static int gpl_read_dbl_impl(gpl_ctx *c, const char *dir, int nfile)
{
  if (!c || !dir || nfile < 0) return GPL_EINVAL;
  hipSetDevice(c->dev);
  if (int rc = gpl_wait_dbl(c)) return rc;      // a dump in flight may be writing these files
  const gpl_config &g = c->cfg;
  const bool mhd = g.physics != GPL_HD;
  const size_t na = (size_t)g.nx*g.ny*g.nz;
  const int nv = c->engine->nvar();
  const size_t count = dump_count(g, nv);
  DumpStep st;
  std::vector<double> buf;
  if (int rc = dbl_read(dir, nfile, dump_names(g), count, &st, buf)) return rc;
  if (!mhd) {
    if (int rc = gpl_upload(c, buf.data(), nullptr, GPL_LAYOUT_ACTIVE)) return rc;
  } else {
    // the reference's array layout carries every face of the dump
    const Layout A = make_layout(g, GPL_LAYOUT_ARY4D);
    std::vector<double> vc((size_t)nv*A.tx*A.ty*A.tz, 0.0);
    for (int q = 0; q < nv; q++)
      for (int k = 0; k < g.nz; k++)
        for (int j = 0; j < g.ny; j++)
          for (int i = 0; i < g.nx; i++)
            vc[A.cell(q, i, j, k)] = buf[(size_t)q*na + ((size_t)k*g.ny + j)*g.nx + i];
    std::vector<std::vector<double>> fa(3);
    size_t o = (size_t)nv*na;
    for (int d = 0; d < 3; d++) {
      fa[d].assign(face_array_size(A, d), 0.0);
      if (d < g.dimensions) {
        int e[3];
        face_dims(g, d, e);
        for (int k = 0; k < e[2]; k++)
          for (int j = 0; j < e[1]; j++)
            for (int i = 0; i < e[0]; i++) fa[d][A.face(d, i, j, k)] = buf[o++];
      } else {
        // a direction outside DIMENSIONS has no staggered component: its
        // slot holds the cell-centred field on both faces of the one zone
        for (int k = 0; k <= g.nz - (d == 2 ? 0 : 1); k++)
          for (int j = 0; j <= g.ny - (d == 1 ? 0 : 1); j++)
            for (int i = 0; i <= g.nx - (d == 0 ? 0 : 1); i++) {
              const int ic = i < g.nx ? i : g.nx - 1, jc = j < g.ny ? j : g.ny - 1,
                        kc = k < g.nz ? k : g.nz - 1;
              fa[d][A.face(d, i, j, k)] = buf[(size_t)(4 + d)*na + ((size_t)kc*g.ny + jc)*g.nx + ic];
            }
      }
    }
    const double *vs[3] = {fa[0].data(), fa[1].data(), fa[2].data()};
    if (int rc = gpl_upload(c, vc.data(), vs, GPL_LAYOUT_ARY4D)) return rc;
  }
  return gpl_set_step(c, st.t, st.dt, st.nstep);
}

extern "C" int gpl_read_dbl(gpl_ctx *c, const char *dir, int nfile)
{
  return c_guard<int>(GPL_ENOMEM, [&] { return gpl_read_dbl_impl(c, dir, nfile); });
}

// ---------------------------------------------------------------------------
// stepping
// ---------------------------------------------------------------------------
// The step-control block and the error bits move between host and device on
// the context's own stream, never through the null stream, so a context is
// not held up by (or made to fail by) another context's work or graph capture
// in another thread.
static bool read_ctl(gpl_ctx *c, StepCtl *h, uint32_t *err)
{
  return hipMemcpyAsync(h, c->ctl, sizeof(*h), hipMemcpyDeviceToHost, c->stream) == hipSuccess &&
         (!err || hipMemcpyAsync(err, c->err, sizeof(*err), hipMemcpyDeviceToHost, c->stream) == hipSuccess) &&
         hipStreamSynchronize(c->stream) == hipSuccess;
}

static bool write_ctl(gpl_ctx *c, const StepCtl *h)
{
  static const uint32_t zero = 0;
  return hipMemcpyAsync(c->ctl, h, sizeof(*h), hipMemcpyHostToDevice, c->stream) == hipSuccess &&
         hipMemcpyAsync(c->err, &zero, sizeof(zero), hipMemcpyHostToDevice, c->stream) == hipSuccess &&
         hipStreamSynchronize(c->stream) == hipSuccess;
}

static uint32_t gpl_advance_step_impl(gpl_ctx *c, double dt, double *inv_dt_hyp)
{
  if (!c) return GPL_STEP_API_ERROR;
  hipSetDevice(c->dev);
  StepCtl h_in, h_out;
  if (!read_ctl(c, &h_in, nullptr)) return GPL_STEP_API_ERROR;
  h_in.dt = dt;
  if (!write_ctl(c, &h_in)) return GPL_STEP_API_ERROR;
  c->engine->enqueue_step(0, c->timing.on ? &c->timing : nullptr);
  uint32_t err = 0;
  if (!read_ctl(c, &h_out, &err)) return GPL_STEP_API_ERROR;
  c->timing.harvest();
  if (inv_dt_hyp) *inv_dt_hyp = h_out.inv_dt;
  return err;
}

extern "C" uint32_t gpl_advance_step(gpl_ctx *c, double dt, double *inv_dt_hyp)
{
  return c_guard<uint32_t>(GPL_STEP_API_ERROR, [&] { return gpl_advance_step_impl(c, dt, inv_dt_hyp); });
}

extern "C" double gpl_next_time_step(const gpl_ctx *c, double dt, double inv_dt_hyp)
{
  if (!c) return 0.0;
  const double dt_new = c->cfg.cfl/inv_dt_hyp;                       // next_time_step.cpp:80
  return dt_new <= c->cfg.cfl_max_var*dt ? dt_new : c->cfg.cfl_max_var*dt;   // :125 MIN()
}

// Capture graph_period() steps starting from the engine's current buffers;
// replaying it leaves the buffers where they started.
static int build_graph(gpl_ctx *c)
{
  if (c->graph_exec && c->graph_key == c->engine->state_key()) return 0;
  if (c->graph_exec) { hipGraphExecDestroy(c->graph_exec); c->graph_exec = nullptr; }
  const void *key = c->engine->state_key();
  hipGraph_t graph;
  if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return GPL_EHIP;
  for (int s = 0; s < c->engine->graph_period(); s++) c->engine->enqueue_step(1, nullptr);
  if (hipStreamEndCapture(c->stream, &graph) != hipSuccess) return GPL_EHIP;
  if (c->engine->state_key() != key) { hipGraphDestroy(graph); return GPL_EHIP; }
  hipError_t e = hipGraphInstantiate(&c->graph_exec, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  if (e != hipSuccess) { c->graph_exec = nullptr; return GPL_EHIP; }
  c->graph_key = key;
  return 0;
}

static uint32_t gpl_run_impl(gpl_ctx *c, int nsteps, double first_dt, double *time_out, double *dt_out)
{
  if (!c || nsteps < 0) return GPL_STEP_API_ERROR;
  hipSetDevice(c->dev);
  StepCtl h_in, h_out;
  if (!read_ctl(c, &h_in, nullptr)) return GPL_STEP_API_ERROR;
  if (first_dt > 0.0) { h_in.dt = first_dt; h_in.time = 0.0; h_in.nstep = 0; }
  if (!write_ctl(c, &h_in)) return GPL_STEP_API_ERROR;
  const bool use_graph = !c->timing.on && getenv("GPL_NO_GRAPH") == nullptr;
  const int period = c->engine->graph_period();
  int done = 0;
  hipEventRecord(c->run_beg, c->stream);
  if (use_graph && nsteps >= period) {
    if (build_graph(c) != 0) return GPL_STEP_API_ERROR;
    while (nsteps - done >= period) {
      if (hipGraphLaunch(c->graph_exec, c->stream) != hipSuccess) return GPL_STEP_API_ERROR;
      done += period;
    }
  }
  while (done < nsteps) {
    c->engine->enqueue_step(1, c->timing.on ? &c->timing : nullptr);
    done++;
  }
  hipEventRecord(c->run_end, c->stream);
  uint32_t err = 0;
  if (!read_ctl(c, &h_out, &err)) return GPL_STEP_API_ERROR;
  float ms = 0.f;
  hipEventElapsedTime(&ms, c->run_beg, c->run_end);
  c->last_run_ms = ms;
  c->timing.harvest();
  if (time_out) *time_out = h_out.time;
  if (dt_out) *dt_out = h_out.dt;
  return err;
}

extern "C" uint32_t gpl_run(gpl_ctx *c, int nsteps, double first_dt, double *time_out, double *dt_out)
{
  return c_guard<uint32_t>(GPL_STEP_API_ERROR, 
                  [&] { return gpl_run_impl(c, nsteps, first_dt, time_out, dt_out); });
}

// ---------------------------------------------------------------------------
// zone-by-zone error reports
// ---------------------------------------------------------------------------
int Engine::set_zone_reports(int on)
{
  const size_t cells = zone_cells();
  if (!cells) return GPL_EUNSUPPORTED;
  if (zerr) { (void)hipFree(zerr); zerr = nullptr; bytes -= 3*cells*sizeof(uint32_t); }
  if (zcnt) { (void)hipFree(zcnt); zcnt = nullptr; bytes -= 4*sizeof(unsigned long long); }
  if (on) {
    if (dev_alloc(bytes, (void **)&zerr, 3*cells*sizeof(uint32_t)) ||
        dev_alloc(bytes, (void **)&zcnt, 4*sizeof(unsigned long long))) {
      zone_changed();
      return GPL_ENOMEM;
    }
    if (hipMemset(zerr, 0, 3*cells*sizeof(uint32_t)) != hipSuccess ||
        hipMemset(zcnt, 0, 4*sizeof(unsigned long long)) != hipSuccess)
      return GPL_EHIP;
  }
  zone_changed();
  return 0;
}

int Engine::get_zone_reports(uint32_t *err, unsigned long long *counts, unsigned long long *total)
{
  if (!zerr) return GPL_EINVAL;
  const int ns = nstages();
  const size_t cells = zone_cells();
  unsigned long long h[4];
  if (hipMemcpy(h, zcnt, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess) return GPL_EHIP;
  if (counts) for (int s = 0; s < ns; s++) counts[s] = h[s];
  if (total) *total = h[3];
  if (err) {
    std::vector<uint32_t> hs(cells);
    const size_t nact = zone_active();
    for (int s = 0; s < ns; s++) {
      if (hipMemcpy(hs.data(), zerr + (size_t)s*cells, cells*sizeof(uint32_t),
                    hipMemcpyDeviceToHost) != hipSuccess)
        return GPL_EHIP;
      zone_extract(hs.data(), err + (size_t)s*nact);
    }
  }
  return ns;
}

static int gpl_set_zone_reports_impl(gpl_ctx *c, int enable)
{
  if (!c) return GPL_EINVAL;
  hipSetDevice(c->dev);
  if (hipStreamSynchronize(c->stream) != hipSuccess) return GPL_EHIP;
  const size_t before = c->engine->bytes;
  const int rc = c->engine->set_zone_reports(enable);
  if (c->graph_exec) {               // the captured kernels carry the old pointers
    hipGraphExecDestroy(c->graph_exec);
    c->graph_exec = nullptr;
  }
  c->bytes = c->bytes - before + c->engine->bytes;
  return rc;
}

extern "C" int gpl_set_zone_reports(gpl_ctx *c, int enable)
{
  return c_guard<int>(GPL_ENOMEM, [&] { return gpl_set_zone_reports_impl(c, enable); });
}

static int gpl_get_zone_reports_impl(gpl_ctx *c, uint32_t *err, unsigned long long *counts,
                                     unsigned long long *total)
{
  if (!c) return GPL_EINVAL;
  hipSetDevice(c->dev);
  if (hipStreamSynchronize(c->stream) != hipSuccess) return GPL_EHIP;
  return c->engine->get_zone_reports(err, counts, total);
}

extern "C" int gpl_get_zone_reports(gpl_ctx *c, uint32_t *err, unsigned long long *counts,
                                    unsigned long long *total)
{
  return c_guard<int>(GPL_ENOMEM, [&] { return gpl_get_zone_reports_impl(c, err, counts, total); });
}

static int gpl_set_c2p_network_impl(gpl_ctx *c, const float *weights)
{
  if (!c) return GPL_EINVAL;
  hipSetDevice(c->dev);
  if (hipStreamSynchronize(c->stream) != hipSuccess) return GPL_EHIP;
  const int rc = c->engine->set_c2p_network(weights);
  if (rc == 0 && c->graph_exec) {      // the captured kernels carry the old argument
    hipGraphExecDestroy(c->graph_exec);
    c->graph_exec = nullptr;
  }
  return rc;
}

extern "C" int gpl_set_c2p_network(gpl_ctx *c, const float *weights)
{
  return c_guard<int>(GPL_ENOMEM, [&] { return gpl_set_c2p_network_impl(c, weights); });
}

static int gpl_diagnostics_impl(gpl_ctx *c, double *vmin, double *vmax, double *mach_max,
                                double *divb_max)
{
  if (!c) return GPL_EINVAL;
  hipSetDevice(c->dev);
  double out[DIAG_MAXQ];
  const int nv = c->engine->nvar();
  if (const int rc = c->engine->diagnostics(out)) return rc;
  for (int a = 0; a < nv; a++) {
    if (vmin) vmin[a] = out[a];
    if (vmax) vmax[a] = out[nv + a];
  }
  if (mach_max) *mach_max = out[2*nv];
  if (divb_max) *divb_max = out[2*nv + 1];
  return 0;
}

extern "C" int gpl_diagnostics(gpl_ctx *c, double *vmin, double *vmax, double *mach_max,
                               double *divb_max)
{
  return c_guard<int>(GPL_ENOMEM, [&] { return gpl_diagnostics_impl(c, vmin, vmax, mach_max, divb_max); });
}

static int gpl_set_timing_impl(gpl_ctx *c, int enable)
{
  if (!c) return GPL_EINVAL;
  c->timing.on = enable != 0;
  for (int k = 0; k < K_COUNT; k++) { c->timing.launches[k] = 0; c->timing.total_ms[k] = 0.0; }
  return 0;
}

extern "C" int gpl_set_timing(gpl_ctx *c, int enable)
{
  return c_guard<int>(GPL_ENOMEM, [&] { return gpl_set_timing_impl(c, enable); });
}

extern "C" int gpl_kernel_count(const gpl_ctx *) { return K_COUNT; }

static int gpl_kernel_stats_impl(gpl_ctx *c, int idx, const char **name, long long *launches,
                                 double *total_ms)
{
  if (!c || idx < 0 || idx >= K_COUNT) return GPL_EINVAL;
  c->timing.harvest();
  if (name) *name = kKernelNames[idx];
  if (launches) *launches = c->timing.launches[idx];
  if (total_ms) *total_ms = c->timing.total_ms[idx];
  return 0;
}

extern "C" int gpl_kernel_stats(gpl_ctx *c, int idx, const char **name, long long *launches,
                                double *total_ms)
{
  return c_guard<int>(GPL_ENOMEM, [&] { return gpl_kernel_stats_impl(c, idx, name, launches, total_ms); });
}

extern "C" double gpl_last_run_ms(const gpl_ctx *c) { return c ? c->last_run_ms : 0.0; }

extern "C" size_t gpl_device_bytes(const gpl_ctx *c) { return c ? c->bytes : 0; }

extern "C" const char *gpl_strerror(int code)
{
  switch (code) {
  case 0: return "ok";
  case GPL_EINVAL: return "invalid argument";
  case GPL_EHIP: return "HIP runtime error";
  case GPL_ENOMEM: return "out of device or host memory";
  case GPL_EUNSUPPORTED: return "configuration not supported";
  case GPL_EIO: return "dump or restart file error";
  case GPL_ESCRATCH: return "request larger than the fixed scratch of a no-allocation entry point";
  default: return "unknown error";
  }
}
