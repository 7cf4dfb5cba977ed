// Host side of the stage update: one Engine per physics module owns the
// device state of a context and enqueues whole AdvanceStep calls on the
// context's stream.  The C ABI (gpl_api.cpp) holds the engine, the stream,
// the step-control block, graph replay and timing.
#pragma once
#include "../../include/gpluto_amd.h"
#include "common.hpp"

#include <hip/hip_runtime.h>

#include <cstddef>
#include <vector>

namespace gpl {

// Per-launch timing hook (begin/end around each kernel launch).
struct LaunchTimer {
  virtual void begin(int kernel) = 0;
  virtual void end(int kernel) = 0;
  virtual ~LaunchTimer() {}
};

struct EngineEnv {
  int dev;
  hipStream_t stream;
  StepCtl *ctl;       // device step-control block
  uint32_t *err;      // device error bits
};

class Engine {
public:
  virtual ~Engine()
  {
    if (zerr) (void)hipFree(zerr);
    if (zcnt) (void)hipFree(zcnt);
  }
  virtual int upload(const double *vc, const double *const *vs, int layout) = 0;
  virtual int download(double *vc, double *const *vs, double *uc, int layout) = 0;
  // Enqueue one AdvanceStep; advance_dt != 0 also applies NextTimeStep and
  // time += dt on the device.  t is null when launches are not timed.
  virtual void enqueue_step(int advance_dt, LaunchTimer *t) = 0;
  // Steps after which the ping-pong buffers are back where they started
  // (a captured graph replays that many steps).
  virtual int graph_period() const = 0;
  // Identifies the buffer holding the current state (graph validity).
  virtual const void *state_key() const = 0;
  // RMHD: learned con2prim guess weights (nullptr: the reference's guess).
  virtual int set_c2p_network(const float *) { return GPL_EUNSUPPORTED; }
  // DataMinMax / CT_CheckDivB of the current state into out[2*nvar + 2]
  // (min[nvar], max[nvar], max Mach, max |div B|), on the context stream.
  virtual int diagnostics(double *out) = 0;
  virtual int nvar() const = 0;
  // Enqueue on stream s the image of a dbl dump of the current state into
  // the device buffer dst: the active zones of each primitive, then (MHD,
  // RMHD) each staggered component with faces 0..n along its direction.
  virtual int dump_gather(double *dst, hipStream_t s) = 0;
  size_t bytes = 0;   // device memory held

  // Zone-by-zone error reports (off until set_zone_reports(1)): one uint32
  // array of zone_cells() entries per RK stage and the counters [stage 0..2,
  // total].  An engine that supports them returns its device cell count from
  // zone_cells(), maps a stage array to the active zones in zone_extract(),
  // passes zone_rep(stage) to its ConsToPrim kernel and resets the stage
  // counters at the start of every step (launch_zone_reset).
  uint32_t *zerr = nullptr;
  unsigned long long *zcnt = nullptr;
  virtual size_t zone_cells() const { return 0; }
  virtual size_t zone_active() const { return 0; }   // active zones of the domain
  virtual int nstages() const = 0;
  // dst[(k*ny + j)*nx + i] for the active zones, from a stage array on the host
  virtual void zone_extract(const uint32_t *src, uint32_t *dst) const { (void)src; (void)dst; }
  virtual void zone_changed() {}
  ZoneRep zone_rep(int stage) const
  {
    if (!zerr) return ZoneRep{nullptr, nullptr, nullptr};
    return ZoneRep{zerr + (size_t)stage*zone_cells(), zcnt + stage, zcnt + 3};
  }
  // (gpl_api.cpp) allocate and clear, or free, the report buffers
  int set_zone_reports(int on);
  // (gpl_api.cpp) stage arrays of the last step (active zones), stage counts,
  // total; returns the number of stages
  int get_zone_reports(uint32_t *err, unsigned long long *counts, unsigned long long *total);
};

// pluto.ini [Boundary] type of side s (X1-beg, X1-end, X2-beg, ...)
inline int side_type(const gpl_config &c, int s) { return c.side_bc[s] ? c.side_bc[s] : c.boundary; }
inline bool all_periodic(const gpl_config &c)
{
  for (int s = 0; s < 2*c.dimensions; s++) if (side_type(c, s) != GPL_PERIODIC) return false;
  return true;
}

int make_mhd3d_engine(const gpl_config &cfg, const EngineEnv &env, Engine **out);
int make_hd_engine(const gpl_config &cfg, const EngineEnv &env, Engine **out);
int make_batch2d_engine(const gpl_config &cfg, const EngineEnv &env, Engine **out);   // 2D MHD

// Host-array layouts accepted by gpl_upload / gpl_download: the reference's
// Ary4D with ghost zones along the used directions (Src/set_grid.cpp:57-58:
// no ghosts along a direction that is not included), or active zones only.
struct Layout {
  int nx, ny, nz;
  int gx, gy, gz;          // ghost zones per side
  int tx, ty, tz;          // totals
  bool ary;
  size_t cell(int nv, int i, int j, int k) const
  {
    if (!ary) return (size_t)nv*nx*ny*nz + ((size_t)k*ny + j)*nx + i;
    return (size_t)nv*tx*ty*tz + ((size_t)(k + gz)*ty + (j + gy))*tx + (i + gx);
  }
  // staggered component d: one extra face along d (Src/initialize.cpp:424-431)
  size_t face(int d, int i, int j, int k) const
  {
    if (!ary) return ((size_t)k*ny + j)*nx + i;
    const int sx = tx + (d == 0), sy = ty + (d == 1);
    return ((size_t)(k + gz)*sy + (j + gy))*sx + (i + gx);
  }
};

inline Layout make_layout(const gpl_config &c, int layout)
{
  Layout L;
  L.nx = c.nx; L.ny = c.ny; L.nz = c.nz;
  L.gx = c.dimensions > 0 ? c.nghost : 0;
  L.gy = c.dimensions > 1 ? c.nghost : 0;
  L.gz = c.dimensions > 2 ? c.nghost : 0;
  L.ary = layout == GPL_LAYOUT_ARY4D;
  L.tx = L.nx + 2*L.gx; L.ty = L.ny + 2*L.gy; L.tz = L.nz + 2*L.gz;
  return L;
}

// device allocation that adds to an engine's byte count
inline int dev_alloc(size_t &counter, void **p, size_t bytes)
{
  if (hipMalloc(p, bytes) != hipSuccess) { *p = nullptr; return GPL_ENOMEM; }
  counter += bytes;
  return 0;
}

}  // namespace gpl
