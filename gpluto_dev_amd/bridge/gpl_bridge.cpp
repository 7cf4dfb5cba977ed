// Reference-side binding of the MI355X stage update: a drop-in for
// AdvanceStep() (Src/rk_step.cpp:88, called at Src/main.cpp:446).
//
// A gPLUTO build links this file instead of rk_step.o, plus
// libgpluto_amd.so.  It compiles against the reference's own headers
// (gpluto.hpp with the run's definitions.hpp) and uses only what they declare
// public: the Data/Grid/timeStep/Runtime structures, the Ary3D/Ary4D element
// accessors and QuitPLUTO.  Everything else the main loop does stays
// unchanged: NextTimeStep (Src/next_time_step.cpp:20, on the invDt_hyp this
// call returns), output, restart and the log.
//
// Data movement: the first call creates the library context from the run's
// definitions.hpp and pluto.ini selections and uploads Data::Vc and Data::Vs
// (the host arrays are authoritative at start-up and after a restart).  Every
// call advances the state on the device and copies Vc, Vs and Uc back, so the
// host arrays hold the step's result whenever the main loop looks at them
// (output, diagnostics), as they do after the reference's own AdvanceStep.
// GPL_BRIDGE_SYNC=0 in the environment skips those copies; the state then
// stays on the device until gpl_bridge_sync() is called (before output).
#include "gpluto.hpp"

#include <cstdlib>
#include <cstring>

#include "../../include/gpluto_amd.h"

namespace {

gpl_ctx *g_ctx = nullptr;
bool g_sync_every_step = true;

int side_type(int reference_bound)
{
  switch (reference_bound) {
  case PERIODIC:   return GPL_PERIODIC;
  case OUTFLOW:    return GPL_OUTFLOW;
  case REFLECTIVE: return GPL_REFLECTIVE;
  default:         return 0;
  }
}

// Data::Vc / Vs / Uc as plain pointers: element (0, 0, 0, 0) of an Ary4D and
// (0, 0, 0) of an Ary3D start their storage (i fastest, Src/Array.hpp:476-479)
double *vc_ptr(Data *d) { return &d->Vc(0, 0, 0, 0); }
double *uc_ptr(Data *d) { return &d->Uc(0, 0, 0, 0); }

void face_ptrs(Data *d, double *vs[3])
{
  vs[0] = vs[1] = vs[2] = nullptr;
#if PHYSICS == MHD || PHYSICS == RMHD
  DIM_EXPAND(vs[0] = &d->Vs[BX1s](0, 0, 0);, vs[1] = &d->Vs[BX2s](0, 0, 0);,
             vs[2] = &d->Vs[BX3s](0, 0, 0);)
#endif
}

void create(Data *data, Grid *grid)
{
  Runtime *rt = RuntimeGet();
  gpl_config c;
  gpl_default_config(&c);
#if PHYSICS == MHD
  c.physics = GPL_MHD;
#elif PHYSICS == RMHD
  c.physics = GPL_RMHD;
#elif PHYSICS == HD
  c.physics = GPL_HD;
#else
  QuitPLUTO("gpl_bridge: PHYSICS is not HD, MHD or RMHD", 1);
#endif
  c.dimensions = DIMENSIONS;
  c.nx = grid->nint[IDIR];
  c.ny = grid->nint[JDIR];
  c.nz = grid->nint[KDIR];
  for (int d = 0; d < 3; d++) { c.xbeg[d] = grid->xbeg[d]; c.xend[d] = grid->xend[d]; }
  c.time_stepping = TIME_STEPPING == RK3 ? GPL_RK3 : GPL_RK2;
  c.limiter = LIMITER == VANLEER_LIM ? GPL_VANLEER_LIM : GPL_MC_LIM;
  const char *s = rt->solv_type;
  c.solver = !strcmp(s, "hlld") ? GPL_SOLVER_HLLD : !strcmp(s, "hllc") ? GPL_SOLVER_HLLC
                                                                      : GPL_SOLVER_HLL;
#if PHYSICS == MHD || PHYSICS == RMHD
  c.ct_emf_average = CT_EMF_AVERAGE == UCT_HLLD ? GPL_UCT_HLLD : GPL_UCT_HLL;
#endif
  c.precision = GPL_FP64;
  c.nghost = grid->nghost[IDIR];
  c.gamma = data->run_config->gamma;
  c.small_rho = data->run_config->small_rho;
  c.small_prs = data->run_config->small_prs;
  c.cfl = rt->cfl;
  c.cfl_max_var = rt->cfl_max_var;
  for (int d = 0; d < DIMENSIONS; d++) {
    c.side_bc[2*d] = side_type(rt->left_bound[d]);
    c.side_bc[2*d + 1] = side_type(rt->right_bound[d]);
    if (!c.side_bc[2*d] || !c.side_bc[2*d + 1])
      QuitPLUTO("gpl_bridge: only periodic, outflow and reflective sides", 1);
  }
  c.boundary = c.side_bc[0];
  if (const int rc = gpl_create(&c, &g_ctx)) {
    printLog("! gpl_bridge: gpl_create: %s\n", gpl_strerror(rc));
    QuitPLUTO("gpl_bridge: gpl_create failed", 1);
  }
  double *vs[3];
  face_ptrs(data, vs);
  if (gpl_upload(g_ctx, vc_ptr(data), vs, GPL_LAYOUT_ARY4D))
    QuitPLUTO("gpl_bridge: gpl_upload failed", 1);
  const char *e = getenv("GPL_BRIDGE_SYNC");
  g_sync_every_step = !(e && !strcmp(e, "0"));
}

}  // namespace

// Copy the device state into Data::Vc, Vs and Uc (ghost zones included).
void gpl_bridge_sync(Data *data)
{
  if (!g_ctx) return;
  double *vs[3];
  face_ptrs(data, vs);
  if (gpl_download(g_ctx, vc_ptr(data), vs, uc_ptr(data), GPL_LAYOUT_ARY4D))
    QuitPLUTO("gpl_bridge: gpl_download failed", 1);
}

// AdvanceStep (Src/rk_step.cpp:88): one full RK step of the configured
// method on the device; Dts->invDt_hyp receives stage 1's maximum of C_dt
// over DIMENSIONS (Src/update_stage.cpp:422-425).  Returns the ERR_* bits.
uint32_t AdvanceStep(Data *data, Grid *grid, timeStep *Dts)
{
  if (!g_ctx) create(data, grid);
  const uint32_t err = gpl_advance_step(g_ctx, Dts->dt, &Dts->invDt_hyp);
  if (err & GPL_STEP_API_ERROR) QuitPLUTO("gpl_bridge: gpl_advance_step failed", 1);
  if (g_sync_every_step) gpl_bridge_sync(data);
  if (err & ERR_QUIT) QuitPLUTO("from AdvanceStep() on the device", (int)err);
  return err;
}
