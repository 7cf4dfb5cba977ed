/* CPU oracle: a plain-C restatement of the reference's Godunov stage update.
 *
 * TEST INFRASTRUCTURE ONLY.  The product path never links this code; only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it, and only as the checker.
 *
 * Data layout and step semantics follow the reference (Src/rk_step.cpp:88,
 * Src/update_stage.cpp:57, Src/main.cpp:320-485): arrays carry nghost = 2
 * ghost zones per active direction, Vc/Uc are [nv][k][j][i], staggered
 * fields carry one extra face along every direction.  Public state transfer
 * uses the periodic "one face per cell" layout of gpluto_dev_amd.ic.
 */
#ifndef GPL_ORACLE_H
#define GPL_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { GPLO_HD = 1, GPLO_MHD = 2, GPLO_RMHD = 3 };
enum { GPLO_HLL = 1, GPLO_HLLC = 2, GPLO_HLLD = 3 };
enum { GPLO_VANLEER_LIM = 26, GPLO_MC_LIM = 27 };    /* Src/gpluto.hpp:330-336 */
enum { GPLO_UCT_HLL = 1, GPLO_UCT_HLLD = 2 };   /* CT_EMF_AVERAGE (Src/MHD/CT/ct.hpp:20-21) */
enum { GPLO_BC_PERIODIC = 1, GPLO_BC_OUTFLOW = 2, GPLO_BC_REFLECTIVE = 3 };   /* pluto.ini [Boundary] */

typedef struct gplo_config {
  int    phys;            /* GPLO_HD, GPLO_MHD or GPLO_RMHD */
  int    dims;            /* 1, 2 or 3 */
  int    nx, ny, nz;      /* active zones (1 for absent directions) */
  int    periodic;        /* 1 periodic on every side, 0 outflow */
  int    time_stepping;   /* 2 (RK2) or 3 (RK3) */
  int    limiter;         /* GPLO_MC_LIM or GPLO_VANLEER_LIM */
  int    solver;          /* GPLO_HLL / GPLO_HLLC / GPLO_HLLD */
  double gamma;
  double cfl, cfl_max_var;
  double small_rho, small_prs;
  int    bc[6];           /* per side X1-beg, X1-end, X2-beg, ..., X3-end: GPLO_BC_*;
                             0 takes periodic ? PERIODIC : OUTFLOW */
  int    emf;             /* GPLO_UCT_HLL (0 too) or GPLO_UCT_HLLD (MHD with HLLD only) */
} gplo_config;

typedef struct gplo_state gplo_state;

/* NULL for an unsupported configuration (a direction periodic on one side
   only) */
gplo_state *gplo_create(const gplo_config *cfg);
void        gplo_destroy(gplo_state *s);
/* vc[nvar][nz][ny][nx]; vs[dims][nz][ny][nx] (lower faces, periodic) or NULL */
int         gplo_set_state(gplo_state *s, const double *vc, const double *vs);
/* as gplo_set_state, with the upper boundary faces of non-periodic
   directions from up[] (the dims slabs of a restart dump, back to back;
   slab d shaped [nz][ny][nx] with n_d = 1); up may be NULL */
int         gplo_set_state_faces(gplo_state *s, const double *vc, const double *vs,
                                 const double *up);
/* uc may be NULL */
int         gplo_get_state(const gplo_state *s, double *vc, double *vs, double *uc);
/* One AdvanceStep with time step dt; returns the ERR_* bits; *inv_dt gets
   Dts->invDt_hyp of stage 1 (Src/update_stage.cpp:422-425). */
uint32_t    gplo_advance_step(gplo_state *s, double dt, double *inv_dt);
/* The main-loop recurrence (Src/main.cpp:446-463, Src/next_time_step.cpp:80,125):
   run nsteps steps from dt = first_dt; returns OR of the step error bits. */
uint32_t    gplo_run(gplo_state *s, int nsteps, double first_dt,
                     double *time_out, double *dt_out);
int         gplo_nvar(const gplo_state *s);
/* Zone-by-zone error reports of the last step: err[stage][nz][ny][nx] holds
   Data::err as ConsToPrim3D left it in each stage (Src/mappers3D.cpp:115-176),
   counts[stage] the zones Where() counted (Src/debug_tools.cpp:587-593),
   *total the flagged zones since creation; any pointer may be NULL.  Returns
   the number of stages. */
int         gplo_zone_errors(const gplo_state *s, uint32_t *err, unsigned long long *counts,
                             unsigned long long *total);
/* pointwise RMHD pieces for tests: HLL interface solve along dir (flux[8],
   Davis speeds) and con2prim of u[8] (rewrites u like the reference) */
uint32_t    gplo_rmhd_face(const double *vl, const double *vr, int dir, double gamma,
                           double *flux, double *sl, double *sr);
/* one MHD (GPLO_HLL/HLLC/HLLD) or RMHD (GPLO_HLL/HLLC) interface solve */
uint32_t    gplo_face(int phys, int solver, const double *vl, const double *vr, int dir,
                      double gamma, double *flux, double *sl, double *sr);
uint32_t    gplo_rmhd_c2p(double *u, double *v, double gamma, double small_rho, double small_prs);
/* acos and sincos for the RMHD resolvent cubic (Src/RMHD/rmhd.hpp:441-444):
   NULL (the default) calls the host libm, as the reference does */
void        gplo_set_cubic_math(double (*acos_fn)(double),
                                void (*sincos_fn)(double, double *, double *));

#ifdef __cplusplus
}
#endif
#endif
