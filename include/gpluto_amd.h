/* gpluto_amd: MI355X (gfx950) Godunov stage update behind a C ABI.
 *
 * Drop-in for the reference's time-step seam: AdvanceStep()
 * (Src/rk_step.cpp:88, called from Src/main.cpp:446) and NextTimeStep()
 * (Src/next_time_step.cpp:20, called from Src/main.cpp:463).  The device
 * owns every buffer from gpl_create() to gpl_destroy(); host arrays are
 * authoritative only across gpl_upload()/gpl_download() (the reference's
 * "acc enter data" / "acc update self" points, Src/startup.cpp:63-73).
 *
 * Plain C types only.  Every entry point returns 0 (or the reference's
 * ERR_* bitmask, Src/gpluto.hpp:247-265, for the stepping calls) on success
 * and a negative GPL_E* code on an API error.
 */
#ifndef GPLUTO_AMD_H
#define GPLUTO_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* PHYSICS of definitions.hpp (Src/gpluto.hpp) */
#define GPL_HD    1
#define GPL_MHD   2
#define GPL_RMHD  3
/* pluto.ini [Solver] names (Src/MHD/set_solver.cpp:35-47, Src/HD/set_solver.cpp:17-19) */
#define GPL_SOLVER_HLL   1
#define GPL_SOLVER_HLLC  2
#define GPL_SOLVER_HLLD  3
/* LIMITER of definitions.hpp (Src/gpluto.hpp:330-336) */
#define GPL_VANLEER_LIM  26
#define GPL_MC_LIM       27
/* CT_EMF_AVERAGE (Src/MHD/CT/ct.hpp:20-21, 73-96) */
#define GPL_UCT_HLL      1
#define GPL_UCT_HLLD     2   /* CT_Store_UCTHlld (Src/MHD/CT/ct_store.hpp:254); MHD with HLLD only */
/* TIME_STEPPING (Src/gpluto.hpp:121-122) */
#define GPL_RK2          2
#define GPL_RK3          3
/* pluto.ini [Boundary] side types */
#define GPL_PERIODIC     1
#define GPL_OUTFLOW      2
#define GPL_REFLECTIVE   3
/* arithmetic of the stage update */
#define GPL_FP64         0   /* drop-in: FP64 everywhere, reference operation order */
#define GPL_MIXED        1   /* FP32 state/PLM/Riemann; FP64 U, faces, CT, dt */
/* array layouts accepted by gpl_upload / gpl_download */
#define GPL_LAYOUT_ACTIVE 0  /* vc[nv][nz][ny][nx]; vs[d][nz][ny][nx] = lower face of each cell */
#define GPL_LAYOUT_ARY4D  1  /* reference Ary4D (Src/Array.hpp:476-479) with nghost ghost
                                zones per active direction; Vs[d] with one extra face along d */

/* Stepping calls return this bit (never raised by the numerics, whose
   ERR_* bits stop at 1<<18) when the call itself failed. */
#define GPL_STEP_API_ERROR 0x80000000u

/* API error codes */
#define GPL_EINVAL       (-1)
#define GPL_EHIP         (-2)
#define GPL_ENOMEM       (-3)   /* device memory, or host memory or threads for the host side */
#define GPL_EUNSUPPORTED (-4)
#define GPL_EIO          (-5)   /* a dump or restart file could not be opened, read or written */
#define GPL_ESCRATCH     (-6)   /* more than a no-allocation entry point's fixed scratch holds */

/* The selections of definitions.hpp and pluto.ini that the stage update
 * reads (Src/structs.hpp:397-470 RunConfig/Runtime, Src/grid_input.hpp). */
typedef struct gpl_config {
  int    physics;          /* GPL_MHD (2D/3D), GPL_RMHD (3D, FP64) or GPL_HD (1D/2D/3D) */
  int    dimensions;       /* DIMENSIONS; unused directions hold 1 zone */
  int    nx, ny, nz;       /* [Grid] active zones per direction */
  double xbeg[3], xend[3]; /* [Grid] domain; dx = (xend - xbeg)/n (Src/set_grid.cpp:196) */
  int    boundary;         /* type of every side whose side_bc entry is 0: GPL_PERIODIC,
                              GPL_OUTFLOW or GPL_REFLECTIVE */
  int    time_stepping;    /* GPL_RK2 or GPL_RK3 */
  int    limiter;          /* GPL_MC_LIM or GPL_VANLEER_LIM */
  int    solver;           /* MHD and RMHD: GPL_SOLVER_HLLD / _HLLC / _HLL; HD: _HLLC / _HLL */
  int    ct_emf_average;   /* GPL_UCT_HLL (MHD, RMHD), or GPL_UCT_HLLD (MHD with HLLD) */
  int    precision;        /* GPL_FP64, or GPL_MIXED (MHD) */
  int    nghost;           /* ghost zones of the ARY4D layout (2 for PLM) */
  int    device;           /* HIP device ordinal */
  double gamma;            /* run_config.gamma */
  double cfl;              /* [Time] CFL */
  double cfl_max_var;      /* [Time] CFL_max_var */
  double small_rho;        /* run_config.small_rho (default 1e-12) */
  double small_prs;        /* run_config.small_prs (default 1e-12) */
  int    side_bc[6];       /* [Boundary] X1-beg, X1-end, X2-beg, X2-end, X3-beg, X3-end;
                              0 takes `boundary`.  A direction is periodic on both
                              sides or on neither; outflow and reflective mix */
} gpl_config;

/* One domain on one device: its state, its stream and its step state.  Calls
 * on one context come from one host thread at a time (they are not locked
 * against each other); separate contexts, on one device or several, are
 * independent.  The dump writer thread of gpl_write_dbl_async touches only
 * its own staging buffer and files. */
typedef struct gpl_ctx gpl_ctx;

/* Fill *cfg with the reference defaults (3D MHD, HLLD, MC, RK2, UCT_HLL,
 * periodic unit cube, gamma 5/3, CFL 0.4, CFL_max_var 1.1). */
void gpl_default_config(gpl_config *cfg);

/* Allocate the device state; replaces Initialize()'s allocations
 * (Src/initialize.cpp:396-483) and AdvanceStep's static scratch
 * (Src/rk_step.cpp:153-219). */
int gpl_create(const gpl_config *cfg, gpl_ctx **out);
int gpl_destroy(gpl_ctx *ctx);

/* Host -> device copy of the primitive state: vc = Data::Vc and
 * vs[d] = Data::Vs[BX1s+d] (Src/structs.hpp:640-712).  Follows Startup()
 * (Src/startup.cpp:388,514-561): the cell-centred field becomes the face
 * average and Uc = PrimToCons(Vc).  MHD: vc holds NVAR = 8 variables in the
 * reference order rho, vx1, vx2, vx3, Bx1, Bx2, Bx3, prs.  HD: NVAR = 5
 * (rho, vx1, vx2, vx3, prs; Src/HD/hd.hpp:21-28) and vs is ignored. */
int gpl_upload(gpl_ctx *ctx, const double *vc, const double *const *vs, int layout);

/* Device -> host copy of Vc, Vs (MHD) and (if uc != NULL) the conserved Uc;
 * a NULL vc or vs is skipped.  ARY4D: Vs[d] carries the upper boundary face
 * (index n along d): the periodic image of face 0, else its own value. */
int gpl_download(gpl_ctx *ctx, double *vc, double *const *vs, double *uc, int layout);

/* The step state the main loop keeps in timeStep and g_stepNumber
 * (Src/main.cpp:207-209, 336, 458-463): time, the next dt and the step
 * counter, device-resident.  gpl_upload resets them to 0; gpl_run with
 * first_dt <= 0 continues from them and advances all three. */
int gpl_set_step(gpl_ctx *ctx, double time, double dt, long long nstep);
int gpl_get_step(gpl_ctx *ctx, double *time, double *dt, long long *nstep);

/* WriteData for the dbl output in single_file mode followed by RestartDump
 * (Src/main.cpp:673-687; Src/write_data.cpp:86-159, 476-505;
 * Src/restart.cpp:417-453): writes <dir>/data.<nfile>.dbl (the active zones
 * of rho vx1 vx2 vx3 [Bx1 Bx2 Bx3] prs, then MHD/RMHD's Bx1s..Bx<D>s with
 * faces 0..n along their direction), line nfile of <dir>/dbl.out (created
 * when nfile is 0) and record nfile of <dir>/restart.out, with the device's
 * time and dt.  nstep >= 0 is written as the step counter, nstep < 0 writes
 * the context's own count (the reference's post-loop dump carries one less,
 * Src/main.cpp:336,482,497).  Returns GPL_EIO on a file error. */
int gpl_write_dbl(gpl_ctx *ctx, const char *dir, int nfile, int nstep);

/* gpl_write_dbl without the wait, for dumps that overlap the computation
 * (the "async download and overlap with compute" of the output path; the
 * reference writes synchronously, Src/main.cpp:385-387): the state's file
 * image and the step state are gathered on the device behind the work already
 * enqueued, copied to pinned host memory on a second stream and written by a
 * host thread, while later calls (gpl_run ...) proceed.  One dump is in flight
 * per context: the next gpl_write_dbl_async, gpl_read_dbl or gpl_destroy waits
 * for it.  The device and pinned staging buffers (one dump's size) stay
 * allocated until gpl_destroy. */
int gpl_write_dbl_async(gpl_ctx *ctx, const char *dir, int nfile, int nstep);
/* Wait for the dump in flight, if any; returns its result (0 if none). */
int gpl_wait_dbl(gpl_ctx *ctx);

/* RestartFromFile for the dbl output (Src/restart.cpp:34-274, RestartGet
 * :342-406): reads the restart.out record of dump nfile (time, dt, step
 * counter), checks line nfile of dbl.out (single_file, endianness, the
 * variables this context dumps), reads data.<nfile>.dbl (byte-swapped if
 * needed) and uploads it as gpl_upload does: Uc = PrimToCons(Vc) and every
 * face of the dump, the upper boundary face of a non-periodic direction
 * included.  Returns GPL_EINVAL when the files do not hold dump nfile of
 * this configuration, GPL_EIO when one cannot be read. */
int gpl_read_dbl(gpl_ctx *ctx, const char *dir, int nfile);

/* AdvanceStep(data, grid, Dts) with Dts->dt = dt (Src/rk_step.cpp:88):
 * one full RK step on the device.  Returns the OR of the ERR_* bits raised;
 * *inv_dt_hyp receives Dts->invDt_hyp of stage 1
 * (Src/update_stage.cpp:422-425). */
uint32_t gpl_advance_step(gpl_ctx *ctx, double dt, double *inv_dt_hyp);

/* NextTimeStep (Src/next_time_step.cpp:20-130) on one rank:
 * min(cfl/inv_dt_hyp, cfl_max_var*dt); 0 for a null context. */
double gpl_next_time_step(const gpl_ctx *ctx, double dt, double inv_dt_hyp);

/* The main-loop recurrence of Src/main.cpp:446-463 for nsteps steps with the
 * time step kept on the device (no host synchronisation between steps; the
 * step is replayed from a HIP graph).  Starts from dt = first_dt at time 0
 * unless first_dt <= 0, in which case it continues from the device's
 * current dt and time.  Returns the OR of the ERR_* bits. */
uint32_t gpl_run(gpl_ctx *ctx, int nsteps, double first_dt, double *time_out, double *dt_out);

/* DataMinMax (Src/main.cpp:752-860) and CT_CheckDivB
 * (Src/MHD/CT/ct_update.cpp:580-635) of the current state, reduced on the
 * device: vmin[NVAR] and vmax[NVAR] of the primitives over the active zones,
 * the largest Mach number max(|v1|, |v2|, |v3|)/sqrt(p/rho), and for MHD and
 * RMHD the largest |div B| with the reference's face-area weighting (0 for
 * HD).  Any output pointer may be NULL.  The results are the reference's
 * formulas bit for bit: min and max do not depend on reduction order. */
int gpl_diagnostics(gpl_ctx *ctx, double *vmin, double *vmax, double *mach_max, double *divb_max);

/* Zone-by-zone error reports: the reference's Data::err (Src/structs.hpp:678)
 * as ConsToPrim3D writes it in every zone (Src/mappers3D.cpp:115-176), and the
 * flagged-zone counts that Where() adds to Data::err_count
 * (Src/debug_tools.cpp:566-684; reset every step, Src/main.cpp:423).  Off by
 * default: the stepping calls then return only the OR of the ERR_* bits (the
 * reference's err_tot).  enable != 0 allocates one uint32 per zone and RK
 * stage and makes every ConsToPrim write each active zone's bits (0 when none)
 * and count the flagged zones; 0 frees them.  A Riemann solver's NaN report
 * (ERR_NAN | ERR_QUIT, fatal in the reference) is not recorded per zone.
 * GPL_EUNSUPPORTED for contexts without zone reports (none at present). */
int gpl_set_zone_reports(gpl_ctx *ctx, int enable);
/* After a stepping call: err[s][nz][ny][nx] = the bits ConsToPrim3D left in
 * stage s of the last step (s < 2 for RK2, s < 3 for RK3; may be NULL),
 * counts[s] = the zones flagged in stage s of the last step (may be NULL),
 * *total = the zones flagged since the reports were enabled (may be NULL).
 * Returns the number of stages, or GPL_EINVAL when the reports are off. */
int gpl_get_zone_reports(gpl_ctx *ctx, uint32_t *err, unsigned long long *counts,
                         unsigned long long *total);

/* Kernel timing with HIP events on the library's stream: enable, then query
 * the average duration of each kernel family over the recorded launches. */
int gpl_set_timing(gpl_ctx *ctx, int enable);
int gpl_kernel_count(const gpl_ctx *ctx);
int gpl_kernel_stats(gpl_ctx *ctx, int idx, const char **name, long long *launches,
                     double *total_ms);
/* GPU time of the last gpl_run call, from events around the whole loop. */
double gpl_last_run_ms(const gpl_ctx *ctx);

/* Bytes of device memory owned by the context. */
size_t gpl_device_bytes(const gpl_ctx *ctx);

const char *gpl_strerror(int code);

/* ---------------------------------------------------------------------------
 * Batched 2D MHD: many independent boxes of uneven sizes in one device
 * state, advanced together (one launch per phase over each of two groups
 * of boxes, the groups on two streams of the batch; one HIP graph per
 * step), each box with its own dt.  GPL_BATCH_STREAMS=n in the environment
 * at gpl_batch_create sets the number of groups (1..4).  No counterpart in the
 * reference, which runs one domain per process; each box computes exactly
 * what gpl_create/gpl_run would compute for it alone (FP64: bit for bit).
 * ------------------------------------------------------------------------- */
typedef struct gpl_batch gpl_batch;

/* cfg: physics GPL_MHD, dimensions 2; GPL_UCT_HLL, or GPL_UCT_HLLD with
 * GPL_SOLVER_HLLD.  Sides X1-beg .. X2-end periodic, outflow or reflective
 * (boundary / side_bc, the same for every box; a direction periodic on both
 * sides or on neither): with an outflow or reflective side every box is
 * stored with the reference's ghost layers, refilled at every stage.  cfg's
 * nx/ny/nz are ignored: box q is nx[q] x ny[q] zones on [xbeg, xend]
 * (3..8192 each; ny == NULL gives square boxes). */
int gpl_batch_create(const gpl_config *cfg, int nbox, const int *nx, const int *ny, gpl_batch **out);
int gpl_batch_destroy(gpl_batch *b);

/* Per-box state transfer, as gpl_upload/gpl_download with dimensions 2:
 * vc holds 8 variables (B3 cell-centred), vs[0..1] the B1/B2 faces; with
 * GPL_LAYOUT_ARY4D the upper boundary faces too (index n along x and y). */
int gpl_batch_upload(gpl_batch *b, int box, const double *vc, const double *const *vs, int layout);
int gpl_batch_download(gpl_batch *b, int box, double *vc, double *const *vs, double *uc, int layout);

/* nsteps steps of every box with per-box dt on the device, as gpl_run.
 * first_dt[nbox] restarts every box at time 0 (NULL: continue);
 * time_out/dt_out receive nbox values each (may be NULL). */
uint32_t gpl_batch_run(gpl_batch *b, int nsteps, const double *first_dt, double *time_out,
                       double *dt_out);

/* Launch shape: march = cells per lane of the y sweep (8, 16, 32) and
 * blocks_per_cu = persistent workgroups per CU (2, 4, 8, or 0 for one
 * workgroup per work-list group).  march 0 asks for the autotuner, which
 * (the default) times every shape on the first steps of the next run of at
 * least 24 steps and keeps the fastest; results do not depend on the shape. */
int gpl_batch_set_tuning(gpl_batch *b, int march, int blocks_per_cu);
int gpl_batch_get_tuning(const gpl_batch *b, int *march, int *blocks_per_cu, double *step_ms);
double gpl_batch_last_run_ms(const gpl_batch *b);
size_t gpl_batch_device_bytes(const gpl_batch *b);

/* ---------------------------------------------------------------------------
 * Learned con2prim initial guess for RMHD.  No counterpart in the reference,
 * whose RMHD_EnergySolve (Src/RMHD/rmhd_energy_solve.cpp:86-96) always starts
 * Newton from the Eq. (A27) root.  An 8-64-64-1 ReLU network of eight
 * scale-free invariants of the conserved state (weights: a float vector of
 * gpl_c2p_net_size() entries) corrects that guess; Newton, its tolerance and
 * the A27 restart on failure decide the result.
 * ------------------------------------------------------------------------- */
int gpl_c2p_net_size(void);
/* n (conserved, target) pairs from sampled primitive states: u[n][8]
 * (D, m1..3, B1..3, E - D), raw features x[n][8], y[n] = log(W/W_A27). */
int gpl_c2p_net_make_data(int device, uint64_t seed, int n, double gamma, double *u_out,
                          float *x_out, float *y_out);
/* Features x[n][8] and targets y[n] of given (conserved u[n][8], primitive
 * v[n][8]) pairs, e.g. the states of an earlier run (downloaded Uc, Vc). */
int gpl_c2p_net_pairs(int device, const double *u, const double *v, int n, double gamma,
                      float *x_out, float *y_out);
/* Train on (x, y) on the GPU (FP32 forward/backward kernels, Adam, the
 * samples shuffled by seed); writes the weights and the last mini-batch loss
 * (normalised target). */
int gpl_c2p_net_fit(int device, uint64_t seed, const float *x, const float *y, int n, int steps,
                    int batch, float lr, float *weights_out, float *loss_out);
/* gpl_c2p_net_make_data with nsamples sampled states, then gpl_c2p_net_fit. */
int gpl_c2p_net_train(int device, uint64_t seed, int nsamples, int steps, int batch, float lr,
                      double gamma, float *weights_out, float *loss_out);
/* Gradient of the mean-squared error over (x, y) -- the training kernel on
 * one batch -- and its loss. */
int gpl_c2p_net_grad(int device, const float *weights, const float *x, const float *y, int n,
                     float *grad_out, float *loss_out);
/* Matrix-core inference (BF16 MFMA, FP32 accumulate): y_out[n] = ysc * net(x),
 * the learned correction of log W_A27. */
int gpl_c2p_net_infer(int device, const float *weights, const float *x, int n, float *y_out);
/* Use the learned guess inside an RMHD context's con2prim (weights copied;
 * NULL restores the reference's A27 start, the bitwise drop-in).  With it on,
 * results agree with the drop-in to the Newton tolerance, not bit for bit. */
int gpl_set_c2p_network(gpl_ctx *ctx, const float *weights);
/* Batched RMHD energy solve of n states u[n][8] from the A27 guess
 * (weights == NULL, the reference method) or from the learned guess with
 * the A27 restart; v_out[n][8], Newton passes and error bits per state, the
 * kernel time, and (restart_out, may be NULL) 1 where the learned start
 * failed and the A27 restart ran. */
int gpl_c2p_solve(int device, const float *weights, const double *u, int n, double gamma,
                  double *v_out, int *passes_out, uint32_t *err_out, double *ms_out,
                  int *restart_out);

/* gpl_c2p_solve that allocates nothing.  No counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The states, results
 * and weights are staged in a fixed device array of the library (one per
 * device, about 9.2 MB, held until the process ends) instead of buffers
 * allocated per call, so n is at most GPL_C2P_NOALLOC_MAX (larger n:
 * GPL_ESCRATCH, nothing run).  The same kernel computes the same bits as
 * gpl_c2p_solve; there is no kernel time output.  Calls that overlap on one
 * device share that array and race; calls on different devices do not. */
#define GPL_C2P_NOALLOC_MAX 65536
int gpl_c2p_solve_noalloc(int device, const float *weights, const double *u, int n, double gamma,
                          double *v_out, int *passes_out, uint32_t *err_out, int *restart_out);

#ifdef __cplusplus
}
#endif
#endif /* GPLUTO_AMD_H */
