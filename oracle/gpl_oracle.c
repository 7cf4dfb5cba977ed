/* CPU oracle for the Godunov stage update -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference algorithm, used by tests/ and by
 * bench.py's cpu_baseline leg as the checker.  Nothing in the product path
 * links it.  Each function names the reference file:line it restates; the
 * order of floating-point operations inside each expression is kept as the
 * reference writes it (C evaluates a+b+c as (a+b)+c, the same as C++), and
 * the file is compiled with -ffp-contract=off, as the x86-64 reference
 * build (no FMA instructions without -march) effectively is.
 *
 * Pinning: see tests/test_oracle_golden.py (bitwise against the reference
 * binary's own output on committed fixtures) and tests/test_oracle_laws.py
 * (conservation, div B, symmetry, closed forms).
 */
#include "gpl_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- reference macros (Src/macros.hpp:12-14,191-201) ------------------ */
#define MAX(a,b)  ( (a) >= (b) ? (a) : (b) )
#define MIN(a,b)  ( (a) <= (b) ? (a) : (b) )
#define ABS_MIN(a,b)  (fabs(a) < fabs(b) ? (a):(b))
#define DSIGN(x)      ( (x) >= 0.0 ? (1.0) : (-1.0))

/* ---- error bits (Src/gpluto.hpp:247-262) ------------------------------- */
#define ERR_NEGATIVE_RHO    (1u<<0)
#define ERR_NEGATIVE_ENG    (1u<<1)
#define ERR_NEGATIVE_PRS    (1u<<2)
#define ERR_CONS2PRIM_FAIL  (1u<<4)
#define ERR_PRESSURE_FIX    (1u<<5)
#define ERR_ENG_SOLVE_NO_SOLUTION (1u<<7)
#define ERR_SUPERLUMINAL_SPEED    (1u<<8)
#define ERR_TOO_MANY_ITERATIONS   (1u<<9)
#define ERR_NAN             (1u<<11)
#define ERR_QUIT            (1u<<15)

/* ---- variable indices (Src/MHD/mhd.hpp:21-33, Src/HD/hd.hpp:21-28) ----- */
#define RHO 0
#define MX1 1
#define VX1 1
#define BX1 4
#define NVMAX 8

#define NG 2   /* Src/get_nghost.cpp:33, PLM */

struct gplo_state {
  gplo_config c;
  int nv;                 /* NVAR */
  int nt[3];              /* ntot per direction */
  int beg[3], end[3];     /* active range (inclusive) */
  int ns[3];              /* staggered array extents (nt + 1) */
  size_t ncell, nstag;
  double dx[3];
  double *Vc, *Uc, *U0, *Rc, *Cdt;
  double *Vs[3], *Bs0[3], *Rs[3];
  /* UCT face data, Src/MHD/CT/ct_emf.cpp:63-147 */
  double *e[3][3], *aL[3], *aR[3], *dL[3], *dR[3];
  double *Ee[3];
  /* pencil scratch, one set per thread (pstride doubles apart; pcmax and
     pBn pstride/NVMAX apart) */
  double *pv, *pvL, *pvR, *pflux, *pcmax, *pBn, *prhs;
  size_t pstride;
  int nthr;
  double *qv[2], *qL[2], *qR[2];
  int stage;
  /* Data::err (Src/structs.hpp:678) of the active zones as ConsToPrim3D
     leaves it in each stage of the last step ([3][nz][ny][nx]), and the
     flagged-zone counts Where() adds to Data::err_count
     (Src/debug_tools.cpp:587-593; reset per step, Src/main.cpp:423) */
  uint32_t *zerr;
  unsigned long long zcount[3], zcount_total;
};

#define IC(s,i,j,k)  (((size_t)(k)*(s)->nt[1] + (size_t)(j))*(s)->nt[0] + (size_t)(i))
#define IS(s,i,j,k)  (((size_t)(k)*(s)->ns[1] + (size_t)(j))*(s)->ns[0] + (size_t)(i))
#define VC(s,nv,i,j,k)  (s)->Vc[(size_t)(nv)*(s)->ncell + IC(s,i,j,k)]
#define UC(s,nv,i,j,k)  (s)->Uc[(size_t)(nv)*(s)->ncell + IC(s,i,j,k)]

/* staggered field + CT (MHD and RMHD) */
static int is_mhd(const gplo_state *s) { return s->c.phys == GPLO_MHD || s->c.phys == GPLO_RMHD; }
static int is_rmhd(const gplo_state *s) { return s->c.phys == GPLO_RMHD; }

/* ======================================================================
 * Pointwise physics
 * ====================================================================== */

/* MHD PrimToCons, Src/MHD/mhd.hpp:906-934 */
static void mhd_prim_to_cons(const double *v, double *u, double gamma)
{
  double gmm1 = gamma - 1.0;
  u[RHO] = v[RHO];
  u[1] = v[RHO]*v[1];
  u[2] = v[RHO]*v[2];
  u[3] = v[RHO]*v[3];
  u[4] = v[4];
  u[5] = v[5];
  u[6] = v[6];
  double kinb2 = v[1]*v[1] + v[2]*v[2] + v[3]*v[3];
  kinb2 = v[RHO]*kinb2 + v[4]*v[4] + v[5]*v[5] + v[6]*v[6];
  kinb2 *= 0.5;
  u[7] = kinb2 + v[7]/gmm1;
}

/* HD PrimToCons, Src/HD/hd.hpp:392-410 */
static void hd_prim_to_cons(const double *v, double *u, double gamma)
{
  u[RHO] = v[RHO];
  u[1] = v[RHO]*v[1];
  u[2] = v[RHO]*v[2];
  u[3] = v[RHO]*v[3];
  double kin = v[1]*v[1] + v[2]*v[2] + v[3]*v[3];
  kin = v[RHO]*kin*0.5;
  u[4] = kin + v[4]/(gamma - 1.0);
}

/* MHD Flux<dir>, Src/MHD/mhd.hpp:164-218 (no background field) */
static void mhd_flux(const double *v, const double *u, int n, double *flux, double *pt)
{
  const int VXn = 1 + n, BXn = 4 + n;
  double vB = v[1]*v[4] + v[2]*v[5] + v[3]*v[6];
  *pt = v[7] + 0.5*(v[4]*v[4] + v[5]*v[5] + v[6]*v[6]);
  flux[RHO] = u[1 + n];
  flux[1] = u[1]*v[VXn] - v[4]*v[BXn];
  flux[2] = u[2]*v[VXn] - v[5]*v[BXn];
  flux[3] = u[3]*v[VXn] - v[6]*v[BXn];
  flux[4] = v[4]*v[VXn] - v[BXn]*v[1];
  flux[5] = v[5]*v[VXn] - v[BXn]*v[2];
  flux[6] = v[6]*v[VXn] - v[BXn]*v[3];
  flux[7] = (u[7] + *pt)*v[VXn] - v[BXn]*vB;
  flux[1 + n] += *pt;
}

/* HD Flux<dir>, Src/HD/hd.hpp:105-125 */
static void hd_flux(const double *v, const double *u, int n, double *flux, double *p)
{
  const int VXn = 1 + n;
  flux[RHO] = u[1 + n];
  flux[1] = u[1]*v[VXn];
  flux[2] = u[2]*v[VXn];
  flux[3] = u[3]*v[VXn];
  flux[4] = (u[4] + v[4])*v[VXn];
  *p = v[4];
  flux[1 + n] += *p;
}

/* SoundSpeed2, Src/eos.hpp:84-99 (ideal EOS, classical) */
static double sound_speed2(const double *v, double gamma, int prs)
{
  return gamma*v[prs]/v[RHO];
}

/* MHD MaxSignalSpeed<dir>, Src/MHD/mhd.hpp:242-280 */
static void mhd_max_signal_speed(const double *v, double a2, int n, double *cmin, double *cmax)
{
  const int VXn = 1 + n;
  const int BXn = 4 + n, BXt = 4 + (n + 1)%3, BXb = 4 + (n + 2)%3;
  double gpr = a2*v[RHO];
  double b1 = v[BXn], b2 = v[BXt], b3 = v[BXb];
  double Btmag2 = b2*b2 + b3*b3;
  double Bmag2  = b1*b1 + Btmag2;
  double cf = gpr - Bmag2;
  cf = gpr + Bmag2 + sqrt(cf*cf + 4.0*gpr*Btmag2);
  cf = sqrt(0.5*cf/v[RHO]);
  *cmax = v[VXn] + cf;
  *cmin = v[VXn] - cf;
}

/* HD MaxSignalSpeed<dir>, Src/HD/hd.hpp:141-150 */
static void hd_max_signal_speed(const double *v, double a2, int n, double *cmin, double *cmax)
{
  const double a = sqrt(a2);
  *cmax = v[1 + n] + a;
  *cmin = v[1 + n] - a;
}

/* ======================================================================
 * Riemann solvers.  Each returns the flux (NFLX components) and the
 * Davis speeds Sl, Sr; cmax(i) = MAX(|Sl|,|Sr|) (Src/MHD/hlld.cpp:178).
 * ====================================================================== */

/* HLLD, Src/MHD/hlld.cpp:62-500 (ideal EOS, no GLM, no background field) */
static uint32_t mhd_hlld(const double *vl, const double *vr, int n, double gamma,
                         double *flux, double *pSl, double *pSr)
{
  const int VXn = 1 + n, VXt = 1 + (n + 1)%3, VXb = 1 + (n + 2)%3;
  const int MXn = VXn, MXt = VXt, MXb = VXb;
  const int BXn = 4 + n, BXt = 4 + (n + 1)%3, BXb = 4 + (n + 2)%3;
  const int NFLX = 8, ENG = 7;
  int nv, switch_to_hll = 0;
  double ul[NVMAX], ur[NVMAX], fl[NVMAX], fr[NVMAX], usl[NVMAX], usr[NVMAX];
  double Uhll[NVMAX], ptl, ptr, Sl, Sr, sl_min, sl_max, sr_min, sr_max;

  mhd_prim_to_cons(vl, ul, gamma);
  mhd_prim_to_cons(vr, ur, gamma);
  double a2L = sound_speed2(vl, gamma, 7);
  double a2R = sound_speed2(vr, gamma, 7);
  mhd_max_signal_speed(vl, a2L, n, &sl_min, &sl_max);   /* HLL_Speed, mhd.hpp:296 */
  mhd_max_signal_speed(vr, a2R, n, &sr_min, &sr_max);
  Sl = MIN(sl_min, sr_min);
  Sr = MAX(sl_max, sr_max);
  *pSl = Sl; *pSr = Sr;

  mhd_flux(vl, ul, n, fl, &ptl);
  mhd_flux(vr, ur, n, fr, &ptr);

  double dul, dur, Bx, Bx1, sBx, pts, SM;
  {
    double scrh = 1.0/(Sr - Sl);
    Bx1 = Bx = (Sr*vr[BXn] - Sl*vl[BXn])*scrh;
    sBx = (Bx > 0.0 ? 1.0 : -1.0);
    dul = Sl - vl[VXn];
    dur = Sr - vr[VXn];
    scrh = 1.0/(dur*ur[RHO] - dul*ul[RHO]);
    SM = (dur*ur[MXn] - dul*ul[MXn] - ptr + ptl)*scrh;
    pts =   dur*ur[RHO]*ptl - dul*ul[RHO]*ptr
          + vl[RHO]*vr[RHO]*dur*dul*(vr[VXn]- vl[VXn]);
    pts *= scrh;
  }
  usl[RHO] = ul[RHO]*dul/(Sl - SM);
  usr[RHO] = ur[RHO]*dur/(Sr - SM);
  if (usl[RHO] < 0.0 || usr[RHO] < 0.0) switch_to_hll = 1;
  double sqrL = sqrt(usl[RHO]);
  double sqrR = sqrt(usr[RHO]);
  double S1L = SM - fabs(Bx)/sqrL;
  double S1R = SM + fabs(Bx)/sqrR;

  if (Sl >= 0.0) {
    for (nv = 0; nv < NFLX; nv++) flux[nv] = fl[nv];
  } else if (Sr <= 0.0) {
    for (nv = 0; nv < NFLX; nv++) flux[nv] = fr[nv];
  } else {
    if (!switch_to_hll) {
      int switch_to_hllc = 0;
      if ( (S1L - Sl) <  1.e-4*(SM - Sl) ) switch_to_hllc = 1;
      if ( (S1R - Sr) > -1.e-4*(Sr - SM) ) switch_to_hllc = 1;
      double scrhL, scrhR;
      if (!switch_to_hllc) {
        scrhL = (ul[RHO]*dul*dul - Bx*Bx)/(ul[RHO]*dul*(Sl - SM) - Bx*Bx);
        scrhR = (ur[RHO]*dur*dur - Bx*Bx)/(ur[RHO]*dur*(Sr - SM) - Bx*Bx);
        usl[BXn] = Bx1;
        usl[BXt] = ul[BXt]*scrhL;
        usl[BXb] = ul[BXb]*scrhL;
        usr[BXn] = Bx1;
        usr[BXt] = ur[BXt]*scrhR;
        usr[BXb] = ur[BXb]*scrhR;
      } else {
        double scrh = 1.0/(Sr - Sl);
        for (nv = 0; nv < NFLX; nv++) {
          Uhll[nv]  = Sr*ur[nv] - Sl*ul[nv] + fl[nv] - fr[nv];
          Uhll[nv] *= scrh;
        }
        usl[BXn] = usr[BXn] = Bx1;
        usl[BXt] = usr[BXt] = Uhll[BXt];
        usl[BXb] = usr[BXb] = Uhll[BXb];
        S1L = S1R = SM;
      }
      scrhL = Bx/(ul[RHO]*dul);
      scrhR = Bx/(ur[RHO]*dur);
      double vsL = vl[VXt] - scrhL*(usl[BXt] - ul[BXt]);
      double vsR = vr[VXt] - scrhR*(usr[BXt] - ur[BXt]);
      double wsL = vl[VXb] - scrhL*(usl[BXb] - ul[BXb]);
      double wsR = vr[VXb] - scrhR*(usr[BXb] - ur[BXb]);
      usl[MXn] = usl[RHO]*SM;
      usr[MXn] = usr[RHO]*SM;
      usl[MXt] = usl[RHO]*vsL;
      usr[MXt] = usr[RHO]*vsR;
      usl[MXb] = usl[RHO]*wsL;
      usr[MXb] = usr[RHO]*wsR;
      scrhL     = vl[VXn]*Bx1 + vl[VXt]*ul[BXt] + vl[VXb]*ul[BXb];
      scrhL    -= SM*Bx1      +    vsL*usl[BXt] + wsL*usl[BXb];
      usl[ENG]  = dul*ul[ENG] - ptl*vl[VXn] + pts*SM + Bx*scrhL;
      usl[ENG] /= Sl - SM;
      scrhR     = vr[VXn]*Bx1 + vr[VXt]*ur[BXt] + vr[VXb]*ur[BXb];
      scrhR    -= SM*Bx1      +    vsR*usr[BXt] +    wsR*usr[BXb];
      usr[ENG]  = dur*ur[ENG] - ptr*vr[VXn] + pts*SM + Bx*scrhR;
      usr[ENG] /= Sr - SM;

      if (S1L >= 0.0) {
        for (nv = 0; nv < NFLX; nv++) flux[nv] = fl[nv] + Sl*(usl[nv] - ul[nv]);
      } else if (S1R <= 0.0) {
        for (nv = 0; nv < NFLX; nv++) flux[nv] = fr[nv] + Sr*(usr[nv] - ur[nv]);
      } else {
        double ussl[NVMAX], ussr[NVMAX];
        ussl[RHO] = usl[RHO];
        ussr[RHO] = usr[RHO];
        double vss  = sqrL*vsL + sqrR*vsR + (usr[BXt] - usl[BXt])*sBx;
        vss /= sqrL + sqrR;
        double wss  = sqrL*wsL + sqrR*wsR + (usr[BXb] - usl[BXb])*sBx;
        wss /= sqrL + sqrR;
        ussl[MXn] = ussl[RHO]*SM;
        ussr[MXn] = ussr[RHO]*SM;
        ussl[MXt] = ussl[RHO]*vss;
        ussr[MXt] = ussr[RHO]*vss;
        ussl[MXb] = ussl[RHO]*wss;
        ussr[MXb] = ussr[RHO]*wss;
        ussl[BXn] = ussr[BXn] = Bx1;
        ussl[BXt]  = sqrL*usr[BXt] + sqrR*usl[BXt] + sqrL*sqrR*(vsR - vsL)*sBx;
        ussl[BXt] /= sqrL + sqrR;
        ussr[BXt]  = ussl[BXt];
        ussl[BXb]  = sqrL*usr[BXb] + sqrR*usl[BXb] + sqrL*sqrR*(wsR - wsL)*sBx;
        ussl[BXb] /= sqrL + sqrR;
        ussr[BXb]  = ussl[BXb];
        scrhL  = SM*Bx1 + vsL*usl [BXt] + wsL*usl [BXb];
        scrhL -= SM*Bx1 + vss*ussl[BXt] + wss*ussl[BXb];
        scrhR  = SM*Bx1 + vsR*usr [BXt] + wsR*usr [BXb];
        scrhR -= SM*Bx1 + vss*ussr[BXt] + wss*ussr[BXb];
        ussl[ENG] = usl[ENG] - sqrL*scrhL*sBx;
        ussr[ENG] = usr[ENG] + sqrR*scrhR*sBx;
        if (SM >= 0.0) {
          for (nv = 0; nv < NFLX; nv++)
            flux[nv] = fl[nv] + S1L*(ussl[nv] - usl[nv]) + Sl*(usl[nv] - ul[nv]);
        } else {
          for (nv = 0; nv < NFLX; nv++)
            flux[nv] = fr[nv] + S1R*(ussr[nv] - usr[nv]) + Sr*(usr[nv] - ur[nv]);
        }
      }
    } else {
      double scrh = 1.0/(Sr - Sl);
      for (nv = 0; nv < NFLX; nv++) {
        flux[nv]  = Sl*Sr*(ur[nv] - ul[nv]) + Sr*fl[nv] - Sl*fr[nv];
        flux[nv] *= scrh;
      }
    }
  }
  uint32_t err = 0;
  for (nv = 0; nv < NFLX; nv++) if (isnan(flux[nv])) err |= ERR_NAN | ERR_QUIT;
  return err;
}

/* MHD HLLC of Li (2005), Src/MHD/hllc.cpp:130-335 (ideal EOS, no GLM, no
   background field; HLL_SWITCH is off in the default configuration,
   Src/gpluto.hpp:877-878) */
static uint32_t mhd_hllc(const double *vl, const double *vr, int n, double gamma,
                         double *flux, double *pSl, double *pSr)
{
  const int VXn = 1 + n;
  const int MXn = VXn, MXt = 1 + (n + 1)%3, MXb = 1 + (n + 2)%3;
  const int BXn = 4 + n, BXt = 4 + (n + 1)%3, BXb = 4 + (n + 2)%3;
  const int NFLX = 8, ENG = 7;
  int nv;
  double ul[NVMAX], ur[NVMAX], fl[NVMAX], fr[NVMAX], usl[NVMAX], usr[NVMAX];
  double Uhll[NVMAX], Fhll[NVMAX], pl, pr, Sl, Sr, sl_min, sl_max, sr_min, sr_max;

  mhd_prim_to_cons(vl, ul, gamma);
  mhd_prim_to_cons(vr, ur, gamma);
  double a2L = sound_speed2(vl, gamma, 7);
  double a2R = sound_speed2(vr, gamma, 7);
  mhd_max_signal_speed(vl, a2L, n, &sl_min, &sl_max);   /* HLL_Speed, mhd.hpp:296 */
  mhd_max_signal_speed(vr, a2R, n, &sr_min, &sr_max);
  Sl = MIN(sl_min, sr_min);
  Sr = MAX(sl_max, sr_max);
  *pSl = Sl; *pSr = Sr;

  mhd_flux(vl, ul, n, fl, &pl);      /* pl, pr: total pressure */
  mhd_flux(vr, ur, n, fr, &pr);

  if (Sl >= 0.0) {
    for (nv = 0; nv < NFLX; nv++) flux[nv] = fl[nv];
  } else if (Sr <= 0.0) {
    for (nv = 0; nv < NFLX; nv++) flux[nv] = fr[nv];
  } else {
    double scrh = 1.0/(Sr - Sl);
    for (nv = 0; nv < NFLX; nv++) {
      Uhll[nv]  = Sr*ur[nv] - Sl*ul[nv] + fl[nv] - fr[nv];
      Uhll[nv] *= scrh;
      Fhll[nv]  =   Sl*Sr*(ur[nv] - ul[nv])
                  + Sr*fl[nv] - Sl*fr[nv];
      Fhll[nv] *= scrh;
    }
    double Bxs = Uhll[BXn];
    double Bys = Uhll[BXt];
    double Bzs = Uhll[BXb];
    double vxL = vl[VXn];
    double vxR = vr[VXn];
    double vxs = Uhll[MXn]/Uhll[RHO];
    {
      double vBL = vl[1]*vl[4] + vl[2]*vl[5] + vl[3]*vl[6];
      double vBR = vr[1]*vr[4] + vr[2]*vr[5] + vr[3]*vr[6];
      double vBs = Uhll[4]*Uhll[1]
                 + Uhll[5]*Uhll[2]
                 + Uhll[6]*Uhll[3];
      vBs /= Uhll[RHO];
      usl[RHO] = ul[RHO]*(Sl - vxL)/(Sl - vxs);
      usr[RHO] = ur[RHO]*(Sr - vxR)/(Sr - vxs);
      double ps = Fhll[MXn] + Bxs*Bxs - Fhll[RHO]*vxs;
      usl[ENG] = (ul[ENG]*(Sl - vxL) +
                  ps*vxs - pl*vxL - Bxs*vBs + vl[BXn]*vBL)/(Sl - vxs);
      usr[ENG] = (ur[ENG]*(Sr - vxR) +
                  ps*vxs - pr*vxR - Bxs*vBs + vr[BXn]*vBR)/(Sr - vxs);
    }
    usl[MXn] = usl[RHO]*vxs;
    usr[MXn] = usr[RHO]*vxs;
    usl[MXt] =   (ul[MXt]*(Sl - vxL)
               - (Bxs*Bys - vl[BXn]*vl[BXt]))/(Sl - vxs);
    usr[MXt] =   (ur[MXt]*(Sr - vxR)
               - (Bxs*Bys - vr[BXn]*vr[BXt]))/(Sr - vxs);
    usl[MXb] =   (ul[MXb]*(Sl - vxL)
               - (Bxs*Bzs - vl[BXn]*vl[BXb]))/(Sl - vxs);
    usr[MXb] =   (ur[MXb]*(Sr - vxR)
               - (Bxs*Bzs - vr[BXn]*vr[BXb]))/(Sr - vxs);
    usl[BXn] = usr[BXn] = Bxs;
    usl[BXt] = usr[BXt] = Bys;
    usl[BXb] = usr[BXb] = Bzs;
    if (vxs > 0.0) {
      for (nv = 0; nv < NFLX; nv++) flux[nv] = fl[nv] + Sl*(usl[nv] - ul[nv]);
    } else {
      for (nv = 0; nv < NFLX; nv++) flux[nv] = fr[nv] + Sr*(usr[nv] - ur[nv]);
    }
  }
  uint32_t err = 0;
  for (nv = 0; nv < NFLX; nv++) if (isnan(flux[nv])) err |= ERR_NAN | ERR_QUIT;
  return err;
}

/* HD HLLC, Src/HD/hllc.cpp:54-250 (ideal EOS) */
static uint32_t hd_hllc(const double *vl, const double *vr, int n, double gamma,
                        double *flux, double *pSl, double *pSr)
{
  const int VXn = 1 + n, VXt = 1 + (n + 1)%3, VXb = 1 + (n + 2)%3;
  const int MXn = VXn, MXt = VXt, MXb = VXb;
  const int NFLX = 5, ENG = 4, PRS = 4;
  int nv;
  double ul[NVMAX], ur[NVMAX], fL[NVMAX], fR[NVMAX], usl[NVMAX], usr[NVMAX];
  double pl, pr, Sl, Sr, sl_min, sl_max, sr_min, sr_max;
  hd_prim_to_cons(vl, ul, gamma);
  hd_prim_to_cons(vr, ur, gamma);
  double a2L = sound_speed2(vl, gamma, PRS);
  double a2R = sound_speed2(vr, gamma, PRS);
  hd_flux(vl, ul, n, fL, &pl);
  hd_flux(vr, ur, n, fR, &pr);
  hd_max_signal_speed(vl, a2L, n, &sl_min, &sl_max);  /* HLL_Speed, hd.hpp:165 */
  hd_max_signal_speed(vr, a2R, n, &sr_min, &sr_max);
  Sl = MIN(sl_min, sr_min);
  Sr = MAX(sl_max, sr_max);
  *pSl = Sl; *pSr = Sr;
  if (Sl > 0.0) {
    for (nv = 0; nv < NFLX; nv++) flux[nv] = fL[nv];
  } else if (Sr < 0.0) {
    for (nv = 0; nv < NFLX; nv++) flux[nv] = fR[nv];
  } else {
    double vs;
    {
      double qL = vl[PRS] + ul[MXn]*(vl[VXn] - Sl);
      double qR = vr[PRS] + ur[MXn]*(vr[VXn] - Sr);
      double wL = vl[RHO]*(vl[VXn] - Sl);
      double wR = vr[RHO]*(vr[VXn] - Sr);
      vs = (qR - qL)/(wR - wL);
    }
    {
      double vxr = vr[VXn];
      double vxl = vl[VXn];
      usl[RHO] = ul[RHO]*(Sl - vxl)/(Sl - vs);
      usr[RHO] = ur[RHO]*(Sr - vxr)/(Sr - vs);
      usl[MXn] = usl[RHO]*vs;      usr[MXn] = usr[RHO]*vs;
      usl[MXt] = usl[RHO]*vl[VXt]; usr[MXt] = usr[RHO]*vr[VXt];
      usl[MXb] = usl[RHO]*vl[VXb]; usr[MXb] = usr[RHO]*vr[VXb];
      usl[ENG] =   ul[ENG]/vl[RHO]
                + (vs - vxl)*(vs + vl[PRS]/(vl[RHO]*(Sl - vxl)));
      usr[ENG] =   ur[ENG]/vr[RHO]
                + (vs - vxr)*(vs + vr[PRS]/(vr[RHO]*(Sr - vxr)));
    }
    usl[ENG] *= usl[RHO];
    usr[ENG] *= usr[RHO];
    if (vs >= 0.0) {
      for (nv = 0; nv < NFLX; nv++) flux[nv] = fL[nv] + Sl*(usl[nv] - ul[nv]);
    } else {
      for (nv = 0; nv < NFLX; nv++) flux[nv] = fR[nv] + Sr*(usr[nv] - ur[nv]);
    }
  }
  uint32_t err = 0;
  for (nv = 0; nv < NFLX; nv++) if (isnan(flux[nv])) err |= ERR_NAN | ERR_QUIT;
  return err;
}

/* HLL (MHD and HD), Src/MHD/hll.cpp, Src/HD/hll.cpp: flux formula as in
   the HLL branch of Src/MHD/hlld.cpp:452-461 */
static uint32_t hll(const double *vl, const double *vr, int n, double gamma, int mhd,
                    double *flux, double *pSl, double *pSr)
{
  int nv, NFLX = mhd ? 8 : 5, PRS = mhd ? 7 : 4;
  double ul[NVMAX], ur[NVMAX], fl[NVMAX], fr[NVMAX], pl, pr;
  double sl_min, sl_max, sr_min, sr_max, Sl, Sr;
  if (mhd) { mhd_prim_to_cons(vl, ul, gamma); mhd_prim_to_cons(vr, ur, gamma); }
  else     { hd_prim_to_cons(vl, ul, gamma);  hd_prim_to_cons(vr, ur, gamma); }
  double a2L = sound_speed2(vl, gamma, PRS);
  double a2R = sound_speed2(vr, gamma, PRS);
  if (mhd) {
    mhd_flux(vl, ul, n, fl, &pl); mhd_flux(vr, ur, n, fr, &pr);
    mhd_max_signal_speed(vl, a2L, n, &sl_min, &sl_max);
    mhd_max_signal_speed(vr, a2R, n, &sr_min, &sr_max);
  } else {
    hd_flux(vl, ul, n, fl, &pl); hd_flux(vr, ur, n, fr, &pr);
    hd_max_signal_speed(vl, a2L, n, &sl_min, &sl_max);
    hd_max_signal_speed(vr, a2R, n, &sr_min, &sr_max);
  }
  Sl = MIN(sl_min, sr_min);
  Sr = MAX(sl_max, sr_max);
  *pSl = Sl; *pSr = Sr;
  if (Sl > 0.0) {
    for (nv = 0; nv < NFLX; nv++) flux[nv] = fl[nv];
  } else if (Sr < 0.0) {
    for (nv = 0; nv < NFLX; nv++) flux[nv] = fr[nv];
  } else {
    double scrh = 1.0/(Sr - Sl);
    for (nv = 0; nv < NFLX; nv++) {
      flux[nv]  = Sl*Sr*(ur[nv] - ul[nv]) + Sr*fl[nv] - Sl*fr[nv];
      flux[nv] *= scrh;
    }
  }
  uint32_t err = 0;
  for (nv = 0; nv < NFLX; nv++) if (isnan(flux[nv])) err |= ERR_NAN | ERR_QUIT;
  return err;
}

/* ======================================================================
 * RMHD (ideal EOS, RMHD_REDUCED_ENERGY = YES, RMHD_FAST_EIGENVALUES = NO,
 * no PCP modules): variables rho, v1..3, B1..3, prs; conserved D, m1..3,
 * B1..3, E - D.
 * ====================================================================== */

/* Enthalpy, Src/eos.hpp:13-51 (relativistic, IDEAL) */
static double rmhd_enthalpy(const double *v, double gamma)
{
  double gmmr  = gamma/(gamma - 1.0);
  double theta = v[7]/v[RHO];
  return 1.0 + gmmr*theta;
}

/* SoundSpeed2, Src/eos.hpp:84-125 (relativistic, IDEAL): a2 and h */
static void rmhd_sound_speed2(const double *v, double gamma, double *a2, double *h)
{
  double hh = rmhd_enthalpy(v, gamma);
  double theta = v[7]/v[RHO];
  *h  = hh;
  *a2 = gamma*theta/hh;
}

/* PrimToCons, Src/RMHD/rmhd.hpp:535-610 (B x (v x B) momentum form) */
static void rmhd_prim_to_cons(const double *v, double *u, double gamma)
{
  double gmmr = gamma/(gamma - 1.0);
  double h = rmhd_enthalpy(v, gamma);
  double vx1 = v[1], vx2 = v[2], vx3 = v[3];
  double vel2 = vx1*vx1 + vx2*vx2 + vx3*vx3;
  double g2 = 1.0/(1.0 - vel2);
  double g  = sqrt(g2);
  double ex = vx2*v[6] - vx3*v[5];
  double ey = vx3*v[4] - vx1*v[6];
  double ez = vx1*v[5] - vx2*v[4];
  double e2 = ex*ex + ey*ey + ez*ez;
  double Bmag2 = v[4]*v[4] + v[5]*v[5] + v[6]*v[6];
  double wt = v[RHO]*h*g2;
  u[RHO] = g*v[RHO];
  u[1] = wt*vx1 + v[5]*ez - v[6]*ey;
  u[2] = wt*vx2 + v[6]*ex - v[4]*ez;
  u[3] = wt*vx3 + v[4]*ey - v[5]*ex;
  u[4] = v[4];
  u[5] = v[5];
  u[6] = v[6];
  u[7] =   v[7]*(g2*gmmr - 1.0) + u[RHO]*g2*vel2/(g + 1.0)
         + 0.5*(Bmag2 + e2);
}

/* Flux<dir>, Src/RMHD/rmhd.hpp:315-366 (reduced energy) */
static void rmhd_flux(const double *v, const double *u, double h, int n, double *flux, double *pt)
{
  const int VXn = 1 + n, VXt = 1 + (n + 1)%3, VXb = 1 + (n + 2)%3;
  const int BXn = 4 + n, BXt = 4 + (n + 1)%3, BXb = 4 + (n + 2)%3;
  double g  = u[RHO]/v[RHO];
  double g2 = g*g;
  double g_2 = 1.0/g2;
  double Bmag2 = v[4]*v[4] + v[5]*v[5] + v[6]*v[6];
  double vB = v[1]*v[4] + v[2]*v[5] + v[3]*v[6];
  double b[3];
  b[0] = g*(v[4]*g_2 + vB*v[1]);
  b[1] = g*(v[5]*g_2 + vB*v[2]);
  b[2] = g*(v[6]*g_2 + vB*v[3]);
  double b2 = Bmag2*g_2 + vB*vB;
  *pt = v[7] + 0.5*b2;
  double wt = v[RHO]*h + b2;
  double wtg2 = wt*g2;
  flux[RHO] = u[RHO]*v[VXn];
  flux[1] = wtg2*v[1]*v[VXn] - b[0]*b[n];
  flux[2] = wtg2*v[2]*v[VXn] - b[1]*b[n];
  flux[3] = wtg2*v[3]*v[VXn] - b[2]*b[n];
  flux[BXn] = 0.0;
  flux[BXt] = v[VXn]*v[BXt] - v[BXn]*v[VXt];
  flux[BXb] = v[VXn]*v[BXb] - v[BXn]*v[VXb];
  flux[7] = u[1 + n];
  flux[7] -= flux[RHO];
  flux[1 + n] += *pt;
}

/* SortArray, Src/RMHD/rmhd.hpp:370-387 (insertion sort) */
static void sort_array(double *z, int n)
{
  for (int j = 1; j < n; j++) {
    double f = z[j];
    int i = j - 1;
    while (i >= 0 && z[i] > f) { z[i+1] = z[i]; i--; }
    z[i+1] = f;
  }
}

/* acos and sin/cos of the resolvent cubic: the host libm unless a test
   substitutes other implementations (gplo_set_cubic_math), to attribute a
   difference to the math library alone */
static double (*cubic_acos)(double) = NULL;
static void (*cubic_sincos)(double, double *, double *) = NULL;

void gplo_set_cubic_math(double (*acos_fn)(double), void (*sincos_fn)(double, double *, double *))
{
  cubic_acos = acos_fn;
  cubic_sincos = sincos_fn;
}

/* CubicSolve, Src/RMHD/rmhd.hpp:390-452 (three real roots, trigonometric form) */
static void cubic_solve(double b, double c, double d, double z[3])
{
  const double one_3 = 1.0/3.0, one_27 = 1.0/27.0, sqrt3 = sqrt(3.0);
  double b2 = b*b;
  double Q  = b2*one_3 - c*(1.0 - 1.e-16);
  double R  = b*(2.0*b2 - 9.0*c)*one_27 + d;
  Q = MAX(Q, 0.0);
  double sQ  = sqrt(Q)/sqrt3;
  double arg = -1.5*R/(Q*sQ);
  arg = MAX(-1.0, arg);
  arg = MIN( 1.0, arg);
  double theta = (cubic_acos ? cubic_acos(arg) : acos(arg))*one_3;
  double cs, sn;
  if (cubic_sincos) {
    double s1;
    cubic_sincos(theta, &s1, &cs);
    sn = sqrt3*s1;
  } else {
    cs = cos(theta);
    sn = sqrt3*sin(theta);
  }
  double p  = -b*one_3;
  z[0] = -sQ*(cs + sn) + p;
  z[1] = -sQ*(cs - sn) + p;
  z[2] =  2.0*sQ*cs    + p;
}

/* QuarticSolve, Src/RMHD/rmhd.hpp:455-505 (resolvent cubic) */
static int quartic_solve(double b, double c, double d, double e, double *z)
{
  double b2 = b*b;
  double p = c - b2*0.375;
  double q = d + b2*b*0.125 - b*c*0.5;
  double r = e - 3.0*b2*b2/256.0 + b2*c/16.0 - 0.25*b*d;
  double a2 = 2.0*p;
  double a1 = p*p - 4.0*r;
  double a0 = -q*q;
  double u[3];
  cubic_solve(a2, a1, a0, u);
  u[0] = MAX(u[0],0.0);
  u[1] = MAX(u[1],0.0);
  u[2] = MAX(u[2],0.0);
  if (u[0] != u[0] || u[1] != u[1] || u[2] != u[2]) return 1;
  double sq  = -0.5*DSIGN(q);
  double sz1 = sq*sqrt(u[0]);
  double sz2 = sq*sqrt(u[1]);
  double sz3 = sq*sqrt(u[2]);
  z[0] = -0.25*b + sz1 + sz2 + sz3;
  z[1] = -0.25*b + sz1 - sz2 - sz3;
  z[2] = -0.25*b - sz1 + sz2 - sz3;
  z[3] = -0.25*b - sz1 - sz2 + sz3;
  sort_array(z, 4);
  return 0;
}

/* MaxSignalSpeed<dir>, Src/RMHD/rmhd.hpp:631-752 (exact fast speeds) */
static void rmhd_max_signal_speed(const double *v, double cs2, double h, int n,
                                  double *cmin, double *cmax)
{
  const int VXn = 1 + n, BXn = 4 + n;
  double lfp, lfm;
  double vB  = v[1]*v[4] + v[2]*v[5] + v[3]*v[6];
  double u02 = v[1]*v[1] + v[2]*v[2] + v[3]*v[3];
  double Bm2 = v[4]*v[4] + v[5]*v[5] + v[6]*v[6];
  if (u02 < 1.e-14) {
    /* zero velocity: biquadratic */
    double w_1  = 1.0/(v[RHO]*h + Bm2);
    double eps2 = cs2 + Bm2*w_1*(1.0 - cs2);
    double a0   = cs2*v[BXn]*v[BXn]*w_1;
    double a1   = - a0 - eps2;
    double del  = a1*a1 - 4.0*a0;
    del = MAX(del, 0.0);
    del = sqrt(del);
    lfp =  sqrt(0.5*(-a1 + del));
    lfm = -lfp;
  } else {
    double vB2 = vB*vB;
    u02 = 1.0/(1.0 - u02);
    double b2  = Bm2/u02 + vB2;
    double w_1 = 1.0/(v[RHO]*h + b2);
    double vx  = v[VXn];
    double vx2 = vx*vx;
    if (fabs(v[BXn]) < 1.e-14) {
      /* zero normal field: quadratic */
      double w     = v[RHO]*h;
      double scrh1 = b2 + cs2*(w - vB2);
      double scrh2 = w*(1.0 - cs2)*u02;
      double del   = 4.0*scrh1*(scrh2*(1.0 - vx2) + scrh1);
      double a2 = scrh1 + scrh2;
      double a1 = -2.0*vx*scrh2;
      lfp = 0.5*(-a1 + sqrt(del))/a2;
      lfm = 0.5*(-a1 - sqrt(del))/a2;
    } else {
      /* general case: quartic */
      double scrh1 = v[BXn]/u02 + vB*vx;
      double scrh2 = scrh1*scrh1;
      double a2_w       = cs2*w_1;
      double eps2       = (cs2*v[RHO]*h + b2)*w_1;
      double one_m_eps2 = u02*v[RHO]*h*(1.0 - cs2)*w_1;
      double scrh = 2.0*(a2_w*vB*scrh1 - eps2*vx);
      double a4 = one_m_eps2 - a2_w*vB2 + eps2;
      double a3 = - 4.0*vx*one_m_eps2 + scrh;
      double a2 =   6.0*vx2*one_m_eps2 + a2_w*(vB2 - scrh2) + eps2*(vx2 - 1.0);
      double a1 = - 4.0*vx*vx2*one_m_eps2 - scrh;
      double a0 = vx2*vx2*one_m_eps2 + a2_w*scrh2 - eps2*vx2;
      scrh = 1.0/a4;
      a3 *= scrh;
      a2 *= scrh;
      a1 *= scrh;
      a0 *= scrh;
      double z[4] = {NAN, NAN, NAN, NAN};
      (void)quartic_solve(a3, a2, a1, a0, z);
      lfm = z[0];
      lfp = z[3];
    }
  }
  *cmax = lfp;
  *cmin = lfm;
}

/* HLL_Solver, Src/RMHD/hll.cpp:150-230 with HLL_Speed (Src/RMHD/rmhd.hpp:818-835,
   Davis): bL = MIN(Sl,0),
   bR = MAX(0,Sr), one averaged flux for every wave configuration */
static uint32_t rmhd_hll(const double *vl, const double *vr, int n, double gamma,
                         double *flux, double *pSl, double *pSr)
{
  double ul[NVMAX], ur[NVMAX], fl[NVMAX], fr[NVMAX], pl, pr;
  double a2L, a2R, hL, hR, sl_min, sl_max, sr_min, sr_max;
  rmhd_prim_to_cons(vl, ul, gamma);
  rmhd_prim_to_cons(vr, ur, gamma);
  rmhd_sound_speed2(vl, gamma, &a2L, &hL);
  rmhd_sound_speed2(vr, gamma, &a2R, &hR);
  rmhd_max_signal_speed(vl, a2L, hL, n, &sl_min, &sl_max);
  rmhd_max_signal_speed(vr, a2R, hR, n, &sr_min, &sr_max);
  double Sl = MIN(sl_min, sr_min);
  double Sr = MAX(sl_max, sr_max);
  *pSl = Sl; *pSr = Sr;
  double bL = MIN(Sl, 0.0);
  double bR = MAX(0.0, Sr);
  rmhd_flux(vl, ul, hL, n, fl, &pl);
  rmhd_flux(vr, ur, hR, n, fr, &pr);
  double scrh = 1.0/(bR - bL);
  uint32_t err = 0;
  for (int nv = 0; nv < 8; nv++) {
    flux[nv]  = bL*bR*(ur[nv] - ul[nv]) + bR*fl[nv] - bL*fr[nv];
    flux[nv] *= scrh;
  }
  for (int nv = 0; nv < 8; nv++) if (isnan(flux[nv])) err |= ERR_NAN | ERR_QUIT;
  return err;
}

/* RMHD HLLC of Mignone & Bodo (2006), Src/RMHD/hllc.cpp:150-470 (ideal EOS,
   reduced energy, no GLM; HLL_SWITCH and RMHD_ENTROPY_FIX off by default,
   Src/gpluto.hpp:877-878, Src/RMHD/rmhd.hpp:143-145).  Every failed star
   construction falls back to the HLL flux. */
#define RMHD_BX_MIN 1.e-6                   /* BX_MIN, Src/RMHD/hllc.cpp:32 */
static uint32_t rmhd_hllc(const double *vl, const double *vr, int n, double gamma,
                          double *flux, double *pSl, double *pSr)
{
  const int VXn = 1 + n, MXn = 1 + n, MXt = 1 + (n + 1)%3, MXb = 1 + (n + 2)%3;
  const int BXn = 4 + n, BXt = 4 + (n + 1)%3, BXb = 4 + (n + 2)%3;
  const int NFLX = 8, ENG = 7;
  double ul[NVMAX], ur[NVMAX], fl[NVMAX], fr[NVMAX], pl, pr;
  double a2L, a2R, hL, hR, sl_min, sl_max, sr_min, sr_max;
  double Uhll[NVMAX], Fhll[NVMAX], usl[NVMAX], usr[NVMAX];
  int nv, switch_to_hll = 0;
  rmhd_prim_to_cons(vl, ul, gamma);
  rmhd_prim_to_cons(vr, ur, gamma);
  rmhd_sound_speed2(vl, gamma, &a2L, &hL);
  rmhd_sound_speed2(vr, gamma, &a2R, &hR);
  rmhd_max_signal_speed(vl, a2L, hL, n, &sl_min, &sl_max);
  rmhd_max_signal_speed(vr, a2R, hR, n, &sr_min, &sr_max);
  double Sl = MIN(sl_min, sr_min);
  double Sr = MAX(sl_max, sr_max);
  *pSl = Sl; *pSr = Sr;
  rmhd_flux(vl, ul, hL, n, fl, &pl);
  rmhd_flux(vr, ur, hR, n, fr, &pr);

  double scrh = 1.0/(Sr - Sl);
  for (nv = 0; nv < NFLX; nv++) {
    Uhll[nv] = Sr*ur[nv] - Sl*ul[nv] + fl[nv] - fr[nv];
    Uhll[nv] *= scrh;
    Fhll[nv]  = Sl*Sr*(ur[nv] - ul[nv]) + Sr*fl[nv] - Sl*fr[nv];
    Fhll[nv] *= scrh;
  }

  if (Sl >= 0.0) {
    for (nv = 0; nv < NFLX; nv++) flux[nv] = fl[nv];
  } else if (Sr <= 0.0) {
    for (nv = 0; nv < NFLX; nv++) flux[nv] = fr[nv];
  } else {
    double vxs = 0.0;
    double vxl = vl[VXn];
    double vxr = vr[VXn];
    double Bx  = Uhll[BXn];
    double Bys = Uhll[BXt];
    double Bzs = Uhll[BXb];
    /* the derivation is done with the total energy */
    Uhll[ENG] += Uhll[RHO];
    Fhll[ENG] += Fhll[RHO];
    /* coefficients of the quadratic for v* (Bx -> 0 limit: Eq. [47]) */
    double a, b, c, BtFBt, Bt2, FBt2;
    if (fabs(Bx) < RMHD_BX_MIN) {
      BtFBt = Bt2 = FBt2 = 0.0;
      a  = Fhll[ENG];
      b  = - (Fhll[MXn] + Uhll[ENG]);
      c  = Uhll[MXn];
    } else {
      BtFBt = Uhll[BXt]*Fhll[BXt] + Uhll[BXb]*Fhll[BXb];
      Bt2   = Uhll[BXt]*Uhll[BXt] + Uhll[BXb]*Uhll[BXb];
      FBt2  = Fhll[BXt]*Fhll[BXt] + Fhll[BXb]*Fhll[BXb];
      a  = Fhll[ENG] - BtFBt;
      b  = Bt2 + FBt2 - (Fhll[MXn] + Uhll[ENG]);
      c  = Uhll[MXn] - BtFBt;
    }
    (void)Bt2;
    const double b2 = b*b;
    if (!isfinite(a) || !isfinite(b) || !isfinite(c) || !isfinite(b2) || !(b2 > 0.0)) {
      switch_to_hll = 1;
    } else {
      const double discriminant = 1.0 - 4.0*a*c/b2;
      if (!isfinite(discriminant) || discriminant < 0.0) {
        switch_to_hll = 1;
      } else {
        scrh = 1.0 + sqrt(discriminant);
        const double denominator = b*scrh;
        if (!isfinite(denominator) || denominator == 0.0) {
          switch_to_hll = 1;
        } else {
          vxs = -2.0*c/denominator;
          if (!isfinite(vxs) || !(fabs(vxs) < 1.0) || !(Sl < vxs && vxs < Sr)) switch_to_hll = 1;
        }
      }
    }
    if (!switch_to_hll) {
      if (fabs(Bx) < RMHD_BX_MIN) {
        /* degenerate case: (vy, vz) do not enter */
        double alpha_l = (Sl - vxl)/(Sl - vxs);
        double alpha_r = (Sr - vxr)/(Sr - vxs);
        double ps;
        usl[RHO] = ul[RHO]*alpha_l;
        usr[RHO] = ur[RHO]*alpha_r;
        ps  = Fhll[MXn] - Fhll[ENG]*vxs;
        usl[ENG] = (Sl*ul[ENG] - fl[ENG] + ps*vxs)/(Sl - vxs);
        usr[ENG] = (Sr*ur[ENG] - fr[ENG] + ps*vxs)/(Sr - vxs);
        usl[MXn] = (usl[ENG] + ps)*vxs;
        usr[MXn] = (usr[ENG] + ps)*vxs;
        usl[MXt] = ul[MXt]*alpha_l;
        usr[MXt] = ur[MXt]*alpha_r;
        usl[MXb] = ul[MXb]*alpha_l;
        usr[MXb] = ur[MXb]*alpha_r;
        usl[MXn] += usl[RHO]*vxs;         /* reduced energy */
        usr[MXn] += usr[RHO]*vxs;
        usl[BXn] = Bx;
        usr[BXn] = Bx;
        usl[BXt] = ul[BXt]*alpha_l;
        usr[BXt] = ur[BXt]*alpha_r;
        usl[BXb] = ul[BXb]*alpha_l;
        usr[BXb] = ur[BXb]*alpha_r;
      } else {
        double vys = (Bys*vxs - Fhll[BXt])/Bx;
        double vzs = (Bzs*vxs - Fhll[BXb])/Bx;
        double gammas_2 = vxs*vxs + vys*vys + vzs*vzs;
        gammas_2 = 1.0 - gammas_2;
        if (!isfinite(gammas_2) || gammas_2 <= 0.0) switch_to_hll = 1;   /* superluminal */
        if (!switch_to_hll) {
          double vBs = vxs*Bx + vys*Bys + vzs*Bzs;
          double alpha_l = (Sl - vxl)/(Sl - vxs);
          double alpha_r = (Sr - vxr)/(Sr - vxs);
          double ps;
          usl[RHO] = ul[RHO]*alpha_l;
          usr[RHO] = ur[RHO]*alpha_r;
          ps = (BtFBt - Fhll[ENG])*vxs + Bx*Bx + Fhll[MXn] - FBt2;
          usl[ENG] = (Sl*ul[ENG] - fl[ENG] + ps*vxs - vBs*Bx)/(Sl - vxs);
          usr[ENG] = (Sr*ur[ENG] - fr[ENG] + ps*vxs - vBs*Bx)/(Sr - vxs);
          usl[MXn] = (usl[ENG] + ps)*vxs - vBs*Bx;
          usr[MXn] = (usr[ENG] + ps)*vxs - vBs*Bx;
          usl[MXt] =  (Sl*ul[MXt] - fl[MXt] - Bx*(Bys*gammas_2 + vBs*vys))
                      /(Sl - vxs);
          usr[MXt] =  (Sr*ur[MXt] - fr[MXt] - Bx*(Bys*gammas_2 + vBs*vys))
                      /(Sr - vxs);
          usl[MXb] =  (Sl*ul[MXb] - fl[MXb] - Bx*(Bzs*gammas_2 + vBs*vzs))
                      /(Sl - vxs);
          usr[MXb] =  (Sr*ur[MXb] - fr[MXb] - Bx*(Bzs*gammas_2 + vBs*vzs))
                      /(Sr - vxs);
          usl[MXn] += usl[RHO]*vxs;       /* reduced energy */
          usr[MXn] += usr[RHO]*vxs;
          usl[BXn] = usr[BXn] = Bx;
          usl[BXt] = usr[BXt] = Bys;
          usl[BXb] = usr[BXb] = Bzs;
        }
      }
      if (!switch_to_hll) {
        if (!(usl[RHO] > 0.0) || !(usr[RHO] > 0.0)) switch_to_hll = 1;
        for (nv = 0; nv < NFLX; nv++)
          if (!isfinite(usl[nv]) || !isfinite(usr[nv])) switch_to_hll = 1;
      }
    }
    if (!switch_to_hll) {
      double press;
      if (vxs > 0.0) {
        for (nv = 0; nv < NFLX; nv++) flux[nv] = fl[nv] + Sl*(usl[nv] - ul[nv]);
        press = pl;
      } else {
        for (nv = 0; nv < NFLX; nv++) flux[nv] = fr[nv] + Sr*(usr[nv] - ur[nv]);
        press = pr;
      }
      for (nv = 0; nv < NFLX; nv++) if (!isfinite(flux[nv])) switch_to_hll = 1;
      if (!isfinite(press)) switch_to_hll = 1;
    }
    if (switch_to_hll) {
      scrh = 1.0/(Sr - Sl);
      for (nv = 0; nv < NFLX; nv++) {
        flux[nv]  = Sl*Sr*(ur[nv] - ul[nv]) + Sr*fl[nv] - Sl*fr[nv];
        flux[nv] *= scrh;
      }
    }
  }
  uint32_t err = 0;
  for (nv = 0; nv < NFLX; nv++) if (!isfinite(flux[nv])) err |= ERR_NAN | ERR_QUIT;
  return err;
}

/* RMHD HLLD of Mignone, Ugliano & Bodo (2009), Src/RMHD/hlld.cpp:95-872
   (ideal EOS, reduced energy, no GLM; HLL_SWITCH and RMHD_ENTROPY_FIX off).
   The state behind each fast wave is a function of the total pressure p
   (HLLD_GetRiemannState); a secant iteration on p closes the jump of the
   normal velocity across the contact (HLLD_Fstar); any failure gives the HLL
   flux. */
typedef struct {
  double S, R[NVMAX], u[NVMAX];
  double vx, vy, vz, Bx, By, Bz, w, sw, Kx, Ky, Kz, K2, Sa;
  int fail;
} rhlld_state;                                  /* Riemann_State, Src/structs.hpp */

static uint32_t rmhd_cons_to_prim(double *u, double *v, const gplo_config *c);

/* HLLD_GetRiemannState, Src/RMHD/hlld.cpp:639-731; 0 when w < 0 (the other
   members then keep their previous values, as in the reference) */
static int rhlld_riemann_state(rhlld_state *Pv, double p, double Bx, int side, int n)
{
  const int MXn = 1 + n, MXt = 1 + (n + 1)%3, MXb = 1 + (n + 2)%3;
  const int BXn = 4 + n, BXt = 4 + (n + 1)%3, BXb = 4 + (n + 2)%3, ENG = 7;
  double A, C, G, X, s, vx, vy, vz, scrh, S = Pv->S, *R = Pv->R;
  A = R[MXn] + p*(1.0 - S*S) - S*R[ENG];          /* Eq. [26] */
  G = R[BXt]*R[BXt] + R[BXb]*R[BXb];              /* Eq. [27] */
  C = R[BXt]*R[MXt] + R[BXb]*R[MXb];              /* Eq. [28] */
  X = Bx*(A*S*Bx + C) - (A + G)*(S*p + R[ENG]);   /* Eq. [30] */
  vx = ( Bx*(A*Bx + C*S) - (R[MXn] + p)*(G + A) );
  vy = ( - (A + G - Bx*Bx*(1.0 - S*S))*R[MXt]
         + R[BXt]*(C + Bx*(S*R[MXn] - R[ENG])) );
  vz = ( - (A + G - Bx*Bx*(1.0 - S*S))*R[MXb]
         + R[BXb]*(C + Bx*(S*R[MXn] - R[ENG])) );
  scrh = vx*R[MXn] + vy*R[MXt] + vz*R[MXb];
  scrh = X*R[ENG] - scrh;
  Pv->w = p + scrh/(X*S - vx);                    /* Eq. [31] */
  if (Pv->w < 0.0) return 0;
  Pv->vx = vx/X;
  Pv->vy = vy/X;
  Pv->vz = vz/X;
  Pv->Bx = Bx;
  Pv->By = -(R[BXt]*(S*p + R[ENG]) - Bx*R[MXt])/A;
  Pv->Bz = -(R[BXb]*(S*p + R[ENG]) - Bx*R[MXb])/A;
  s  = Bx > 0.0 ? 1.0 : -1.0;
  if (side < 0) s *= -1.0;
  Pv->sw = s*sqrt(Pv->w);
  scrh = 1.0/(S*p +  R[ENG] + Bx*Pv->sw);
  Pv->Kx = scrh*(R[MXn] + p + R[BXn]*Pv->sw);
  Pv->Ky = scrh*(R[MXt]     + R[BXt]*Pv->sw);
  Pv->Kz = scrh*(R[MXb]     + R[BXb]*Pv->sw);
  Pv->K2 = Pv->Kx*Pv->Kx + Pv->Ky*Pv->Ky + Pv->Kz*Pv->Kz;
  return 1;
}

/* HLLD_Fstar, Src/RMHD/hlld.cpp:560-636: the normal velocity jump across
   the contact at total pressure p (Eq. [48]) */
static double rhlld_fstar(rhlld_state *PaL, rhlld_state *PaR, double p, double Bx, double *Sc, int n)
{
  int success = 1;
  double dK, Bxc, Byc, Bzc, fun, vxcL, KLBc, vxcR, KRBc;
  success *= rhlld_riemann_state(PaL, p, Bx, -1, n);
  success *= rhlld_riemann_state(PaR, p, Bx,  1, n);
  dK  = PaR->Kx - PaL->Kx + 1.e-12;
  Bxc = Bx*dK;
  Byc =   PaR->By*(PaR->Kx - PaR->vx)
        - PaL->By*(PaL->Kx - PaL->vx)
        + Bx*(PaR->vy - PaL->vy);
  Bzc =   PaR->Bz*(PaR->Kx - PaR->vx)
        - PaL->Bz*(PaL->Kx - PaL->vx)
        + Bx*(PaR->vz - PaL->vz);
  KLBc = PaL->Kx*Bxc + PaL->Ky*Byc + PaL->Kz*Bzc;
  KRBc = PaR->Kx*Bxc + PaR->Ky*Byc + PaR->Kz*Bzc;
  vxcL = PaL->Kx - dK*Bx*(1.0 - PaL->K2)/(PaL->sw*dK - KLBc);
  vxcR = PaR->Kx - dK*Bx*(1.0 - PaR->K2)/(PaR->sw*dK - KRBc);
  PaL->Sa = PaL->Kx;
  PaR->Sa = PaR->Kx;
  *Sc     = 0.5*(vxcL + vxcR);
  fun     = vxcL - vxcR;
  /* the reference overwrites the two state flags here */
  success  = (vxcL - PaL->Kx) > -1.e-6;
  success *= (PaR->Kx  - vxcR) > -1.e-6;
  success *= (PaL->S - PaL->vx) < 0.0;
  success *= (PaR->S - PaR->vx) > 0.0;
  success *= (PaR->w - p) > 0.0;
  success *= (PaL->w - p) > 0.0;
  success *= (PaL->Sa - PaL->S)  > -1.e-6;
  success *= (PaR->S  - PaR->Sa) > -1.e-6;
  PaL->fail = !success;
  return fun;
}

/* HLLD_GetAState, Src/RMHD/hlld.cpp:734-768: the state behind a fast wave */
static void rhlld_astate(rhlld_state *Pa, double p, double Bx, int n)
{
  const int MXn = 1 + n, MXt = 1 + (n + 1)%3, MXb = 1 + (n + 2)%3;
  const int BXn = 4 + n, BXt = 4 + (n + 1)%3, BXb = 4 + (n + 2)%3, ENG = 7;
  double vB, *ua = Pa->u, S = Pa->S, *R = Pa->R, scrh;
  scrh = 1.0/(S - Pa->vx);
  ua[RHO] =  R[RHO]*scrh;
  ua[BXn] =  Bx;
  ua[BXt] = (R[BXt] - Bx*Pa->vy)*scrh;
  ua[BXb] = (R[BXb] - Bx*Pa->vz)*scrh;
  vB     = Pa->vx*ua[BXn] + Pa->vy*ua[BXt] + Pa->vz*ua[BXb];
  ua[ENG] = (R[ENG] + p*Pa->vx - vB*Bx)*scrh;
  ua[MXn] = (ua[ENG] + p)*Pa->vx - vB*ua[BXn];
  ua[MXt] = (ua[ENG] + p)*Pa->vy - vB*ua[BXt];
  ua[MXb] = (ua[ENG] + p)*Pa->vz - vB*ua[BXb];
}

/* HLLD_GetCState, Src/RMHD/hlld.cpp:771-855: the state across the contact */
static void rhlld_cstate(rhlld_state *PaL, rhlld_state *PaR, double p, double Bx, double *Uc, int n)
{
  const int MXn = 1 + n, MXt = 1 + (n + 1)%3, MXb = 1 + (n + 2)%3;
  const int BXn = 4 + n, BXt = 4 + (n + 1)%3, BXb = 4 + (n + 2)%3, ENG = 7;
  double dK, *ua, vxcL, vycL, vzcL, KLBc, vxcR, vycR, vzcR, KRBc;
  double vxc, vyc, vzc, vBc, Bxc, Byc, Bzc, Sa, vxa, scrhL, scrhR;
  rhlld_astate(PaL, p, Bx, n);
  rhlld_astate(PaR, p, Bx, n);
  dK = (PaR->Kx - PaL->Kx) + 1.e-12;
  Bxc = Bx*dK;
  Byc =   PaR->By*(PaR->Kx - PaR->vx)
        - PaL->By*(PaL->Kx - PaL->vx)
        + Bx*(PaR->vy - PaL->vy);
  Bzc =   PaR->Bz*(PaR->Kx - PaR->vx)
        - PaL->Bz*(PaL->Kx - PaL->vx)
        + Bx*(PaR->vz - PaL->vz);
  Bxc  = Bx;
  Byc /= dK;
  Bzc /= dK;
  Uc[BXn] = Bxc;
  Uc[BXt] = Byc;
  Uc[BXb] = Bzc;
  KLBc = PaL->Kx*Bxc + PaL->Ky*Byc + PaL->Kz*Bzc;
  KRBc = PaR->Kx*Bxc + PaR->Ky*Byc + PaR->Kz*Bzc;
  scrhL = (1.0 - PaL->K2)/(PaL->sw - KLBc);
  scrhR = (1.0 - PaR->K2)/(PaR->sw - KRBc);
  vxcL = PaL->Kx - Uc[BXn]*scrhL;                 /* Eq. [47] */
  vxcR = PaR->Kx - Uc[BXn]*scrhR;
  vycL = PaL->Ky - Uc[BXt]*scrhL;
  vycR = PaR->Ky - Uc[BXt]*scrhR;
  vzcL = PaL->Kz - Uc[BXb]*scrhL;
  vzcR = PaR->Kz - Uc[BXb]*scrhR;
  vxc = 0.5*(vxcL + vxcR);
  vyc = 0.5*(vycL + vycR);
  vzc = 0.5*(vzcL + vzcR);
  if (vxc > 0.0) {
    rhlld_astate(PaL, p, Bx, n);
    ua = PaL->u; Sa = PaL->Sa; vxa = PaL->vx;
  } else {
    rhlld_astate(PaR, p, Bx, n);
    ua = PaR->u; Sa = PaR->Sa; vxa = PaR->vx;
  }
  vBc = vxc*Uc[BXn] + vyc*Uc[BXt] + vzc*Uc[BXb];
  Uc[RHO] = ua[RHO]*(Sa - vxa)/(Sa - vxc);                       /* Eq. [32] */
  Uc[ENG] = (Sa*ua[ENG] - ua[MXn] + p*vxc - vBc*Bx)/(Sa - vxc);  /* Eq. [33] */
  Uc[MXn] = (Uc[ENG] + p)*vxc - vBc*Bx;                          /* Eq. [34] */
  Uc[MXt] = (Uc[ENG] + p)*vyc - vBc*Uc[BXt];
  Uc[MXb] = (Uc[ENG] + p)*vzc - vBc*Uc[BXb];
}

/* HLLD_TotalPressure, Src/RMHD/hlld.cpp:857-872 */
static double rhlld_total_pressure(const double *v)
{
  double Bmag2 = v[4]*v[4] + v[5]*v[5] + v[6]*v[6];
  double vel2  = v[1]*v[1] + v[2]*v[2] + v[3]*v[3];
  double vB    = v[1]*v[4] + v[2]*v[5] + v[3]*v[6];
  return v[7] + 0.5*(Bmag2*(1.0 - vel2) + vB*vB);
}

static uint32_t rmhd_hlld(const double *vl, const double *vr, int n, const gplo_config *cfg,
                          double *flux, double *pSl, double *pSr)
{
  const int MXn = 1 + n, BXn = 4 + n, NFLX = 8, ENG = 7;
  const double gamma = cfg->gamma;
  double ul[NVMAX], ur[NVMAX], fl[NVMAX], fr[NVMAX], pl, pr;
  double a2L, a2R, hL, hR, sl_min, sl_max, sr_min, sr_max;
  int nv;
  uint32_t err = 0;
  rmhd_prim_to_cons(vl, ul, gamma);
  rmhd_prim_to_cons(vr, ur, gamma);
  rmhd_sound_speed2(vl, gamma, &a2L, &hL);
  rmhd_sound_speed2(vr, gamma, &a2R, &hR);
  rmhd_flux(vl, ul, hL, n, fl, &pl);
  rmhd_flux(vr, ur, hR, n, fr, &pr);
  rmhd_max_signal_speed(vl, a2L, hL, n, &sl_min, &sl_max);
  rmhd_max_signal_speed(vr, a2R, hR, n, &sr_min, &sr_max);
  double Sl = MIN(sl_min, sr_min);
  double Sr = MAX(sl_max, sr_max);
  *pSl = Sl; *pSr = Sr;

  if (Sl >= 0.0) {
    for (nv = 0; nv < NFLX; nv++) flux[nv] = fl[nv];
  } else if (Sr <= 0.0) {
    for (nv = 0; nv < NFLX; nv++) flux[nv] = fr[nv];
  } else {
    int switch_to_hll = 0;
    double Uhll[NVMAX], Fhll[NVMAX], Vhll[NVMAX], Uc[NVMAX];
    double p0, f0, p, f, dp, Sc = 0.0, scrh;
    double dS_1 = 1.0/(Sr - Sl);
    for (nv = 0; nv < NFLX; nv++) {
      Uhll[nv]  = Sr*ur[nv] - Sl*ul[nv] + fl[nv] - fr[nv];
      Uhll[nv] *= dS_1;
      Fhll[nv]  = Sr*fl[nv] - Sl*fr[nv] + Sl*Sr*(ur[nv] - ul[nv]);
      Fhll[nv] *= dS_1;
    }
    rhlld_state PaL, PaR;
    memset(&PaL, 0, sizeof(PaL));
    memset(&PaR, 0, sizeof(PaR));
    PaL.S = Sl;
    PaR.S = Sr;
    double Bx = Uhll[BXn];
    for (nv = 0; nv < NFLX; nv++) {
      PaL.R[nv] = Sl*ul[nv] - fl[nv];
      PaR.R[nv] = Sr*ur[nv] - fr[nv];
    }
    PaL.R[ENG] += PaL.R[RHO];                     /* reduced energy */
    PaR.R[ENG] += PaR.R[RHO];
    /* initial guess: the B -> 0 limit, or the total pressure of the HLL state */
    scrh = MAX(pl, pr);
    if (Bx*Bx/scrh < 0.01) {
      double a = Sr - Sl;
      double b = PaR.R[ENG] - PaL.R[ENG] + Sr*PaL.R[MXn] - Sl*PaR.R[MXn];
      double c = PaL.R[MXn]*PaR.R[ENG] - PaR.R[MXn]*PaL.R[ENG];
      scrh = b*b - 4.0*a*c;
      scrh = MAX(scrh, 0.0);
      p0 = 0.5*(- b + sqrt(scrh))*dS_1;
    } else {
      uint32_t local_err = rmhd_cons_to_prim(Uhll, Vhll, cfg);
      err |= local_err;
      if (local_err) switch_to_hll = 1;
      p0 = rhlld_total_pressure(Vhll);
    }
    f0 = rhlld_fstar(&PaL, &PaR, p0, Bx, &Sc, n);
    if (isnan(f0) || PaL.fail) switch_to_hll = 1;
    if (fabs(f0) > 1.e-12 && !switch_to_hll) {
      p = 1.025*p0; f = f0;
      for (int k1 = 1; k1 < 20; k1++) {           /* MAX_ITER, Src/RMHD/hlld.cpp:45 */
        f = rhlld_fstar(&PaL, &PaR, p, Bx, &Sc, n);
        if (f != f || PaL.fail || (k1 > 7) || (fabs(f) > fabs(f0) && k1 > 4)) {
          switch_to_hll = 1;
          break;
        }
        dp = (p - p0)/(f - f0)*f;
        p0 = p; f0 = f;
        p -= dp;
        if (p < 0.0) p = 1.e-6;
        if (fabs(dp) < 1.e-5*p || fabs(f) < 1.e-6) break;
      }
    } else p = p0;
    if (PaL.fail) switch_to_hll = 1;
    if (!switch_to_hll) {
      if (PaL.Sa >= -1.e-6) {
        rhlld_astate(&PaL, p, Bx, n);
        PaL.u[ENG] -= PaL.u[RHO];
        for (nv = 0; nv < NFLX; nv++) flux[nv] = fl[nv] + Sl*(PaL.u[nv] - ul[nv]);
      } else if (PaR.Sa <= 1.e-6) {
        rhlld_astate(&PaR, p, Bx, n);
        PaR.u[ENG] -= PaR.u[RHO];
        for (nv = 0; nv < NFLX; nv++) flux[nv] = fr[nv] + Sr*(PaR.u[nv] - ur[nv]);
      } else {
        rhlld_cstate(&PaL, &PaR, p, Bx, Uc, n);
        if (Sc > 0.0) {
          PaL.u[ENG] -= PaL.u[RHO];
          Uc[ENG]    -= Uc[RHO];
          for (nv = 0; nv < NFLX; nv++)
            flux[nv] = fl[nv] + Sl*(PaL.u[nv] - ul[nv]) + PaL.Sa*(Uc[nv] - PaL.u[nv]);
        } else {
          PaR.u[ENG] -= PaR.u[RHO];
          Uc[ENG]    -= Uc[RHO];
          for (nv = 0; nv < NFLX; nv++)
            flux[nv] = fr[nv] + Sr*(PaR.u[nv] - ur[nv]) + PaR.Sa*(Uc[nv] - PaR.u[nv]);
        }
      }
    } else {
      for (nv = 0; nv < NFLX; nv++) flux[nv] = Fhll[nv];
    }
  }
  for (nv = 0; nv < NFLX; nv++) if (isnan(flux[nv])) err |= ERR_NAN | ERR_QUIT;
  return err;
}

/* RMHD_EnergySolve, Src/RMHD/rmhd_energy_solve.cpp:35-209 (IDEAL, reduced
   energy, Newton-Raphson on W' = W - D, tolerance 1e-11, at most 50 passes) */
static uint32_t rmhd_energy_solve(double *u, double *v, double gamma)
{
  const double acc = 1.e-11;
  int k, done;
  double Y1, Y2, scrh, W, W2, S2_W2, fW, dfW, dW, dv2_dW, chi, p, lor = 1.0, lor2;
  double dp, dp_drho, dp_dchi, dchi_dW, drho_dW, one_m_v2, vel2, Wp;
  double D  = u[RHO];
  double E  = u[7];
  double S  = u[1]*u[4] + u[2]*u[5] + u[3]*u[6];
  double m2 = u[1]*u[1] + u[2]*u[2] + u[3]*u[3];
  double B2 = u[4]*u[4] + u[5]*u[5] + u[6]*u[6];
  double S2 = S*S;
  if (E < 0.0) return ERR_NEGATIVE_ENG;
  /* initial guess: positive root of Eq. (A27) */
  Y1 = -4.0*(E + D - B2);
  Y2 = m2 - 2.0*(E + D)*B2 + B2*B2;
  chi = Y1*Y1 - 12.0*Y2;
  chi = MAX(0.0,chi);
  W = ( - Y1 + sqrt(chi))/6.0;
  W = MAX(D, W);
  Wp = W - D;
  done = 0; p = -1.0;
  for (k = 1; k < 50; k++) {
    W = Wp + D;
    W2    = W*W;
    S2_W2 = S2/W2;
    Y1    = 1.0/(W + B2);
    Y2    = Y1*Y1;
    vel2 = S2_W2*Y1*(Y1*W + 1.0) + m2*Y2;
    one_m_v2 = 1.0 - vel2;
    lor2 = 1.0/one_m_v2;
    lor  = sqrt(lor2);
    if (vel2 > 1.0) return ERR_ENG_SOLVE_NO_SOLUTION;
    chi = Wp/lor2 - D*vel2/(lor + 1.0);
    dv2_dW  = -2.0*Y2*(3.0*S2_W2 + Y1*(S2_W2*B2*B2/W + m2));
    dchi_dW =  one_m_v2 - 0.5*lor*(D + 2.0*chi*lor)*dv2_dW;
    drho_dW = -0.5*D*lor*dv2_dW;
    dp_dchi = (gamma - 1.0)/gamma;
    dp_drho = 0.0;
    p       = chi*dp_dchi;
    if (done) break;
    dp  = dp_dchi*dchi_dW + dp_drho*drho_dW;
    fW  = Wp + 0.5*(B2 + (B2*m2 - S2)*Y2) - (E + p);
    dfW = 1.0 - dp - (B2*m2 - S2)*Y2*Y1;
    dW  = fW/dfW;
    Wp -= dW;
    if (fabs(dW) < acc*Wp || fabs(fW) < acc) done = 1;
  }
  if (p < 0.0)  return ERR_NEGATIVE_PRS;
  if (k == 50)  return ERR_TOO_MANY_ITERATIONS;
  if (isnan(p)) return ERR_NAN;
  W = Wp + D;
  v[RHO] = u[RHO]/lor;
  v[7] = p;
  Y1   = 1.0/(W + B2);
  scrh = S/W;
  v[1] = Y1*(u[1] + scrh*u[4]);
  v[2] = Y1*(u[2] + scrh*u[5]);
  v[3] = Y1*(u[3] + scrh*u[6]);
  return 0;
}

/* RMHD_PressureFix, Src/RMHD/rmhd_pressure_fix.cpp:22-179: p = small_prs,
   bisection on v^2 of Eq. (A3), then E recomputed */
static uint32_t rmhd_pressure_fix(double *u, double *v, double gamma, double small_prs)
{
  int k;
  double v2, f, W, W2, S2_W2, fmin, v2min, v2max, lor, plor, scrh, w_1;
  double D  = u[RHO];
  double S  = u[1]*u[4] + u[2]*u[5] + u[3]*u[6];
  double m2 = u[1]*u[1] + u[2]*u[2] + u[3]*u[3];
  double B2 = u[4]*u[4] + u[5]*u[5] + u[6]*u[6];
  double S2 = S*S;
  double p = small_prs;
  v2min = v2 = 0.0;
  lor  = 1.0/sqrt(1.0 - v2);
  plor = p*lor;
  W = (D + plor*gamma/(gamma - 1.0))*lor;
  W2  = W*W;
  f  = (S2*(2.0*W + B2) + m2*W2)/((W + B2)*(W + B2)*W2);
  f -= v2;
  fmin = f;
  v2max = v2 = 1.0-1.e-9;
  lor  = 1.0/sqrt(1.0 - v2);
  plor = p*lor;
  W = (D + plor*gamma/(gamma - 1.0))*lor;
  W2  = W*W;
  f  = (S2*(2.0*W + B2) + m2*W2)/((W + B2)*(W + B2)*W2);
  f -= v2;
  for (k = 1; k < 50; k++) {
    v2 = 0.5*(v2min + v2max);
    lor  = 1.0/sqrt(1.0 - v2);
    plor = p*lor;
    W = (D + plor*gamma/(gamma - 1.0))*lor;
    W2  = W*W;
    f  = (S2*(2.0*W + B2) + m2*W2)/((W + B2)*(W + B2)*W2);
    f -= v2;
    if (f*fmin > 0.0) { v2min = v2; fmin = f; }
    else              { v2max = v2; }
    if (fabs(f) < 1.e-9) break;
  }
  if (v2 >= 1.0 && k == 50) return ERR_PRESSURE_FIX;
  v[RHO] = D/lor;
  v[7] = p;
  w_1  = 1.0/(W + B2);
  scrh = S/W;
  v[1] = w_1*(u[1] + scrh*u[4]);
  v[2] = w_1*(u[2] + scrh*u[5]);
  v[3] = w_1*(u[3] + scrh*u[6]);
  S2_W2 = S2/W2;
  u[7] = W - D - p + 0.5*(1.0 + v2)*B2 - 0.5*S2_W2;
  return 0;
}

/* ConsToPrim, Src/RMHD/mappers.cpp:33-157 (energy solve, pressure fix on
   failure, subluminal check) */
static uint32_t rmhd_cons_to_prim(double *u, double *v, const gplo_config *c)
{
  int fix = 0;
  uint32_t errm = 0, err;
  v[4] = u[4];
  v[5] = u[5];
  v[6] = u[6];
  if (u[RHO] < 0.0) { u[RHO] = c->small_rho; errm |= ERR_NEGATIVE_RHO; }
  if (u[7] < 0.0)   { fix = 1; errm |= ERR_NEGATIVE_ENG; }
  if (!fix) {
    err = rmhd_energy_solve(u, v, c->gamma);
    if (err) fix = 1;
    errm |= err;
  }
  if (fix) {
    err = rmhd_pressure_fix(u, v, c->gamma, c->small_prs);
    if (err != 0) { errm |= err; errm |= ERR_CONS2PRIM_FAIL; }
  }
  double scrh = v[1]*v[1] + v[2]*v[2] + v[3]*v[3];
  if (scrh >= 1.0) { errm |= ERR_SUPERLUMINAL_SPEED; errm |= ERR_CONS2PRIM_FAIL; }
  return errm;
}

/* ======================================================================
 * PLM slope, Src/States/plm_states.cpp:121-138 with SET_MC_LIMITER /
 * SET_VL_LIMITER on a uniform Cartesian grid (Src/plm_coeffs.hpp:92-100)
 * ====================================================================== */
static double plm_slope(double vm, double v0, double vp, int lim)
{
  double dvp = (vp - v0)*1.0;   /* geometry.wp = 1 (reconstruction_geometry.hpp:11) */
  double dvm = (v0 - vm)*1.0;
  double dv;
  if (lim == GPLO_MC_LIM) {
    if ( (dvp)*(dvm) > 0.0) {
      double _qc  = 0.5*( (dvm) + (dvp) ), _scrh = 2.0*ABS_MIN( (dvp), (dvm) );
      dv   = ABS_MIN(_qc, _scrh);
    } else dv = 0.0;
  } else {
    dv = ( (dvp)*(dvm) > 0.0 ? 2.0*(dvp)*(dvm)/((dvp) + (dvm)) :0.0);
  }
  return dv;
}

/* CT_Store_UCTHlld<dir>, Src/MHD/CT/ct_store.hpp:254-354 (ideal EOS): the
   UCT weights and diffusion coefficients from the HLLD fan, with the face's
   own normal field Bn (= vl[BXn] = vr[BXn] after the PLM override), the HLL
   values where a star density is negative */
static void uct_hlld_store(gplo_state *s, const double *vl, const double *vr, int dir,
                           int t1, int t2, double Sl, double Sr, size_t id)
{
  const int VXn = 1 + dir, VXt = 1 + t1, VXb = 1 + t2, MXn = VXn;
  double ul[NVMAX], ur[NVMAX], scrh;
  int switch_to_hll = 0;
  mhd_prim_to_cons(vl, ul, s->c.gamma);
  mhd_prim_to_cons(vr, ur, s->c.gamma);
  double aR_1 = MAX(0.0, Sr);
  double aL_1 = MIN(0.0, Sl);
  double Bn = vl[4 + dir];
  double dul = Sl - vl[VXn];
  double dur = Sr - vr[VXn];
  double pL = vl[7] + 0.5*(vl[4]*vl[4] + vl[5]*vl[5] + vl[6]*vl[6]);
  double pR = vr[7] + 0.5*(vr[4]*vr[4] + vr[5]*vr[5] + vr[6]*vr[6]);
  scrh  = 1.0/(dur*ur[RHO] - dul*ul[RHO]);
  double Sc = (dur*ur[MXn] - dul*ul[MXn] - pR + pL)*scrh;
  double usl_RHO = ul[RHO]*dul/(Sl - Sc);
  double usr_RHO = ur[RHO]*dur/(Sr - Sc);
  if (usl_RHO < 0.0 || usr_RHO < 0.0) switch_to_hll = 1;
  double sqrL = sqrt(usl_RHO);
  double sqrR = sqrt(usr_RHO);
  double Sal = Sc - fabs(Bn)/sqrL;
  double Sar = Sc + fabs(Bn)/sqrR;
  double chiL  = (vl[VXn] - Sc)*(Sl - Sc)/(Sal + Sl - 2.0*Sc);
  double chiR  = (vr[VXn] - Sc)*(Sr - Sc)/(Sar + Sr - 2.0*Sc);
  double nuL  = (Sal + Sl)/(fabs(Sal) + fabs(Sl));
  double nuR  = (Sar + Sr)/(fabs(Sar) + fabs(Sr));
  double nuLR = (Sal + Sar)/(fabs(Sal) + fabs(Sar));
  if (fabs(Sar - Sal) > 1.e-9*fabs(Sr - Sl)) {
    s->dL[dir][id] =   0.5*(chiL*nuL - chiL*nuLR)
                     + 0.5*(fabs(Sal) - nuLR*Sal);
    s->dR[dir][id] =   0.5*(chiR*nuR - chiR*nuLR)
                     + 0.5*(fabs(Sar) - nuLR*Sar);
    s->aL[dir][id] = 0.5*(1.0 + nuLR);
    s->aR[dir][id] = 0.5*(1.0 - nuLR);
  } else {   /* HLLC, degenerate limit Bn -> 0 */
    s->dL[dir][id] = 0.5*chiL*nuL + 0.5*fabs(Sal);
    s->dR[dir][id] = 0.5*chiR*nuR + 0.5*fabs(Sar);
    s->aL[dir][id] = 0.5;
    s->aR[dir][id] = 0.5;
  }
  scrh = 1.0/(aR_1 - aL_1);
  if (switch_to_hll) {
    s->aL[dir][id] =  aR_1*scrh;
    s->aR[dir][id] = -aL_1*scrh;
    s->dR[dir][id] = -aL_1*aR_1*scrh;
    s->dL[dir][id] = -aL_1*aR_1*scrh;
  }
  s->e[t2][dir][id] =  (aR_1*vl[VXt] - aL_1*vr[VXt])*scrh;
  s->e[t1][dir][id] = -(aR_1*vl[VXb] - aL_1*vr[VXb])*scrh;
}

/* ======================================================================
 * Boundary conditions
 * ====================================================================== */

/* Cell-centred B of ghost zones from faces: Src/boundary.cpp:275-303 with
   CT_AverageStaggeredFields, Src/MHD/CT/ct_field_average.cpp:182-184 */
static void ghost_field_average(gplo_state *s, const int lo[3], const int hi[3])
{
  int i, j, k, d;
  for (k = lo[2]; k <= hi[2]; k++)
  for (j = lo[1]; j <= hi[1]; j++)
  for (i = lo[0]; i <= hi[0]; i++) {
    for (d = 0; d < s->c.dims; d++) {
      int o[3] = {0, 0, 0}; o[d] = 1;
      double b = 0.5*(s->Vs[d][IS(s,i,j,k)] + s->Vs[d][IS(s,i+o[0],j+o[1],k+o[2])]);
      UC(s, BX1 + d, i, j, k) = b;
      VC(s, BX1 + d, i, j, k) = b;
    }
  }
}

/* CT_FillStaggeredFields, Src/MHD/CT/ct_fill_mag_field.cpp:39-330: the
   normal staggered component in the ghost layers of one side from
   div B = 0, outward from the boundary face, over every transverse index
   (ghosts included).  Face areas of the uniform Cartesian grid
   (Src/set_geometry.cpp:124,147,166); the missing direction of a 2D box adds
   its 0.0. */
static void fill_staggered(gplo_state *s, int dir, int upper)
{
  const int dims = s->c.dims, t1 = (dir + 1)%3, t2 = (dir + 2)%3;
  const double *dx = s->dx;
  double A[3];
  A[0] = dims == 3 ? 1.0*dx[1]*dx[2] : 1.0*dx[1];
  A[1] = dims == 3 ? dx[0]*1.0*dx[2] : dx[0]*1.0;
  A[2] = dx[0]*dx[1]*1.0;
  int a, b, i;
  for (b = 0; b < s->nt[t2]; b++)
  for (a = 0; a < s->nt[t1]; a++) {
    const int n = s->nt[dir];
    const int i0 = upper ? s->end[dir] + 1 : s->beg[dir] - 1;
    const int i1 = upper ? n - 1 : 0, st = upper ? 1 : -1;
    for (i = i0; upper ? i <= i1 : i >= i1; i += st) {
      int p[3];
      p[dir] = i; p[t1] = a; p[t2] = b;
      double Bp[3], Bm[3], dB[3] = {0.0, 0.0, 0.0};
      for (int d = 0; d < dims; d++) {
        int q[3] = {p[0], p[1], p[2]};
        q[d] += 1;
        Bp[d] = s->Vs[d][IS(s, q[0], q[1], q[2])];
        Bm[d] = s->Vs[d][IS(s, p[0], p[1], p[2])];
        dB[d] = A[d]*Bp[d] - A[d]*Bm[d];
      }
      /* the two transverse terms in the reference's order (x, y, z skipping dir) */
      double ta = 0.0, tb = 0.0;
      if (dir == 0)      { ta = dB[1]; tb = dB[2]; }
      else if (dir == 1) { ta = dB[0]; tb = dB[2]; }
      else               { ta = dB[0]; tb = dB[1]; }
      if (!upper) {
        s->Vs[dir][IS(s, p[0], p[1], p[2])] = (A[dir]*Bp[dir] + ta + tb)/A[dir];
      } else {
        int q[3] = {p[0], p[1], p[2]};
        q[dir] += 1;
        s->Vs[dir][IS(s, q[0], q[1], q[2])] = (A[dir]*Bm[dir] - (ta + tb))/A[dir];
      }
    }
  }
}

/* Boundary(), Src/boundary.cpp:67-303: sides X1_BEG..X3_END in order;
   PeriodicBoundary :502-669; OutflowBoundary :677-760 with
   CT_FillStaggeredFields for MHD/RMHD; ReflectiveBoundary :931-1100 with
   FlipSign :840-925; then
   the ghost cells' field average (:275-303). */
static void boundary(gplo_state *s)
{
  int side, i, j, k, nv, d;
  for (side = 0; side < 2*s->c.dims; side++) {
    int dir = side/2, upper = side%2;
    int lo[3] = {0, 0, 0}, hi[3];
    for (d = 0; d < 3; d++) hi[d] = s->nt[d] - 1;
    if (!upper) { lo[dir] = 0; hi[dir] = s->beg[dir] - 1; }
    else        { lo[dir] = s->end[dir] + 1; hi[dir] = s->nt[dir] - 1; }
    int nint = s->c.dims > dir ? (dir == 0 ? s->c.nx : dir == 1 ? s->c.ny : s->c.nz) : 1;
    const int type = s->c.bc[side];
    if (type == GPLO_BC_REFLECTIVE) {
      /* ReflectiveBoundary, Src/boundary.cpp:931-1100, with FlipSign
         (:840-925): ghost cell i = vsign*Vc(i0 - i), i0 = 2*beg - 1 on the
         lower side and 2*end + 1 on the upper side; vsign flips the normal
         velocity and the normal field */
      const int i0 = upper ? 2*s->end[dir] + 1 : 2*s->beg[dir] - 1;
      int vsign[NVMAX];
      for (nv = 0; nv < s->nv; nv++) vsign[nv] = 1;
      vsign[1 + dir] = -1;
      if (is_mhd(s)) vsign[BX1 + dir] = -1;
      for (nv = 0; nv < s->nv; nv++)
        for (k = lo[2]; k <= hi[2]; k++)
        for (j = lo[1]; j <= hi[1]; j++)
        for (i = lo[0]; i <= hi[0]; i++) {
          int p[3] = {i, j, k};
          p[dir] = i0 - p[dir];
          VC(s,nv,i,j,k) = vsign[nv]*VC(s,nv,p[0],p[1],p[2]);
        }
      if (is_mhd(s)) {
        /* staggered part (Src/boundary.cpp:955-985 and analogues): the
           tangential components mirrored over BOX{J,K}_LOOP, the normal one
           mirrored about the boundary face with its sign flipped */
        for (d = 0; d < s->c.dims; d++) {
          int ext[3] = {0, 0, 0};
          if (d != dir) ext[d] = 1;
          for (k = lo[2]; k <= hi[2] + ext[2]; k++)
          for (j = lo[1]; j <= hi[1] + ext[1]; j++)
          for (i = lo[0]; i <= hi[0] + ext[0]; i++) {
            int p[3] = {i, j, k}, q[3] = {i, j, k};
            if (d != dir) {
              p[dir] = i0 - p[dir];
            } else if (!upper) {
              p[dir] = i0 - p[dir] + 1;
            } else {
              p[dir] = i0 - p[dir];
              q[dir] += 1;
            }
            s->Vs[d][IS(s,q[0],q[1],q[2])] = vsign[BX1 + d]*s->Vs[d][IS(s,p[0],p[1],p[2])];
          }
        }
      }
    } else if (type == GPLO_BC_PERIODIC) {
      int sh = upper ? -nint : nint;
      int o[3] = {0, 0, 0}; o[dir] = sh;
      for (nv = 0; nv < s->nv; nv++)
        for (k = lo[2]; k <= hi[2]; k++)
        for (j = lo[1]; j <= hi[1]; j++)
        for (i = lo[0]; i <= hi[0]; i++)
          VC(s,nv,i,j,k) = VC(s,nv,i+o[0],j+o[1],k+o[2]);
      if (is_mhd(s)) {
        for (d = 0; d < s->c.dims; d++) {
          /* BOX_LOOP on the lower side for the normal component, BOX{I,J,K}_LOOP
             (one more face along d) otherwise */
          int ext[3] = {0, 0, 0};
          if (!(d == dir && !upper)) ext[d] = 1;
          for (k = lo[2]; k <= hi[2] + ext[2]; k++)
          for (j = lo[1]; j <= hi[1] + ext[1]; j++)
          for (i = lo[0]; i <= hi[0] + ext[0]; i++)
            s->Vs[d][IS(s,i,j,k)] = s->Vs[d][IS(s,i+o[0],j+o[1],k+o[2])];
        }
      }
    } else {
      /* OutflowBoundary, Src/boundary.cpp:677-760: every variable and the
         tangential staggered components (BOX{J,K}_LOOP: one more face along
         the component) copied from the edge layer; then the normal
         component from div B = 0 (Src/boundary.cpp:193-199) */
      for (nv = 0; nv < s->nv; nv++)
        for (k = lo[2]; k <= hi[2]; k++)
        for (j = lo[1]; j <= hi[1]; j++)
        for (i = lo[0]; i <= hi[0]; i++) {
          int p[3] = {i, j, k};
          p[dir] = upper ? s->end[dir] : s->beg[dir];
          VC(s,nv,i,j,k) = VC(s,nv,p[0],p[1],p[2]);
        }
      if (is_mhd(s)) {
        for (d = 0; d < s->c.dims; d++) {
          if (d == dir) continue;
          int ext[3] = {0, 0, 0};
          ext[d] = 1;
          for (k = lo[2]; k <= hi[2] + ext[2]; k++)
          for (j = lo[1]; j <= hi[1] + ext[1]; j++)
          for (i = lo[0]; i <= hi[0] + ext[0]; i++) {
            int p[3] = {i, j, k};
            p[dir] = upper ? s->end[dir] : s->beg[dir];
            s->Vs[d][IS(s,i,j,k)] = s->Vs[d][IS(s,p[0],p[1],p[2])];
          }
        }
        fill_staggered(s, dir, upper);
      }
    }
    if (is_mhd(s)) ghost_field_average(s, lo, hi);
  }
}

/* ======================================================================
 * One directional sweep: DataToSweep + Reconstruct + Riemann + CT store +
 * RightHandSide + dUc accumulation (Src/update_stage.cpp:96-187 and the
 * J/K analogues).
 * ====================================================================== */
/* one pencil (a, b) of the sweep along dir, with the calling thread's scratch */
static uint32_t sweep_pencil(gplo_state *s, int dir, int a, int b, double dtdx, double *v,
                             double *vL, double *vR, double *flux, double *cmax, double *Bn)
{
  const int nvar = s->nv, mhd = is_mhd(s);
  const int t1 = (dir + 1)%3, t2 = (dir + 2)%3;
  const int n = s->nt[dir];
  uint32_t err = 0;
  int i, nv;
  int p[3];
  p[t1] = a; p[t2] = b;
  /* DataToSweep, Src/data2sweep.cpp:75-115 */
  for (i = 0; i < n; i++) {
    p[dir] = i;
    for (nv = 0; nv < nvar; nv++) v[i*NVMAX + nv] = VC(s, nv, p[0], p[1], p[2]);
    if (mhd) Bn[i] = s->Vs[dir][IS(s, p[0], p[1], p[2])];
  }
  /* Reconstruct, Src/States/plm_states.cpp:99-141 then 184-189 */
  int ib = s->beg[dir] - 1, ie = s->end[dir] + 1;
  for (nv = 0; nv < nvar; nv++)
    for (i = ib; i <= ie; i++) {
      double dv = plm_slope(v[(i-1)*NVMAX + nv], v[i*NVMAX + nv], v[(i+1)*NVMAX + nv], s->c.limiter);
      vL[(i+1)*NVMAX + nv] = v[i*NVMAX + nv] + 0.5*dv;
      vR[i*NVMAX + nv]     = v[i*NVMAX + nv] - 0.5*dv;
    }
  if (is_rmhd(s)) {
    /* VelocityLimiter, Src/vel_limiter.hpp:14-31 (plm_states.cpp:158-161):
       a cell whose reconstructed velocity is superluminal on either side
       is flattened */
    for (i = ib; i <= ie; i++) {
      const double *ra = vR + i*NVMAX, *rb = vL + (i+1)*NVMAX;
      double v2m = ra[1]*ra[1] + ra[2]*ra[2] + ra[3]*ra[3];
      double v2p = rb[1]*rb[1] + rb[2]*rb[2] + rb[3]*rb[3];
      if (v2m >= 1.0 || v2p >= 1.0)
        for (nv = 0; nv < nvar; nv++) vR[i*NVMAX + nv] = vL[(i+1)*NVMAX + nv] = v[i*NVMAX + nv];
    }
  }
  if (mhd) {
    for (i = ib; i <= ie; i++) {
      vL[(i+1)*NVMAX + BX1 + dir] = Bn[i+1];
      vR[i*NVMAX + BX1 + dir]     = Bn[i];
    }
  }
  /* Riemann solver on faces beg..end+1 (Src/MHD/hlld.cpp:100-114) */
  for (i = s->beg[dir]; i <= s->end[dir] + 1; i++) {
    double Sl, Sr;
    const double *vl = vL + i*NVMAX, *vr = vR + i*NVMAX;
    if (is_rmhd(s)) {
      if (s->c.solver == GPLO_HLLC)      err |= rmhd_hllc(vl, vr, dir, s->c.gamma, flux + i*NVMAX, &Sl, &Sr);
      else if (s->c.solver == GPLO_HLLD) err |= rmhd_hlld(vl, vr, dir, &s->c, flux + i*NVMAX, &Sl, &Sr);
      else err |= rmhd_hll(vl, vr, dir, s->c.gamma, flux + i*NVMAX, &Sl, &Sr);
    } else if (mhd) {
      if (s->c.solver == GPLO_HLLD)      err |= mhd_hlld(vl, vr, dir, s->c.gamma, flux + i*NVMAX, &Sl, &Sr);
      else if (s->c.solver == GPLO_HLLC) err |= mhd_hllc(vl, vr, dir, s->c.gamma, flux + i*NVMAX, &Sl, &Sr);
      else err |= hll(vl, vr, dir, s->c.gamma, 1, flux + i*NVMAX, &Sl, &Sr);
    } else {
      if (s->c.solver == GPLO_HLLC) err |= hd_hllc(vl, vr, dir, s->c.gamma, flux + i*NVMAX, &Sl, &Sr);
      else err |= hll(vl, vr, dir, s->c.gamma, 0, flux + i*NVMAX, &Sl, &Sr);
    }
    cmax[i] = MAX(fabs(Sl), fabs(Sr));
    if (mhd && s->c.emf == GPLO_UCT_HLLD) {
      p[dir] = i;
      uct_hlld_store(s, vl, vr, dir, t1, t2, Sl, Sr, IC(s, p[0], p[1], p[2]));
    } else if (mhd) {
      /* CT_Store_UCTHll<dir>, Src/MHD/CT/ct_store.hpp:156-177 */
      const int VXt = 1 + t1, VXb = 1 + t2;
      p[dir] = i;
      size_t id = IC(s, p[0], p[1], p[2]);
      double aR_1 = MAX(0.0, Sr);
      double aL_1 = MIN(0.0, Sl);
      double scrh = 1.0/(aR_1 - aL_1);
      s->e[t2][dir][id] =  (aR_1*vl[VXt] - aL_1*vr[VXt])*scrh;
      s->e[t1][dir][id] = -(aR_1*vl[VXb] - aL_1*vr[VXb])*scrh;
      s->aL[dir][id] =  aR_1*scrh;
      s->aR[dir][id] = -aL_1*scrh;
      s->dR[dir][id] = -aL_1*aR_1*scrh;
      s->dL[dir][id] = -aL_1*aR_1*scrh;
    }
  }
  /* RightHandSide on the active box only, Src/MHD/rhs.cpp:330-527 */
  if (a < s->beg[t1] || a > s->end[t1] || b < s->beg[t2] || b > s->end[t2]) return err;
  for (i = s->beg[dir]; i <= s->end[dir]; i++) {
    p[dir] = i;
    size_t id = IC(s, p[0], p[1], p[2]);
    for (nv = 0; nv < nvar; nv++) {
      double rhs = -dtdx*(flux[(i+1)*NVMAX + nv] - flux[i*NVMAX + nv]);
      /* MHD adds src[] = 0 (Src/MHD/rhs.cpp:519); RMHD with CT has no
         source block (Src/RMHD/rhs.cpp:14-21,400-422) */
      if (!is_rmhd(s)) rhs += 0.0;
      /* dUc = rhs (x); dUc += rhs (y, z): Src/update_stage.cpp:173-184,272-284,369-380 */
      if (dir == 0) s->Rc[(size_t)nv*s->ncell + id] = rhs;
      else          s->Rc[(size_t)nv*s->ncell + id] += rhs;
    }
    double cdt = 0.5*(cmax[i] + cmax[i+1])/s->dx[dir];
    if (dir == 0) s->Cdt[id] = cdt;      /* rhs.cpp:523 */
    else          s->Cdt[id] += cdt;     /* rhs.cpp:756,981 */
  }
  return err;
}

/* the sweep along dir over the box widened by ntr transverse zones; pencils
   are independent (each writes its own cells), so threads split them */
static uint32_t sweep(gplo_state *s, int dir, double dt)
{
  const int mhd = is_mhd(s);
  const int t1 = (dir + 1)%3, t2 = (dir + 2)%3;
  const int ntr = mhd ? NG : 0;           /* Src/update_stage.cpp:80-90 */
  uint32_t err = 0;
  int lo[3], hi[3];
  for (int d = 0; d < 3; d++) { lo[d] = s->beg[d]; hi[d] = s->end[d]; }
  for (int d = 0; d < 3; d++) if (d != dir && d < s->c.dims) { lo[d] -= ntr; hi[d] += ntr; }
  const double dtdx = dt/s->dx[dir];
  const long na = hi[t1] - lo[t1] + 1, nb = hi[t2] - lo[t2] + 1;
  long q;
#pragma omp parallel for reduction(|:err) schedule(static)
  for (q = 0; q < na*nb; q++) {
#ifdef _OPENMP
    const size_t th = (size_t)omp_get_thread_num();
#else
    const size_t th = 0;
#endif
    const size_t o = th*s->pstride, oc = th*(s->pstride/NVMAX);
    err |= sweep_pencil(s, dir, lo[t1] + (int)(q % na), lo[t2] + (int)(q / na), dtdx,
                        s->pv + o, s->pvL + o, s->pvR + o, s->pflux + o, s->pcmax + oc, s->pBn + oc);
  }
  return err;
}

/* ArrayReconstruct<dir> linear branch, Src/array_reconstruct.cpp:46-206:
   reconstruct array Q (staggered or cell layout via 'stag') along dir for
   every pencil in the box; qL(i+1)/qR(i) for i in [1, ntot-2]. */
static void array_reconstruct(gplo_state *s, int dir, const double *Q, int stag, int lim,
                              const int lo[3], const int hi[3], double *qL, double *qR)
{
  const int t1 = (dir + 1)%3, t2 = (dir + 2)%3;
  const int n = s->nt[dir];
  int b;
#pragma omp parallel for schedule(static)
  for (b = lo[t2]; b <= hi[t2]; b++)
  for (int a = lo[t1]; a <= hi[t1]; a++) {
    int p[3];
    p[t1] = a; p[t2] = b;
    for (int i = 1; i < n - 1; i++) {
      p[dir] = i;
      size_t id = IS(s, p[0], p[1], p[2]);
      int pm[3] = {p[0], p[1], p[2]}, pp[3] = {p[0], p[1], p[2]};
      pm[dir] = i - 1; pp[dir] = i + 1;
      double q0, qm, qp;
      if (stag) { q0 = Q[id]; qm = Q[IS(s,pm[0],pm[1],pm[2])]; qp = Q[IS(s,pp[0],pp[1],pp[2])]; }
      else      { q0 = Q[IC(s,p[0],p[1],p[2])]; qm = Q[IC(s,pm[0],pm[1],pm[2])]; qp = Q[IC(s,pp[0],pp[1],pp[2])]; }
      double dq = plm_slope(qm, q0, qp, lim);
      int pn[3] = {p[0], p[1], p[2]}; pn[dir] = i + 1;
      qL[IS(s,pn[0],pn[1],pn[2])] = q0 + 0.5*dq;
      qR[id] = q0 - 0.5*dq;
    }
  }
}

/* CT_EMF_Riemann2D for UCT_HLL, Src/MHD/CT/ct_emf_average.cpp:535-989 */
static void ct_emf_uct(gplo_state *s)
{
  const int D = s->c.dims;
  int lo[3], hi[3], i, j, k;
  for (int d = 0; d < 3; d++) { lo[d] = s->beg[d]; hi[d] = s->end[d] + (d < D ? 1 : 0); }
  const int recB = GPLO_MC_LIM, recV = s->c.limiter;
  double *bL = s->qL[0], *bR = s->qR[0], *vL = s->qL[1], *vR = s->qR[1];
  double *Ex = s->Ee[0], *Ey = s->Ee[1], *Ez = s->Ee[2];
  /* (dir of reconstruction, face component Q, e array, target edge, sign of
     phi, accumulate?) -- blocks X1a, X1b, X2a, X2b, X3a, X3b */
  struct blk { int rdir, comp, ea, eb, edge, sgn, acc, cdir; int need[3]; };
  const struct blk B[6] = {
    {0, 1, 2, 1, 2, +1, 0, 1, {1,1,0}},  /* X1a: By, e[K][J] -> Ez, avg over j */
    {0, 2, 1, 2, 1, -1, 0, 2, {1,0,1}},  /* X1b: Bz, e[J][K] -> Ey, avg over k */
    {1, 2, 0, 2, 0, +1, 0, 2, {0,1,1}},  /* X2a: Bz, e[I][K] -> Ex, avg over k */
    {1, 0, 2, 0, 2, -1, 1, 0, {1,1,0}},  /* X2b: Bx, e[K][I] -> Ez +=, avg over i */
    {2, 0, 1, 0, 1, +1, 1, 0, {1,0,1}},  /* X3a: Bx, e[J][I] -> Ey +=, avg over i */
    {2, 1, 0, 1, 0, -1, 1, 1, {0,1,1}},  /* X3b: By, e[I][J] -> Ex +=, avg over j */
  };
  for (int q = 0; q < 6; q++) {
    const struct blk *bk = &B[q];
    if (bk->rdir >= D || bk->comp >= D) continue;
    int ok = 1;
    for (int d = 0; d < 3; d++) if (bk->need[d] && d >= D) ok = 0;
    if (!ok) continue;
    array_reconstruct(s, bk->rdir, s->Vs[bk->comp], 1, recB, lo, hi, bL, bR);
    array_reconstruct(s, bk->rdir, s->e[bk->ea][bk->eb], 0, recV, lo, hi, vL, vR);
    double *E = bk->edge == 0 ? Ex : bk->edge == 1 ? Ey : Ez;
    const int fd = bk->rdir;   /* face direction of the a/d coefficients */
#pragma omp parallel for private(i, j) schedule(static)
    for (k = lo[2]; k <= hi[2]; k++)
    for (j = lo[1]; j <= hi[1]; j++)
    for (i = lo[0]; i <= hi[0]; i++) {
      int m[3] = {i, j, k}; m[bk->cdir] -= 1;
      size_t id0 = IC(s, m[0], m[1], m[2]), id1 = IC(s, i, j, k);
      size_t ie = IS(s, i, j, k);
      double aL = 0.5*(s->aL[fd][id0] + s->aL[fd][id1]);
      double aR = 0.5*(s->aR[fd][id0] + s->aR[fd][id1]);
      double dL = 0.5*(s->dL[fd][id0] + s->dL[fd][id1]);
      double dR = 0.5*(s->dR[fd][id0] + s->dR[fd][id1]);
      double phi = dR*bR[ie] - dL*bL[ie];
      double val = bk->sgn > 0 ? (aL*vL[ie]*bL[ie] + aR*vR[ie]*bR[ie]) + phi
                               : (aL*vL[ie]*bL[ie] + aR*vR[ie]*bR[ie]) - phi;
      if (bk->acc) E[ie] += val; else E[ie] = val;
    }
  }
}

/* CT_Update, Src/MHD/CT/ct_update.cpp:41-430 (Cartesian) */
static void ct_update(gplo_state *s, double dt)
{
  const int D = s->c.dims;
  int i, j, k;
  double *Ex = s->Ee[0], *Ey = s->Ee[1], *Ez = s->Ee[2];
  /* Bx1 faces: i in [ibeg, iend+1] */
#pragma omp parallel for private(i, j) schedule(static)
  for (k = s->beg[2]; k <= s->end[2]; k++)
  for (j = s->beg[1]; j <= s->end[1]; j++)
  for (i = s->beg[0]; i <= s->end[0] + 1; i++) {
    double curlE = 0.0;
    if (D >= 2) { double dtdx2 = dt/s->dx[1]; curlE = 0.0 + dtdx2*(Ez[IS(s,i,j+1,k)] - Ez[IS(s,i,j,k)]); }
    if (D == 3) { double dtdx3 = dt/s->dx[2]; curlE = curlE - dtdx3*(Ey[IS(s,i,j,k+1)] - Ey[IS(s,i,j,k)]); }
    s->Rs[0][IS(s,i,j,k)] = -curlE;
  }
  if (D >= 2) {
#pragma omp parallel for private(i, j) schedule(static)
    for (k = s->beg[2]; k <= s->end[2]; k++)
    for (j = s->beg[1]; j <= s->end[1] + 1; j++)
    for (i = s->beg[0]; i <= s->end[0]; i++) {
      double dtdx1 = dt/s->dx[0];
      double curlE = - dtdx1*(Ez[IS(s,i+1,j,k)] - Ez[IS(s,i,j,k)]);
      if (D == 3) { double dtdx3 = dt/s->dx[2]; curlE = curlE + dtdx3*(Ex[IS(s,i,j,k+1)] - Ex[IS(s,i,j,k)]); }
      s->Rs[1][IS(s,i,j,k)] = -curlE;
    }
  }
  if (D == 3) {
#pragma omp parallel for private(i, j) schedule(static)
    for (k = s->beg[2]; k <= s->end[2] + 1; k++)
    for (j = s->beg[1]; j <= s->end[1]; j++)
    for (i = s->beg[0]; i <= s->end[0]; i++) {
      double dtdx1 = dt/s->dx[0], dtdx2 = dt/s->dx[1];
      double curlE = dtdx1*(Ey[IS(s,i+1,j,k)] - Ey[IS(s,i,j,k)])
                   - dtdx2*(Ex[IS(s,i,j+1,k)] - Ex[IS(s,i,j,k)]);
      s->Rs[2][IS(s,i,j,k)] = -curlE;
    }
  }
}

/* UpdateStage, Src/update_stage.cpp:57-440 */
static uint32_t update_stage(gplo_state *s, double dt, double *inv_dt)
{
  uint32_t err = 0;
  for (int d = 0; d < s->c.dims; d++) err |= sweep(s, d, dt);
  if (is_mhd(s)) { ct_emf_uct(s); ct_update(s, dt); }
  if (s->stage == 1) {
    double local = -1.0;     /* getMax, :446-458; max is exact in any order */
#pragma omp parallel for reduction(max:local) schedule(static)
    for (int k = s->beg[2]; k <= s->end[2]; k++)
    for (int j = s->beg[1]; j <= s->end[1]; j++)
    for (int i = s->beg[0]; i <= s->end[0]; i++) local = MAX(local, s->Cdt[IC(s,i,j,k)]);
    *inv_dt = local/(double)s->c.dims;
  }
  return err;
}

/* ConsToPrim MHD (Src/MHD/mappers.cpp:32-128) / HD (Src/HD/mappers.cpp:33-110)
   wrapped by ConsToPrim3D (Src/mappers3D.cpp:76-194) */
/* the zone's Data::err entry (Src/mappers3D.cpp:115,121,173,176) and its
   count in Where (Src/debug_tools.cpp:587-593) */
static void zone_report(gplo_state *s, int i, int j, int k, uint32_t err)
{
  const int nx = s->end[0] - s->beg[0] + 1, ny = s->end[1] - s->beg[1] + 1;
  const int nz = s->end[2] - s->beg[2] + 1;
  const size_t a = ((size_t)(k - s->beg[2])*ny + (size_t)(j - s->beg[1]))*nx + (size_t)(i - s->beg[0]);
  s->zerr[(size_t)(s->stage - 1)*nx*ny*nz + a] = err;
}

static uint32_t cons_to_prim_3d(gplo_state *s)
{
  const double gmm1 = s->c.gamma - 1.0;
  const int mhd = is_mhd(s), ENG = mhd ? 7 : 4;
  uint32_t err_all = 0;
  unsigned long long flagged = 0;
#pragma omp parallel for reduction(|:err_all) reduction(+:flagged) schedule(static)
  for (int k = s->beg[2]; k <= s->end[2]; k++)
  for (int j = s->beg[1]; j <= s->end[1]; j++)
  for (int i = s->beg[0]; i <= s->end[0]; i++) {
    double u[NVMAX], v[NVMAX];
    uint32_t err = 0;
    size_t id = IC(s,i,j,k);
    for (int nv = 0; nv < s->nv; nv++) {
      u[nv] = s->Uc[(size_t)nv*s->ncell + id];
      if (isnan(u[nv])) err |= ERR_NAN | ERR_QUIT;
    }
    if (is_rmhd(s)) {
      uint32_t le = rmhd_cons_to_prim(u, v, &s->c);
      err |= le;
      if (le & ERR_CONS2PRIM_FAIL) err |= ERR_QUIT;
      if (le) for (int nv = 0; nv < s->nv; nv++) s->Uc[(size_t)nv*s->ncell + id] = u[nv];
      for (int nv = 0; nv < s->nv; nv++) s->Vc[(size_t)nv*s->ncell + id] = v[nv];
      zone_report(s, i, j, k, err);
      flagged += err != 0;
      err_all |= err;
      continue;
    }
    uint32_t le = 0;
    if (u[RHO] < 0.0) { u[RHO] = s->c.small_rho; le |= ERR_NEGATIVE_RHO | ERR_CONS2PRIM_FAIL; }
    v[RHO] = u[RHO];
    double tau = 1.0/u[RHO];
    v[1] = u[1]*tau;
    v[2] = u[2]*tau;
    v[3] = u[3]*tau;
    double m2 = u[1]*u[1] + u[2]*u[2] + u[3]*u[3];
    double kin;
    if (mhd) {
      v[4] = u[4]; v[5] = u[5]; v[6] = u[6];
      double b2 = u[4]*u[4] + u[5]*u[5] + u[6]*u[6];
      kin = 0.5*(m2*tau + b2);
    } else {
      kin = 0.5*m2*tau;
    }
    if (u[ENG] > 0.0) {
      v[ENG] = gmm1*(u[ENG] - kin);
      if (v[ENG] < 0.0) {
        v[ENG] = s->c.small_prs;
        u[ENG] = v[ENG]/gmm1 + kin;
        le |= ERR_NEGATIVE_PRS;
      }
    } else {
      u[ENG] = s->c.small_prs/gmm1 + kin;
      v[ENG] = s->c.small_prs;
      le |= ERR_NEGATIVE_ENG;
    }
    err |= le;
    if (le & ERR_CONS2PRIM_FAIL) err |= ERR_QUIT;
    if (le) for (int nv = 0; nv < s->nv; nv++) s->Uc[(size_t)nv*s->ncell + id] = u[nv];
    for (int nv = 0; nv < s->nv; nv++) s->Vc[(size_t)nv*s->ncell + id] = v[nv];
    zone_report(s, i, j, k, err);
    flagged += err != 0;
    err_all |= err;
  }
  s->zcount[s->stage - 1] += flagged;
  s->zcount_total += flagged;
  return err_all;
}

/* CT_AverageStaggeredFields on the active box, Src/MHD/CT/ct_field_average.cpp:182-184,393-395 */
static void ct_average(gplo_state *s)
{
#pragma omp parallel for schedule(static)
  for (int k = s->beg[2]; k <= s->end[2]; k++)
  for (int j = s->beg[1]; j <= s->end[1]; j++)
  for (int i = s->beg[0]; i <= s->end[0]; i++)
    for (int d = 0; d < s->c.dims; d++) {
      int o[3] = {0, 0, 0}; o[d] = 1;
      UC(s, BX1 + d, i, j, k) = 0.5*(s->Vs[d][IS(s,i,j,k)] + s->Vs[d][IS(s,i+o[0],j+o[1],k+o[2])]);
    }
}

/* AdvanceStep RK2/RK3, Src/rk_step.cpp:88-720 */
uint32_t gplo_advance_step(gplo_state *s, double dt, double *inv_dt)
{
  static const double W[4][3] = { {0,0,0}, {1.0, 1.0, 1.0}, {0.5, 0.5, 0.5}, {0,0,0} };
  const int nstage = s->c.time_stepping == 3 ? 3 : 2;
  uint32_t err = 0;
  double w[4][3];
  memcpy(w, W, sizeof(w));
  if (nstage == 3) {   /* Src/rk_step.cpp:35-51 */
    w[2][0] = 0.75; w[2][1] = 0.25; w[2][2] = 0.25;
    w[3][0] = 1.0/3.0; w[3][1] = 2.0/3.0; w[3][2] = 2.0/3.0;
  }
  const int D = s->c.dims, mhd = is_mhd(s);
  s->zcount[0] = s->zcount[1] = s->zcount[2] = 0;
  for (int st = 1; st <= nstage; st++) {
    s->stage = st;
    boundary(s);
    err |= update_stage(s, dt, inv_dt);
#pragma omp parallel for collapse(2) schedule(static)
    for (int nv = 0; nv < s->nv; nv++)
    for (int k = s->beg[2]; k <= s->end[2]; k++)
    for (int j = s->beg[1]; j <= s->end[1]; j++)
    for (int i = s->beg[0]; i <= s->end[0]; i++) {
      size_t id = (size_t)nv*s->ncell + IC(s,i,j,k);
      if (st == 1) {
        s->U0[id] = s->Uc[id];
        s->Uc[id] = s->U0[id] + w[1][2]*s->Rc[id];
      } else {
        s->Uc[id] = w[st][0]*s->U0[id] + w[st][1]*s->Uc[id] + w[st][2]*s->Rc[id];
      }
    }
    if (mhd) {   /* BOX_STAG_LOOP, Src/macros.hpp:82-84 */
#pragma omp parallel for schedule(static)
      for (int k = s->beg[2]; k <= s->end[2] + (D == 3); k++)
      for (int j = s->beg[1]; j <= s->end[1] + (D >= 2); j++)
      for (int i = s->beg[0]; i <= s->end[0] + 1; i++) {
        size_t id = IS(s,i,j,k);
        for (int d = 0; d < D; d++) {
          if (st == 1) {
            s->Bs0[d][id] = s->Vs[d][id];
            s->Vs[d][id] = s->Bs0[d][id] + w[1][2]*s->Rs[d][id];
          } else {
            s->Vs[d][id] = w[st][0]*s->Bs0[d][id] + w[st][1]*s->Vs[d][id] + w[st][2]*s->Rs[d][id];
          }
        }
      }
      ct_average(s);
    }
    err |= cons_to_prim_3d(s);
  }
  return err;
}

int gplo_zone_errors(const gplo_state *s, uint32_t *err, unsigned long long *counts,
                     unsigned long long *total)
{
  const size_t na = (size_t)s->c.nx*s->c.ny*s->c.nz;
  const int nstage = s->c.time_stepping == 3 ? 3 : 2;
  if (err) memcpy(err, s->zerr, (size_t)nstage*na*sizeof(uint32_t));
  if (counts) for (int q = 0; q < nstage; q++) counts[q] = s->zcount[q];
  if (total) *total = s->zcount_total;
  return nstage;
}

uint32_t gplo_run(gplo_state *s, int nsteps, double first_dt, double *time_out, double *dt_out)
{
  double t = 0.0, dt = first_dt, inv_dt = 0.0;
  uint32_t err = 0;
  for (int n = 0; n < nsteps; n++) {
    err |= gplo_advance_step(s, dt, &inv_dt);
    t += dt;
    double dt_new = s->c.cfl/inv_dt;                  /* next_time_step.cpp:80 */
    dt = MIN(dt_new, s->c.cfl_max_var*dt);            /* next_time_step.cpp:125 */
  }
  if (time_out) *time_out = t;
  if (dt_out) *dt_out = dt;
  return err;
}

/* ======================================================================
 * Allocation and state transfer
 * ====================================================================== */
gplo_state *gplo_create(const gplo_config *cfg)
{
  /* boundary type of every side */
  gplo_config c = *cfg;
  for (int side = 0; side < 6; side++) {
    if (c.bc[side] == 0) c.bc[side] = c.periodic ? GPLO_BC_PERIODIC : GPLO_BC_OUTFLOW;
    if (c.bc[side] < GPLO_BC_PERIODIC || c.bc[side] > GPLO_BC_REFLECTIVE) return NULL;
  }
  for (int d = 0; d < c.dims; d++) {
    if ((c.bc[2*d] == GPLO_BC_PERIODIC) != (c.bc[2*d + 1] == GPLO_BC_PERIODIC)) return NULL;
  }
  /* UCT_HLLD: only with the MHD HLLD solver (Src/MHD/hll.cpp:75-76,
     Src/MHD/hllc.cpp:81-83 quit otherwise) */
  if (c.emf == 0) c.emf = GPLO_UCT_HLL;
  if (c.emf != GPLO_UCT_HLL && c.emf != GPLO_UCT_HLLD) return NULL;
  if (c.emf == GPLO_UCT_HLLD && (c.phys != GPLO_MHD || c.solver != GPLO_HLLD)) return NULL;
  gplo_state *s = (gplo_state *)calloc(1, sizeof(*s));
  s->c = c;
  if (s->c.small_rho == 0.0) s->c.small_rho = 1.e-12;   /* Src/structs.hpp:407-408 */
  if (s->c.small_prs == 0.0) s->c.small_prs = 1.e-12;
  s->nv = cfg->phys == GPLO_HD ? 5 : 8;
  int nint[3] = {cfg->nx, cfg->ny, cfg->nz};
  for (int d = 0; d < 3; d++) {
    if (d < cfg->dims) { s->nt[d] = nint[d] + 2*NG; s->beg[d] = NG; s->end[d] = NG + nint[d] - 1; }
    else               { s->nt[d] = 1; s->beg[d] = 0; s->end[d] = 0; }
    s->ns[d] = s->nt[d] + 1;
    s->dx[d] = (1.0 - 0.0)/nint[d];   /* Src/set_grid.cpp:196 on [0,1] */
  }
  s->ncell = (size_t)s->nt[0]*s->nt[1]*s->nt[2];
  s->nstag = (size_t)s->ns[0]*s->ns[1]*s->ns[2];
  size_t nc = s->ncell, ns = s->nstag;
  s->Vc = calloc(nc*s->nv, 8); s->Uc = calloc(nc*s->nv, 8);
  s->U0 = calloc(nc*s->nv, 8); s->Rc = calloc(nc*s->nv, 8);
  s->Cdt = calloc(nc, 8);
  s->zerr = calloc((size_t)3*c.nx*c.ny*c.nz, sizeof(uint32_t));
  for (int d = 0; d < 3; d++) {
    s->Vs[d] = calloc(ns, 8); s->Bs0[d] = calloc(ns, 8); s->Rs[d] = calloc(ns, 8);
    s->Ee[d] = calloc(ns, 8);
    s->aL[d] = calloc(nc, 8); s->aR[d] = calloc(nc, 8);
    s->dL[d] = calloc(nc, 8); s->dR[d] = calloc(nc, 8);
    for (int e = 0; e < 3; e++) s->e[d][e] = calloc(nc, 8);
  }
  int nmax = s->nt[0]; if (s->nt[1] > nmax) nmax = s->nt[1]; if (s->nt[2] > nmax) nmax = s->nt[2];
#ifdef _OPENMP
  s->nthr = omp_get_max_threads();
#else
  s->nthr = 1;
#endif
  const size_t nth = (size_t)s->nthr;
  s->pstride = (size_t)(nmax + 2)*NVMAX;
  s->pv = calloc(nth*s->pstride, 8); s->pvL = calloc(nth*s->pstride, 8);
  s->pvR = calloc(nth*s->pstride, 8); s->pflux = calloc(nth*s->pstride, 8);
  s->pcmax = calloc(nth*(nmax + 2), 8); s->pBn = calloc(nth*(nmax + 2), 8);
  s->prhs = calloc(s->pstride, 8);
  for (int q = 0; q < 2; q++) { s->qv[q] = calloc(ns, 8); s->qL[q] = calloc(ns, 8); s->qR[q] = calloc(ns, 8); }
  return s;
}

void gplo_destroy(gplo_state *s)
{
  if (!s) return;
  free(s->Vc); free(s->Uc); free(s->U0); free(s->Rc); free(s->Cdt); free(s->zerr);
  for (int d = 0; d < 3; d++) {
    free(s->Vs[d]); free(s->Bs0[d]); free(s->Rs[d]); free(s->Ee[d]);
    free(s->aL[d]); free(s->aR[d]); free(s->dL[d]); free(s->dR[d]);
    for (int e = 0; e < 3; e++) free(s->e[d][e]);
  }
  free(s->pv); free(s->pvL); free(s->pvR); free(s->pflux); free(s->pcmax); free(s->pBn); free(s->prhs);
  for (int q = 0; q < 2; q++) { free(s->qv[q]); free(s->qL[q]); free(s->qR[q]); }
  free(s);
}

int gplo_nvar(const gplo_state *s) { return s->nv; }

static int wrap(int a, int n) { a %= n; return a < 0 ? a + n : a; }

/* Startup semantics for the fields handed over (Src/startup.cpp:344-561):
   InitDomain fills Vc and Vs everywhere (periodic images), the cell field
   is the face average (CT_AverageStaggeredFields, :388), then
   Uc = PrimToCons(Vc) on the active box (:514-561). */
int gplo_set_state(gplo_state *s, const double *vc, const double *vs)
{
  return gplo_set_state_faces(s, vc, vs, NULL);
}

/* As gplo_set_state; on a direction whose sides are not periodic the upper
   boundary face (index n along d) comes from up, the dims slabs of
   RestartFromFile's staggered read (Src/restart.cpp:145-172 with
   ReadOutputBuffer, Src/write_data.cpp:685-723: faces 0..n along d), back
   to back, slab d shaped as the active box with n_d = 1. */
int gplo_set_state_faces(gplo_state *s, const double *vc, const double *vs, const double *up)
{
  const int nx = s->c.nx, ny = s->c.ny, nz = s->c.nz;
  const size_t na = (size_t)nx*ny*nz;
  for (int k = 0; k < s->nt[2]; k++)
  for (int j = 0; j < s->nt[1]; j++)
  for (int i = 0; i < s->nt[0]; i++) {
    int a = i - s->beg[0], b = j - s->beg[1], c = k - s->beg[2];
    if (s->c.periodic) { a = wrap(a, nx); b = wrap(b, ny); c = wrap(c, nz); }
    else {
      a = a < 0 ? 0 : (a >= nx ? nx - 1 : a);
      b = b < 0 ? 0 : (b >= ny ? ny - 1 : b);
      c = c < 0 ? 0 : (c >= nz ? nz - 1 : c);
    }
    for (int nv = 0; nv < s->nv; nv++)
      VC(s,nv,i,j,k) = vc[(size_t)nv*na + ((size_t)c*ny + b)*nx + a];
  }
  if (is_mhd(s)) {
    if (!vs) return -1;
    for (int d = 0; d < s->c.dims; d++)
      for (int k = 0; k < s->ns[2]; k++)
      for (int j = 0; j < s->ns[1]; j++)
      for (int i = 0; i < s->ns[0]; i++) {
        int a = wrap(i - s->beg[0], nx), b = wrap(j - s->beg[1], ny), c = wrap(k - s->beg[2], nz);
        const int p[3] = {i - s->beg[0], j - s->beg[1], k - s->beg[2]}, n[3] = {nx, ny, nz};
        if (up && s->c.bc[2*d] != GPLO_BC_PERIODIC && p[d] == n[d]) {
          size_t off = 0;
          for (int e = 0; e < d; e++) off += na/n[e];
          const int sx = d == 0 ? 1 : nx, sy = d == 1 ? 1 : ny;
          const int q[3] = {d == 0 ? 0 : a, d == 1 ? 0 : b, d == 2 ? 0 : c};
          s->Vs[d][IS(s,i,j,k)] = up[off + ((size_t)q[2]*sy + q[1])*sx + q[0]];
        } else {
          s->Vs[d][IS(s,i,j,k)] = vs[(size_t)d*na + ((size_t)c*ny + b)*nx + a];
        }
      }
    int lo[3] = {0, 0, 0}, hi[3] = {s->nt[0] - 1, s->nt[1] - 1, s->nt[2] - 1};
    ghost_field_average(s, lo, hi);
  }
  for (int k = s->beg[2]; k <= s->end[2]; k++)
  for (int j = s->beg[1]; j <= s->end[1]; j++)
  for (int i = s->beg[0]; i <= s->end[0]; i++) {
    double v[NVMAX], u[NVMAX];
    for (int nv = 0; nv < s->nv; nv++) v[nv] = VC(s,nv,i,j,k);
    if (is_rmhd(s))     rmhd_prim_to_cons(v, u, s->c.gamma);
    else if (is_mhd(s)) mhd_prim_to_cons(v, u, s->c.gamma);
    else                hd_prim_to_cons(v, u, s->c.gamma);
    for (int nv = 0; nv < s->nv; nv++) UC(s,nv,i,j,k) = u[nv];
  }
  return 0;
}

int gplo_get_state(const gplo_state *s, double *vc, double *vs, double *uc)
{
  const int nx = s->c.nx, ny = s->c.ny, nz = s->c.nz;
  const size_t na = (size_t)nx*ny*nz;
  for (int c = 0; c < nz; c++)
  for (int b = 0; b < ny; b++)
  for (int a = 0; a < nx; a++) {
    size_t o = ((size_t)c*ny + b)*nx + a;
    size_t id = IC(s, a + s->beg[0], b + s->beg[1], c + s->beg[2]);
    for (int nv = 0; nv < s->nv; nv++) {
      if (vc) vc[(size_t)nv*na + o] = s->Vc[(size_t)nv*s->ncell + id];
      if (uc) uc[(size_t)nv*na + o] = s->Uc[(size_t)nv*s->ncell + id];
    }
    if (vs && is_mhd(s))
      for (int d = 0; d < s->c.dims; d++)
        vs[(size_t)d*na + o] = s->Vs[d][IS(s, a + s->beg[0], b + s->beg[1], c + s->beg[2])];
  }
  return 0;
}

/* ======================================================================
 * Pointwise entry points for tests: one RMHD interface solve and one RMHD
 * con2prim, as restated above.
 * ====================================================================== */
uint32_t gplo_rmhd_face(const double *vl, const double *vr, int dir, double gamma,
                        double *flux, double *sl, double *sr)
{
  return rmhd_hll(vl, vr, dir, gamma, flux, sl, sr);
}

uint32_t gplo_face(int phys, int solver, const double *vl, const double *vr, int dir, double gamma,
                   double *flux, double *sl, double *sr)
{
  if (phys == GPLO_RMHD) {
    gplo_config c;
    memset(&c, 0, sizeof(c));
    c.gamma = gamma; c.small_rho = 1.e-12; c.small_prs = 1.e-12;   /* Src/structs.hpp:407-408 */
    if (solver == GPLO_HLLD) return rmhd_hlld(vl, vr, dir, &c, flux, sl, sr);
    return solver == GPLO_HLLC ? rmhd_hllc(vl, vr, dir, gamma, flux, sl, sr)
                               : rmhd_hll(vl, vr, dir, gamma, flux, sl, sr);
  }
  if (solver == GPLO_HLLD) return mhd_hlld(vl, vr, dir, gamma, flux, sl, sr);
  if (solver == GPLO_HLLC) return mhd_hllc(vl, vr, dir, gamma, flux, sl, sr);
  return hll(vl, vr, dir, gamma, 1, flux, sl, sr);
}

uint32_t gplo_rmhd_c2p(double *u, double *v, double gamma, double small_rho, double small_prs)
{
  gplo_config c;
  memset(&c, 0, sizeof(c));
  c.gamma = gamma; c.small_rho = small_rho; c.small_prs = small_prs;
  return rmhd_cons_to_prim(u, v, &c);
}
