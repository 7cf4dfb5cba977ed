// Pointwise physics of the Godunov stage update, shared by every kernel.
//
// Templated on the arithmetic type R: R = double is the drop-in path and
// keeps the reference's operation order exactly (the translation unit is
// compiled with -ffp-contract=off, so a*b+c stays a rounded multiply and a
// rounded add); R = float is the mixed-precision form's Riemann/PLM type.
//
// Variable order (reference Src/MHD/mhd.hpp:21-33): rho, v1, v2, v3, B1, B2,
// B3, prs; HD (Src/HD/hd.hpp:21-28): rho, v1, v2, v3, prs.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "crmath.hpp"

#define GPL_HD __host__ __device__ __forceinline__

namespace gpl {

// Reference error bits (Src/gpluto.hpp:247-262)
constexpr uint32_t ERR_NEGATIVE_RHO   = 1u << 0;
constexpr uint32_t ERR_NEGATIVE_ENG   = 1u << 1;
constexpr uint32_t ERR_NEGATIVE_PRS   = 1u << 2;
constexpr uint32_t ERR_CONS2PRIM_FAIL = 1u << 4;
constexpr uint32_t ERR_NAN            = 1u << 11;
constexpr uint32_t ERR_QUIT           = 1u << 15;

constexpr int MC_LIM = 27, VANLEER_LIM = 26;   // Src/gpluto.hpp:335-336

template <typename R> GPL_HD R rmax(R a, R b) { return a >= b ? a : b; }   // MAX, Src/macros.hpp:191
template <typename R> GPL_HD R rmin(R a, R b) { return a <= b ? a : b; }   // MIN, Src/macros.hpp:194
template <typename R> GPL_HD R rabs(R a) { return a < R(0) ? -a : a; }
template <typename R> GPL_HD R abs_min(R a, R b) { return rabs(a) < rabs(b) ? a : b; }  // ABS_MIN
template <typename R> GPL_HD R rsqrt_(R a);
template <> GPL_HD double rsqrt_<double>(double a) { return sqrt(a); }
template <> GPL_HD float rsqrt_<float>(float a) { return sqrtf(a); }
template <typename R> GPL_HD bool risnan(R a) { return a != a; }
template <typename R> GPL_HD bool risfinite(R a) { return __builtin_isfinite(a); }   // isfinite

// PLM limited slope, Src/States/plm_states.cpp:121-133 with SET_MC_LIMITER /
// SET_VL_LIMITER on a uniform Cartesian grid (Src/plm_coeffs.hpp:92-100).
template <typename R, int LIM>
GPL_HD R plm_slope(R vm, R v0, R vp)
{
  R dvp = (vp - v0)*R(1);
  R dvm = (v0 - vm)*R(1);
  R dv;
  if (LIM == MC_LIM) {
    if (dvp*dvm > R(0)) {
      R qc = R(0.5)*(dvm + dvp), scrh = R(2)*abs_min(dvp, dvm);
      dv = abs_min(qc, scrh);
    } else dv = R(0);
  } else {
    dv = (dvp*dvm > R(0) ? R(2)*dvp*dvm/(dvp + dvm) : R(0));
  }
  return dv;
}

// ---------------------------------------------------------------------
// MHD
// ---------------------------------------------------------------------

// PrimToCons, Src/MHD/mhd.hpp:906-934
template <typename R>
GPL_HD void mhd_prim_to_cons(const R *v, R *u, R gmm1)
{
  u[0] = v[0];
  u[1] = v[0]*v[1];
  u[2] = v[0]*v[2];
  u[3] = v[0]*v[3];
  u[4] = v[4];
  u[5] = v[5];
  u[6] = v[6];
  R kinb2 = v[1]*v[1] + v[2]*v[2] + v[3]*v[3];
  kinb2 = v[0]*kinb2 + v[4]*v[4] + v[5]*v[5] + v[6]*v[6];
  kinb2 *= R(0.5);
  u[7] = kinb2 + v[7]/gmm1;
}

// Flux<dir>, Src/MHD/mhd.hpp:164-218
template <typename R, int N>
GPL_HD void mhd_flux(const R *v, const R *u, R *flux, R &pt)
{
  constexpr int VXn = 1 + N, BXn = 4 + N;
  R vB = v[1]*v[4] + v[2]*v[5] + v[3]*v[6];
  pt = v[7] + R(0.5)*(v[4]*v[4] + v[5]*v[5] + v[6]*v[6]);
  flux[0] = u[1 + N];
  flux[1] = u[1]*v[VXn] - v[4]*v[BXn];
  flux[2] = u[2]*v[VXn] - v[5]*v[BXn];
  flux[3] = u[3]*v[VXn] - v[6]*v[BXn];
  flux[4] = v[4]*v[VXn] - v[BXn]*v[1];
  flux[5] = v[5]*v[VXn] - v[BXn]*v[2];
  flux[6] = v[6]*v[VXn] - v[BXn]*v[3];
  flux[7] = (u[7] + pt)*v[VXn] - v[BXn]*vB;
  flux[1 + N] += pt;
}

// MaxSignalSpeed<dir>, Src/MHD/mhd.hpp:242-280; a2 = SoundSpeed2 (Src/eos.hpp:92)
template <typename R, int N>
GPL_HD void mhd_max_signal_speed(const R *v, R gamma, R &cmin, R &cmax)
{
  constexpr int VXn = 1 + N;
  constexpr int BXn = 4 + N, BXt = 4 + (N + 1)%3, BXb = 4 + (N + 2)%3;
  R a2 = gamma*v[7]/v[0];
  R gpr = a2*v[0];
  R b1 = v[BXn], b2 = v[BXt], b3 = v[BXb];
  R Btmag2 = b2*b2 + b3*b3;
  R Bmag2  = b1*b1 + Btmag2;
  R cf = gpr - Bmag2;
  cf = gpr + Bmag2 + rsqrt_(cf*cf + R(4)*gpr*Btmag2);
  cf = rsqrt_(R(0.5)*cf/v[0]);
  cmax = v[VXn] + cf;
  cmin = v[VXn] - cf;
}

// Result of one interface solve: hydro flux (rho, m1..3, E), the three
// magnetic flux components (kept for 2D, where B3 is cell-centred) and the
// Davis speeds.
template <typename R>
struct FaceFlux {
  R f[8];
  R Sl, Sr;
  uint32_t err;
};

// HLLD, Src/MHD/hlld.cpp:62-500 (ideal EOS, no GLM, no background field)
template <typename R, int N>
GPL_HD void mhd_hlld(const R *vl, const R *vr, R gamma, FaceFlux<R> &out)
{
  constexpr int VXn = 1 + N, VXt = 1 + (N + 1)%3, VXb = 1 + (N + 2)%3;
  constexpr int MXn = VXn, MXt = VXt, MXb = VXb;
  constexpr int BXn = 4 + N, BXt = 4 + (N + 1)%3, BXb = 4 + (N + 2)%3;
  constexpr int ENG = 7;
  const R gmm1 = gamma - R(1);
  R ul[8], ur[8], fl[8], fr[8], usl[8], usr[8];
  R ptl, ptr, sl_min, sl_max, sr_min, sr_max;
  int switch_to_hll = 0;

  mhd_prim_to_cons(vl, ul, gmm1);
  mhd_prim_to_cons(vr, ur, gmm1);
  mhd_max_signal_speed<R, N>(vl, gamma, sl_min, sl_max);    // HLL_Speed, mhd.hpp:296
  mhd_max_signal_speed<R, N>(vr, gamma, sr_min, sr_max);
  const R Sl = rmin(sl_min, sr_min);
  const R Sr = rmax(sl_max, sr_max);
  out.Sl = Sl; out.Sr = Sr;
  mhd_flux<R, N>(vl, ul, fl, ptl);
  mhd_flux<R, N>(vr, ur, fr, ptr);

  R dul, dur, Bx, Bx1, sBx, pts, SM;
  {
    R scrh = R(1)/(Sr - Sl);
    Bx1 = Bx = (Sr*vr[BXn] - Sl*vl[BXn])*scrh;
    sBx = (Bx > R(0) ? R(1) : R(-1));
    dul = Sl - vl[VXn];
    dur = Sr - vr[VXn];
    scrh = R(1)/(dur*ur[0] - dul*ul[0]);
    SM = (dur*ur[MXn] - dul*ul[MXn] - ptr + ptl)*scrh;
    pts =   dur*ur[0]*ptl - dul*ul[0]*ptr
          + vl[0]*vr[0]*dur*dul*(vr[VXn] - vl[VXn]);
    pts *= scrh;
  }
  usl[0] = ul[0]*dul/(Sl - SM);
  usr[0] = ur[0]*dur/(Sr - SM);
  if (usl[0] < R(0) || usr[0] < R(0)) switch_to_hll = 1;
  R sqrL = rsqrt_(usl[0]);
  R sqrR = rsqrt_(usr[0]);
  R S1L = SM - rabs(Bx)/sqrL;
  R S1R = SM + rabs(Bx)/sqrR;

  R *flux = out.f;
  if (Sl >= R(0)) {
    for (int nv = 0; nv < 8; nv++) flux[nv] = fl[nv];
  } else if (Sr <= R(0)) {
    for (int nv = 0; nv < 8; nv++) flux[nv] = fr[nv];
  } else if (!switch_to_hll) {
    int switch_to_hllc = 0;
    if ((S1L - Sl) <  R(1.e-4)*(SM - Sl)) switch_to_hllc = 1;
    if ((S1R - Sr) > -R(1.e-4)*(Sr - SM)) switch_to_hllc = 1;
    R scrhL, scrhR;
    if (!switch_to_hllc) {
      scrhL = (ul[0]*dul*dul - Bx*Bx)/(ul[0]*dul*(Sl - SM) - Bx*Bx);
      scrhR = (ur[0]*dur*dur - Bx*Bx)/(ur[0]*dur*(Sr - SM) - Bx*Bx);
      usl[BXn] = Bx1;
      usl[BXt] = ul[BXt]*scrhL;
      usl[BXb] = ul[BXb]*scrhL;
      usr[BXn] = Bx1;
      usr[BXt] = ur[BXt]*scrhR;
      usr[BXb] = ur[BXb]*scrhR;
    } else {
      R scrh = R(1)/(Sr - Sl);
      R ut, ub;
      ut = Sr*ur[BXt] - Sl*ul[BXt] + fl[BXt] - fr[BXt];  ut *= scrh;
      ub = Sr*ur[BXb] - Sl*ul[BXb] + fl[BXb] - fr[BXb];  ub *= scrh;
      usl[BXn] = usr[BXn] = Bx1;
      usl[BXt] = usr[BXt] = ut;
      usl[BXb] = usr[BXb] = ub;
      S1L = S1R = SM;
    }
    scrhL = Bx/(ul[0]*dul);
    scrhR = Bx/(ur[0]*dur);
    R vsL = vl[VXt] - scrhL*(usl[BXt] - ul[BXt]);
    R vsR = vr[VXt] - scrhR*(usr[BXt] - ur[BXt]);
    R wsL = vl[VXb] - scrhL*(usl[BXb] - ul[BXb]);
    R wsR = vr[VXb] - scrhR*(usr[BXb] - ur[BXb]);
    usl[MXn] = usl[0]*SM;
    usr[MXn] = usr[0]*SM;
    usl[MXt] = usl[0]*vsL;
    usr[MXt] = usr[0]*vsR;
    usl[MXb] = usl[0]*wsL;
    usr[MXb] = usr[0]*wsR;
    scrhL  = vl[VXn]*Bx1 + vl[VXt]*ul[BXt] + vl[VXb]*ul[BXb];
    scrhL -= SM*Bx1 + vsL*usl[BXt] + wsL*usl[BXb];
    usl[ENG]  = dul*ul[ENG] - ptl*vl[VXn] + pts*SM + Bx*scrhL;
    usl[ENG] /= Sl - SM;
    scrhR  = vr[VXn]*Bx1 + vr[VXt]*ur[BXt] + vr[VXb]*ur[BXb];
    scrhR -= SM*Bx1 + vsR*usr[BXt] + wsR*usr[BXb];
    usr[ENG]  = dur*ur[ENG] - ptr*vr[VXn] + pts*SM + Bx*scrhR;
    usr[ENG] /= Sr - SM;

    if (S1L >= R(0)) {
      for (int nv = 0; nv < 8; nv++) flux[nv] = fl[nv] + Sl*(usl[nv] - ul[nv]);
    } else if (S1R <= R(0)) {
      for (int nv = 0; nv < 8; nv++) flux[nv] = fr[nv] + Sr*(usr[nv] - ur[nv]);
    } else {
      R ussl[8], ussr[8];
      ussl[0] = usl[0];
      ussr[0] = usr[0];
      R vss = sqrL*vsL + sqrR*vsR + (usr[BXt] - usl[BXt])*sBx;
      vss /= sqrL + sqrR;
      R wss = sqrL*wsL + sqrR*wsR + (usr[BXb] - usl[BXb])*sBx;
      wss /= sqrL + sqrR;
      ussl[MXn] = ussl[0]*SM;
      ussr[MXn] = ussr[0]*SM;
      ussl[MXt] = ussl[0]*vss;
      ussr[MXt] = ussr[0]*vss;
      ussl[MXb] = ussl[0]*wss;
      ussr[MXb] = ussr[0]*wss;
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
      if (SM >= R(0)) {
        for (int nv = 0; nv < 8; nv++)
          flux[nv] = fl[nv] + S1L*(ussl[nv] - usl[nv]) + Sl*(usl[nv] - ul[nv]);
      } else {
        for (int nv = 0; nv < 8; nv++)
          flux[nv] = fr[nv] + S1R*(ussr[nv] - usr[nv]) + Sr*(usr[nv] - ur[nv]);
      }
    }
  } else {
    R scrh = R(1)/(Sr - Sl);
    for (int nv = 0; nv < 8; nv++) {
      flux[nv]  = Sl*Sr*(ur[nv] - ul[nv]) + Sr*fl[nv] - Sl*fr[nv];
      flux[nv] *= scrh;
    }
  }
  uint32_t err = 0;
  for (int nv = 0; nv < 8; nv++) if (risnan(flux[nv])) err |= ERR_NAN | ERR_QUIT;
  out.err = err;
}

// HLL with Davis speeds (Src/MHD/hll.cpp; formula as Src/MHD/hlld.cpp:452-461)
template <typename R, int N>
GPL_HD void mhd_hll(const R *vl, const R *vr, R gamma, FaceFlux<R> &out)
{
  const R gmm1 = gamma - R(1);
  R ul[8], ur[8], fl[8], fr[8], ptl, ptr, sl_min, sl_max, sr_min, sr_max;
  mhd_prim_to_cons(vl, ul, gmm1);
  mhd_prim_to_cons(vr, ur, gmm1);
  mhd_max_signal_speed<R, N>(vl, gamma, sl_min, sl_max);
  mhd_max_signal_speed<R, N>(vr, gamma, sr_min, sr_max);
  const R Sl = rmin(sl_min, sr_min);
  const R Sr = rmax(sl_max, sr_max);
  out.Sl = Sl; out.Sr = Sr;
  mhd_flux<R, N>(vl, ul, fl, ptl);
  mhd_flux<R, N>(vr, ur, fr, ptr);
  R *flux = out.f;
  if (Sl > R(0)) {
    for (int nv = 0; nv < 8; nv++) flux[nv] = fl[nv];
  } else if (Sr < R(0)) {
    for (int nv = 0; nv < 8; nv++) flux[nv] = fr[nv];
  } else {
    R scrh = R(1)/(Sr - Sl);
    for (int nv = 0; nv < 8; nv++) {
      flux[nv]  = Sl*Sr*(ur[nv] - ul[nv]) + Sr*fl[nv] - Sl*fr[nv];
      flux[nv] *= scrh;
    }
  }
  uint32_t err = 0;
  for (int nv = 0; nv < 8; nv++) if (risnan(flux[nv])) err |= ERR_NAN | ERR_QUIT;
  out.err = err;
}

// HLLC of Li (2005), Src/MHD/hllc.cpp:130-335 (ideal EOS, no GLM, no
// background field; HLL_SWITCH is off in the default configuration,
// Src/gpluto.hpp:877-878, so the star states are always used)
template <typename R, int N>
GPL_HD void mhd_hllc(const R *vl, const R *vr, R gamma, FaceFlux<R> &out)
{
  constexpr int VXn = 1 + N;
  constexpr int MXn = VXn, MXt = 1 + (N + 1)%3, MXb = 1 + (N + 2)%3;
  constexpr int BXn = 4 + N, BXt = 4 + (N + 1)%3, BXb = 4 + (N + 2)%3;
  constexpr int ENG = 7;
  const R gmm1 = gamma - R(1);
  R ul[8], ur[8], fl[8], fr[8], pl, pr, sl_min, sl_max, sr_min, sr_max;
  mhd_prim_to_cons(vl, ul, gmm1);
  mhd_prim_to_cons(vr, ur, gmm1);
  mhd_max_signal_speed<R, N>(vl, gamma, sl_min, sl_max);    // HLL_Speed, mhd.hpp:296
  mhd_max_signal_speed<R, N>(vr, gamma, sr_min, sr_max);
  const R Sl = rmin(sl_min, sr_min);
  const R Sr = rmax(sl_max, sr_max);
  out.Sl = Sl; out.Sr = Sr;
  mhd_flux<R, N>(vl, ul, fl, pl);     // pl, pr: total pressure
  mhd_flux<R, N>(vr, ur, fr, pr);
  R *flux = out.f;
  if (Sl >= R(0)) {
    for (int nv = 0; nv < 8; nv++) flux[nv] = fl[nv];
  } else if (Sr <= R(0)) {
    for (int nv = 0; nv < 8; nv++) flux[nv] = fr[nv];
  } else {
    // the HLL average state and flux
    R Uhll[8], Fhll[8], usl[8], usr[8];
    const R scrh = R(1)/(Sr - Sl);
    for (int nv = 0; nv < 8; nv++) {
      Uhll[nv]  = Sr*ur[nv] - Sl*ul[nv] + fl[nv] - fr[nv];
      Uhll[nv] *= scrh;
      Fhll[nv]  =   Sl*Sr*(ur[nv] - ul[nv])
                  + Sr*fl[nv] - Sl*fr[nv];
      Fhll[nv] *= scrh;
    }
    const R Bxs = Uhll[BXn], Bys = Uhll[BXt], Bzs = Uhll[BXb];
    const R vxL = vl[VXn], vxR = vr[VXn];
    const R vxs = Uhll[MXn]/Uhll[0];
    {
      const R vBL = vl[1]*vl[4] + vl[2]*vl[5] + vl[3]*vl[6];
      const R vBR = vr[1]*vr[4] + vr[2]*vr[5] + vr[3]*vr[6];
      R vBs = Uhll[4]*Uhll[1]
            + Uhll[5]*Uhll[2]
            + Uhll[6]*Uhll[3];
      vBs /= Uhll[0];
      usl[0] = ul[0]*(Sl - vxL)/(Sl - vxs);
      usr[0] = ur[0]*(Sr - vxR)/(Sr - vxs);
      const R ps = Fhll[MXn] + Bxs*Bxs - Fhll[0]*vxs;
      usl[ENG] = (ul[ENG]*(Sl - vxL) +
                  ps*vxs - pl*vxL - Bxs*vBs + vl[BXn]*vBL)/(Sl - vxs);
      usr[ENG] = (ur[ENG]*(Sr - vxR) +
                  ps*vxs - pr*vxR - Bxs*vBs + vr[BXn]*vBR)/(Sr - vxs);
    }
    usl[MXn] = usl[0]*vxs;
    usr[MXn] = usr[0]*vxs;
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
    if (vxs > R(0)) {
      for (int nv = 0; nv < 8; nv++) flux[nv] = fl[nv] + Sl*(usl[nv] - ul[nv]);
    } else {
      for (int nv = 0; nv < 8; nv++) flux[nv] = fr[nv] + Sr*(usr[nv] - ur[nv]);
    }
  }
  uint32_t err = 0;
  for (int nv = 0; nv < 8; nv++) if (risnan(flux[nv])) err |= ERR_NAN | ERR_QUIT;
  out.err = err;
}

// CT_Store_UCTHlld<dir>, Src/MHD/CT/ct_store.hpp:254-354 (ideal EOS): the
// UCT weights and diffusion coefficients from the HLLD fan, with the face's
// own normal field Bn = vl[BXn] = vr[BXn] (after the PLM override), replaced
// by the HLL values where a star density is negative.
template <typename R, int N>
GPL_HD void uct_hlld_store(const R *vl, const R *vr, R gamma, R Sl, R Sr, R &et, R &eb, R &aL,
                           R &aR, R &dL, R &dR)
{
  constexpr int VXn = 1 + N, VXt = 1 + (N + 1)%3, VXb = 1 + (N + 2)%3, MXn = VXn;
  R ul[8], ur[8];
  mhd_prim_to_cons(vl, ul, gamma - R(1));
  mhd_prim_to_cons(vr, ur, gamma - R(1));
  const R aR_1 = rmax(R(0), Sr);
  const R aL_1 = rmin(R(0), Sl);
  const R Bn = vl[4 + N];
  const R dul = Sl - vl[VXn];
  const R dur = Sr - vr[VXn];
  const R pL = vl[7] + R(0.5)*(vl[4]*vl[4] + vl[5]*vl[5] + vl[6]*vl[6]);
  const R pR = vr[7] + R(0.5)*(vr[4]*vr[4] + vr[5]*vr[5] + vr[6]*vr[6]);
  R scrh = R(1)/(dur*ur[0] - dul*ul[0]);
  const R Sc = (dur*ur[MXn] - dul*ul[MXn] - pR + pL)*scrh;
  const R usl_rho = ul[0]*dul/(Sl - Sc);
  const R usr_rho = ur[0]*dur/(Sr - Sc);
  const bool switch_to_hll = usl_rho < R(0) || usr_rho < R(0);
  const R sqrL = rsqrt_(usl_rho);
  const R sqrR = rsqrt_(usr_rho);
  const R Sal = Sc - rabs(Bn)/sqrL;
  const R Sar = Sc + rabs(Bn)/sqrR;
  const R chiL = (vl[VXn] - Sc)*(Sl - Sc)/(Sal + Sl - R(2)*Sc);
  const R chiR = (vr[VXn] - Sc)*(Sr - Sc)/(Sar + Sr - R(2)*Sc);
  const R nuL  = (Sal + Sl)/(rabs(Sal) + rabs(Sl));
  const R nuR  = (Sar + Sr)/(rabs(Sar) + rabs(Sr));
  const R nuLR = (Sal + Sar)/(rabs(Sal) + rabs(Sar));
  if (rabs(Sar - Sal) > R(1.e-9)*rabs(Sr - Sl)) {
    dL =   R(0.5)*(chiL*nuL - chiL*nuLR)
         + R(0.5)*(rabs(Sal) - nuLR*Sal);
    dR =   R(0.5)*(chiR*nuR - chiR*nuLR)
         + R(0.5)*(rabs(Sar) - nuLR*Sar);
    aL = R(0.5)*(R(1) + nuLR);
    aR = R(0.5)*(R(1) - nuLR);
  } else {             // HLLC, degenerate limit Bn -> 0
    dL = R(0.5)*chiL*nuL + R(0.5)*rabs(Sal);
    dR = R(0.5)*chiR*nuR + R(0.5)*rabs(Sar);
    aL = R(0.5);
    aR = R(0.5);
  }
  scrh = R(1)/(aR_1 - aL_1);
  if (switch_to_hll) {
    aL =  aR_1*scrh;
    aR = -aL_1*scrh;
    dR = -aL_1*aR_1*scrh;
    dL = -aL_1*aR_1*scrh;
  }
  et =  (aR_1*vl[VXt] - aL_1*vr[VXt])*scrh;
  eb = -(aR_1*vl[VXb] - aL_1*vr[VXb])*scrh;
}

// The weights aL, aR and the diffusion coefficient d of CT_Store_UCTHll
// (Src/MHD/CT/ct_store.hpp:156-177) from the clipped signal speeds
// aL_1 = min(0, Sl), aR_1 = max(0, Sr).  The sweeps store aL_1, aR_1 and the
// EMF kernels call this, so the weights come out of the same operations on
// the same values as in the reference's store.
template <typename R>
GPL_HD void uct_hll_coef(R aL_1, R aR_1, R &aL, R &aR, R &d)
{
  const R scrh = R(1)/(aR_1 - aL_1);
  aL =  aR_1*scrh;
  aR = -aL_1*scrh;
  d  = -aL_1*aR_1*scrh;
}

// CT_Store_UCTHll<dir>, Src/MHD/CT/ct_store.hpp:156-177.  Per face:
//   et = e[DIRb][DIRn] (upwinded transverse velocity v_t),
//   eb = e[DIRt][DIRn] (minus upwinded v_b),
//   aL, aR weights and the diffusion coefficient d (= dL = dR).
template <typename R, int N>
GPL_HD void uct_hll_store(const R *vl, const R *vr, R Sl, R Sr, R &et, R &eb, R &aL, R &aR, R &d)
{
  constexpr int VXt = 1 + (N + 1)%3, VXb = 1 + (N + 2)%3;
  R aR_1 = rmax(R(0), Sr);
  R aL_1 = rmin(R(0), Sl);
  R scrh = R(1)/(aR_1 - aL_1);
  et =  (aR_1*vl[VXt] - aL_1*vr[VXt])*scrh;
  eb = -(aR_1*vl[VXb] - aL_1*vr[VXb])*scrh;
  uct_hll_coef<R>(aL_1, aR_1, aL, aR, d);
}

// ConsToPrim MHD (Src/MHD/mappers.cpp:32-128) for the hydro part; the cell
// field b[] is the face average (CT_AverageStaggeredFields).  u[] holds
// rho, m1, m2, m3, E and may be rewritten by the pressure/energy floors.
GPL_HD uint32_t mhd_cons_to_prim(double *u, const double *b, double gmm1,
                                  double small_rho, double small_prs, double *v)
{
  uint32_t err = 0;
  for (int nv = 0; nv < 5; nv++) if (u[nv] != u[nv]) err |= ERR_NAN | ERR_QUIT;
  for (int nv = 0; nv < 3; nv++) if (b[nv] != b[nv]) err |= ERR_NAN | ERR_QUIT;
  uint32_t le = 0;
  if (u[0] < 0.0) { u[0] = small_rho; le |= ERR_NEGATIVE_RHO | ERR_CONS2PRIM_FAIL; }
  v[0] = u[0];
  double tau = 1.0/u[0];
  v[1] = u[1]*tau;
  v[2] = u[2]*tau;
  v[3] = u[3]*tau;
  v[4] = b[0]; v[5] = b[1]; v[6] = b[2];
  double m2 = u[1]*u[1] + u[2]*u[2] + u[3]*u[3];
  double b2 = b[0]*b[0] + b[1]*b[1] + b[2]*b[2];
  double kinb2 = 0.5*(m2*tau + b2);
  if (u[4] > 0.0) {
    v[7] = gmm1*(u[4] - kinb2);
    if (v[7] < 0.0) {
      v[7] = small_prs;
      u[4] = v[7]/gmm1 + kinb2;
      le |= ERR_NEGATIVE_PRS;
    }
  } else {
    u[4] = small_prs/gmm1 + kinb2;
    v[7] = small_prs;
    le |= ERR_NEGATIVE_ENG;
  }
  err |= le;
  if (le & ERR_CONS2PRIM_FAIL) err |= ERR_QUIT;
  return err;
}


// ---------------------------------------------------------------------
// HD (ideal EOS): variables rho, v1, v2, v3, prs; conserved rho, m1..3, E
// ---------------------------------------------------------------------

// PrimToCons, Src/HD/hd.hpp:392-410
template <typename R>
GPL_HD void hd_prim_to_cons(const R *v, R *u, R gmm1)
{
  u[0] = v[0];
  u[1] = v[0]*v[1];
  u[2] = v[0]*v[2];
  u[3] = v[0]*v[3];
  R kin = v[1]*v[1] + v[2]*v[2] + v[3]*v[3];
  kin = v[0]*kin*R(0.5);
  u[4] = kin + v[4]/gmm1;
}

// Flux<dir>, Src/HD/hd.hpp:105-125
template <typename R, int N>
GPL_HD void hd_flux(const R *v, const R *u, R *flux)
{
  constexpr int VXn = 1 + N;
  flux[0] = u[1 + N];
  flux[1] = u[1]*v[VXn];
  flux[2] = u[2]*v[VXn];
  flux[3] = u[3]*v[VXn];
  flux[4] = (u[4] + v[4])*v[VXn];
  flux[1 + N] += v[4];
}

// MaxSignalSpeed<dir>, Src/HD/hd.hpp:141-150 with SoundSpeed2 (Src/eos.hpp:84-99)
template <typename R, int N>
GPL_HD void hd_max_signal_speed(const R *v, R gamma, R &cmin, R &cmax)
{
  const R a = rsqrt_(gamma*v[4]/v[0]);
  cmax = v[1 + N] + a;
  cmin = v[1 + N] - a;
}

// HLLC, Src/HD/hllc.cpp:54-250 (ideal EOS, Davis speeds, HLL_Speed hd.hpp:165).
// out.f[0..4] = rho, m1, m2, m3, E fluxes.
template <typename R, int N>
GPL_HD void hd_hllc(const R *vl, const R *vr, R gamma, FaceFlux<R> &out)
{
  constexpr int VXn = 1 + N, VXt = 1 + (N + 1)%3, VXb = 1 + (N + 2)%3;
  const R gmm1 = gamma - R(1);
  R ul[5], ur[5], fl[5], fr[5], sl_min, sl_max, sr_min, sr_max;
  hd_prim_to_cons(vl, ul, gmm1);
  hd_prim_to_cons(vr, ur, gmm1);
  hd_flux<R, N>(vl, ul, fl);
  hd_flux<R, N>(vr, ur, fr);
  hd_max_signal_speed<R, N>(vl, gamma, sl_min, sl_max);
  hd_max_signal_speed<R, N>(vr, gamma, sr_min, sr_max);
  const R Sl = rmin(sl_min, sr_min);
  const R Sr = rmax(sl_max, sr_max);
  out.Sl = Sl; out.Sr = Sr;
  R *flux = out.f;
  if (Sl > R(0)) {
    for (int nv = 0; nv < 5; nv++) flux[nv] = fl[nv];
  } else if (Sr < R(0)) {
    for (int nv = 0; nv < 5; nv++) flux[nv] = fr[nv];
  } else {
    // contact speed (hllc.cpp:194-200)
    const R qL = vl[4] + ul[VXn]*(vl[VXn] - Sl);
    const R qR = vr[4] + ur[VXn]*(vr[VXn] - Sr);
    const R wL = vl[0]*(vl[VXn] - Sl);
    const R wR = vr[0]*(vr[VXn] - Sr);
    const R vs = (qR - qL)/(wR - wL);
    // star states of the side that is used (hllc.cpp:202-232)
    const bool left = vs >= R(0);
    const R *v = left ? vl : vr;
    const R *u = left ? ul : ur;
    const R *f = left ? fl : fr;
    const R S = left ? Sl : Sr;
    R us[5];
    us[0] = u[0]*(S - v[VXn])/(S - vs);
    us[VXn] = us[0]*vs;
    us[VXt] = us[0]*v[VXt];
    us[VXb] = us[0]*v[VXb];
    us[4] =   u[4]/v[0]
            + (vs - v[VXn])*(vs + v[4]/(v[0]*(S - v[VXn])));
    us[4] *= us[0];
    for (int nv = 0; nv < 5; nv++) flux[nv] = f[nv] + S*(us[nv] - u[nv]);
  }
  uint32_t err = 0;
  for (int nv = 0; nv < 5; nv++) if (risnan(flux[nv])) err |= ERR_NAN | ERR_QUIT;
  out.err = err;
}

// HLL with Davis speeds, Src/HD/hll.cpp (flux formula as Src/MHD/hlld.cpp:452-461)
template <typename R, int N>
GPL_HD void hd_hll(const R *vl, const R *vr, R gamma, FaceFlux<R> &out)
{
  const R gmm1 = gamma - R(1);
  R ul[5], ur[5], fl[5], fr[5], sl_min, sl_max, sr_min, sr_max;
  hd_prim_to_cons(vl, ul, gmm1);
  hd_prim_to_cons(vr, ur, gmm1);
  hd_flux<R, N>(vl, ul, fl);
  hd_flux<R, N>(vr, ur, fr);
  hd_max_signal_speed<R, N>(vl, gamma, sl_min, sl_max);
  hd_max_signal_speed<R, N>(vr, gamma, sr_min, sr_max);
  const R Sl = rmin(sl_min, sr_min);
  const R Sr = rmax(sl_max, sr_max);
  out.Sl = Sl; out.Sr = Sr;
  R *flux = out.f;
  if (Sl > R(0)) {
    for (int nv = 0; nv < 5; nv++) flux[nv] = fl[nv];
  } else if (Sr < R(0)) {
    for (int nv = 0; nv < 5; nv++) flux[nv] = fr[nv];
  } else {
    R scrh = R(1)/(Sr - Sl);
    for (int nv = 0; nv < 5; nv++) {
      flux[nv]  = Sl*Sr*(ur[nv] - ul[nv]) + Sr*fl[nv] - Sl*fr[nv];
      flux[nv] *= scrh;
    }
  }
  uint32_t err = 0;
  for (int nv = 0; nv < 5; nv++) if (risnan(flux[nv])) err |= ERR_NAN | ERR_QUIT;
  out.err = err;
}

// ConsToPrim HD (Src/HD/mappers.cpp:33-110) inside ConsToPrim3D
// (Src/mappers3D.cpp:104-194).  u[] may be rewritten by the floors.
GPL_HD uint32_t hd_cons_to_prim(double *u, double gmm1, double small_rho, double small_prs,
                                 double *v)
{
  uint32_t err = 0;
  for (int nv = 0; nv < 5; nv++) if (u[nv] != u[nv]) err |= ERR_NAN | ERR_QUIT;
  uint32_t le = 0;
  if (u[0] < 0.0) { u[0] = small_rho; le |= ERR_NEGATIVE_RHO | ERR_CONS2PRIM_FAIL; }
  v[0] = u[0];
  double tau = 1.0/u[0];
  v[1] = u[1]*tau;
  v[2] = u[2]*tau;
  v[3] = u[3]*tau;
  double m2 = u[1]*u[1] + u[2]*u[2] + u[3]*u[3];
  double kin = 0.5*m2*tau;
  if (u[4] > 0.0) {
    v[4] = gmm1*(u[4] - kin);
    if (v[4] < 0.0) {
      v[4] = small_prs;
      u[4] = v[4]/gmm1 + kin;
      le |= ERR_NEGATIVE_PRS;
    }
  } else {
    u[4] = small_prs/gmm1 + kin;
    v[4] = small_prs;
    le |= ERR_NEGATIVE_ENG;
  }
  err |= le;
  if (le & ERR_CONS2PRIM_FAIL) err |= ERR_QUIT;
  return err;
}

// ---------------------------------------------------------------------
// RMHD (ideal EOS, reduced energy E - D, exact fast speeds): variables
// rho, v1..3, B1..3, prs (Src/RMHD/rmhd.hpp:21-42)
// ---------------------------------------------------------------------

// Enthalpy and SoundSpeed2 (Src/eos.hpp:13-51, 84-125), relativistic IDEAL
template <typename R>
GPL_HD R rmhd_enthalpy(const R *v, R gamma)
{
  const R gmmr = gamma/(gamma - R(1));
  const R theta = v[7]/v[0];
  return R(1) + gmmr*theta;
}

// PrimToCons, Src/RMHD/rmhd.hpp:535-610 (B x (v x B) momentum form)
template <typename R>
GPL_HD void rmhd_prim_to_cons(const R *v, R *u, R gamma)
{
  const R gmmr = gamma/(gamma - R(1));
  const R h = rmhd_enthalpy(v, gamma);
  const R vx1 = v[1], vx2 = v[2], vx3 = v[3];
  const R vel2 = vx1*vx1 + vx2*vx2 + vx3*vx3;
  const R g2 = R(1)/(R(1) - vel2);
  const R g  = rsqrt_(g2);
  const R ex = vx2*v[6] - vx3*v[5];
  const R ey = vx3*v[4] - vx1*v[6];
  const R ez = vx1*v[5] - vx2*v[4];
  const R e2 = ex*ex + ey*ey + ez*ez;
  const R Bmag2 = v[4]*v[4] + v[5]*v[5] + v[6]*v[6];
  const R wt = v[0]*h*g2;
  u[0] = g*v[0];
  u[1] = wt*vx1 + v[5]*ez - v[6]*ey;
  u[2] = wt*vx2 + v[6]*ex - v[4]*ez;
  u[3] = wt*vx3 + v[4]*ey - v[5]*ex;
  u[4] = v[4];
  u[5] = v[5];
  u[6] = v[6];
  u[7] =   v[7]*(g2*gmmr - R(1)) + u[0]*g2*vel2/(g + R(1))
         + R(0.5)*(Bmag2 + e2);
}

// Flux<dir>, Src/RMHD/rmhd.hpp:315-366 (reduced energy); pt_out receives the
// total pressure the reference returns alongside
template <typename R, int N>
GPL_HD void rmhd_flux(const R *v, const R *u, R h, R *flux, R &pt_out)
{
  constexpr int VXn = 1 + N, VXt = 1 + (N + 1)%3, VXb = 1 + (N + 2)%3;
  constexpr int BXn = 4 + N, BXt = 4 + (N + 1)%3, BXb = 4 + (N + 2)%3;
  const R g  = u[0]/v[0];
  const R g2 = g*g;
  const R g_2 = R(1)/g2;
  const R Bmag2 = v[4]*v[4] + v[5]*v[5] + v[6]*v[6];
  const R vB = v[1]*v[4] + v[2]*v[5] + v[3]*v[6];
  R b[3];
  b[0] = g*(v[4]*g_2 + vB*v[1]);
  b[1] = g*(v[5]*g_2 + vB*v[2]);
  b[2] = g*(v[6]*g_2 + vB*v[3]);
  const R b2 = Bmag2*g_2 + vB*vB;
  const R pt = v[7] + R(0.5)*b2;
  const R wt = v[0]*h + b2;
  const R wtg2 = wt*g2;
  flux[0] = u[0]*v[VXn];
  flux[1] = wtg2*v[1]*v[VXn] - b[0]*b[N];
  flux[2] = wtg2*v[2]*v[VXn] - b[1]*b[N];
  flux[3] = wtg2*v[3]*v[VXn] - b[2]*b[N];
  flux[BXn] = R(0);
  flux[BXt] = v[VXn]*v[BXt] - v[BXn]*v[VXt];
  flux[BXb] = v[VXn]*v[BXb] - v[BXn]*v[VXb];
  flux[7] = u[1 + N];
  flux[7] -= flux[0];
  flux[1 + N] += pt;
  pt_out = pt;
}

template <typename R, int N>
GPL_HD void rmhd_flux(const R *v, const R *u, R h, R *flux)
{
  R pt;
  rmhd_flux<R, N>(v, u, h, flux, pt);
}

// CubicSolve, Src/RMHD/rmhd.hpp:390-452 (three real roots, trigonometric form)
GPL_HD void rmhd_cubic_solve(double b, double c, double d, double z[3])
{
  const double one_3 = 1.0/3.0, one_27 = 1.0/27.0, sqrt3 = sqrt(3.0);
  const double b2 = b*b;
  double Q = b2*one_3 - c*(1.0 - 1.e-16);
  const double R = b*(2.0*b2 - 9.0*c)*one_27 + d;
  Q = rmax(Q, 0.0);
  const double sQ = sqrt(Q)/sqrt3;
  double arg = -1.5*R/(Q*sQ);
  arg = rmax(-1.0, arg);
  arg = rmin( 1.0, arg);
  // acos, cos and sin correctly rounded, as the host libm returns them for
  // all but ~1e-3 of the arguments (the device library's are faithful only).
  // GPL_CUBIC_HOST_LIBM (host test builds) calls the host libm itself, so the
  // rest of the arithmetic can be compared with the oracle bit for bit.
#if defined(GPL_CUBIC_HOST_LIBM)
  const double theta = acos(arg)*one_3;
  double sn1, cs;
  sincos(theta, &sn1, &cs);   // the reference's g++ build merges cos and sin
#else
  const double theta = cr::acos_cr(arg)*one_3;
  double sn1, cs;
  cr::sincos_cr(theta, &sn1, &cs);
#endif
  const double sn = sqrt3*sn1;
  const double p  = -b*one_3;
  z[0] = -sQ*(cs + sn) + p;
  z[1] = -sQ*(cs - sn) + p;
  z[2] =  2.0*sQ*cs    + p;
}

// QuarticSolve, Src/RMHD/rmhd.hpp:455-505: the smallest and largest roots
// (the sorted z[0] and z[3]); returns false when the resolvent fails.
GPL_HD bool rmhd_quartic_extremes(double b, double c, double d, double e, double &zmin, double &zmax)
{
  const double b2 = b*b;
  const double p = c - b2*0.375;
  const double q = d + b2*b*0.125 - b*c*0.5;
  const double r = e - 3.0*b2*b2/256.0 + b2*c/16.0 - 0.25*b*d;
  double u[3];
  rmhd_cubic_solve(2.0*p, p*p - 4.0*r, -q*q, u);
  u[0] = rmax(u[0], 0.0);
  u[1] = rmax(u[1], 0.0);
  u[2] = rmax(u[2], 0.0);
  if (u[0] != u[0] || u[1] != u[1] || u[2] != u[2]) return false;
  const double sq  = -0.5*(q >= 0.0 ? 1.0 : -1.0);          // DSIGN, Src/macros.hpp:197
  const double sz1 = sq*sqrt(u[0]);
  const double sz2 = sq*sqrt(u[1]);
  const double sz3 = sq*sqrt(u[2]);
  double z[4];
  z[0] = -0.25*b + sz1 + sz2 + sz3;
  z[1] = -0.25*b + sz1 - sz2 - sz3;
  z[2] = -0.25*b - sz1 + sz2 - sz3;
  z[3] = -0.25*b - sz1 - sz2 + sz3;
  // SortArray (insertion sort, strict >): the extremes of the sorted list
  for (int j = 1; j < 4; j++) {
    const double f = z[j];
    int i = j - 1;
    while (i >= 0 && z[i] > f) { z[i + 1] = z[i]; i--; }
    z[i + 1] = f;
  }
  zmin = z[0];
  zmax = z[3];
  return true;
}

// MaxSignalSpeed<dir>, Src/RMHD/rmhd.hpp:631-752 (exact fast speeds)
template <int N>
GPL_HD void rmhd_max_signal_speed(const double *v, double cs2, double h, double &cmin, double &cmax)
{
  constexpr int VXn = 1 + N, BXn = 4 + N;
  double lfp, lfm;
  const double vB  = v[1]*v[4] + v[2]*v[5] + v[3]*v[6];
  double u02 = v[1]*v[1] + v[2]*v[2] + v[3]*v[3];
  const double Bm2 = v[4]*v[4] + v[5]*v[5] + v[6]*v[6];
  if (u02 < 1.e-14) {
    const double w_1  = 1.0/(v[0]*h + Bm2);
    const double eps2 = cs2 + Bm2*w_1*(1.0 - cs2);
    const double a0   = cs2*v[BXn]*v[BXn]*w_1;
    const double a1   = - a0 - eps2;
    double del  = a1*a1 - 4.0*a0;
    del = rmax(del, 0.0);
    del = sqrt(del);
    lfp =  sqrt(0.5*(-a1 + del));
    lfm = -lfp;
  } else {
    const double vB2 = vB*vB;
    u02 = 1.0/(1.0 - u02);
    const double b2  = Bm2/u02 + vB2;
    const double w_1 = 1.0/(v[0]*h + b2);
    const double vx  = v[VXn];
    const double vx2 = vx*vx;
    if (fabs(v[BXn]) < 1.e-14) {
      const double w     = v[0]*h;
      const double scrh1 = b2 + cs2*(w - vB2);
      const double scrh2 = w*(1.0 - cs2)*u02;
      const double del   = 4.0*scrh1*(scrh2*(1.0 - vx2) + scrh1);
      const double a2 = scrh1 + scrh2;
      const double a1 = -2.0*vx*scrh2;
      lfp = 0.5*(-a1 + sqrt(del))/a2;
      lfm = 0.5*(-a1 - sqrt(del))/a2;
    } else {
      const double scrh1 = v[BXn]/u02 + vB*vx;
      const double scrh2 = scrh1*scrh1;
      const double a2_w       = cs2*w_1;
      const double eps2       = (cs2*v[0]*h + b2)*w_1;
      const double one_m_eps2 = u02*v[0]*h*(1.0 - cs2)*w_1;
      double scrh = 2.0*(a2_w*vB*scrh1 - eps2*vx);
      const double a4 = one_m_eps2 - a2_w*vB2 + eps2;
      double a3 = - 4.0*vx*one_m_eps2 + scrh;
      double a2 =   6.0*vx2*one_m_eps2 + a2_w*(vB2 - scrh2) + eps2*(vx2 - 1.0);
      double a1 = - 4.0*vx*vx2*one_m_eps2 - scrh;
      double a0 = vx2*vx2*one_m_eps2 + a2_w*scrh2 - eps2*vx2;
      scrh = 1.0/a4;
      a3 *= scrh;
      a2 *= scrh;
      a1 *= scrh;
      a0 *= scrh;
      lfm = lfp = __builtin_nan("");
      (void)rmhd_quartic_extremes(a3, a2, a1, a0, lfm, lfp);
    }
  }
  cmax = lfp;
  cmin = lfm;
}

// HLL_Solver, Src/RMHD/hll.cpp:150-230 with HLL_Speed (Src/RMHD/rmhd.hpp:818-835):
// bL = MIN(Sl,0), bR = MAX(0,Sr), one averaged flux for every wave pattern.
template <int N>
GPL_HD void rmhd_hll(const double *vl, const double *vr, double gamma, FaceFlux<double> &out)
{
  double ul[8], ur[8], fl[8], fr[8], sl_min, sl_max, sr_min, sr_max;
  rmhd_prim_to_cons<double>(vl, ul, gamma);
  rmhd_prim_to_cons<double>(vr, ur, gamma);
  const double hL = rmhd_enthalpy<double>(vl, gamma), hR = rmhd_enthalpy<double>(vr, gamma);
  const double a2L = gamma*(vl[7]/vl[0])/hL, a2R = gamma*(vr[7]/vr[0])/hR;
  rmhd_max_signal_speed<N>(vl, a2L, hL, sl_min, sl_max);
  rmhd_max_signal_speed<N>(vr, a2R, hR, sr_min, sr_max);
  const double Sl = rmin(sl_min, sr_min);
  const double Sr = rmax(sl_max, sr_max);
  out.Sl = Sl; out.Sr = Sr;
  const double bL = rmin(Sl, 0.0);
  const double bR = rmax(0.0, Sr);
  rmhd_flux<double, N>(vl, ul, hL, fl);
  rmhd_flux<double, N>(vr, ur, hR, fr);
  const double scrh = 1.0/(bR - bL);
  double *flux = out.f;
  for (int nv = 0; nv < 8; nv++) {
    flux[nv]  = bL*bR*(ur[nv] - ul[nv]) + bR*fl[nv] - bL*fr[nv];
    flux[nv] *= scrh;
  }
  uint32_t err = 0;
  for (int nv = 0; nv < 8; nv++) if (risnan(flux[nv])) err |= ERR_NAN | ERR_QUIT;
  out.err = err;
}

// HLLC of Mignone & Bodo (2006), Src/RMHD/hllc.cpp:150-470 (ideal EOS,
// reduced energy, no GLM; HLL_SWITCH and RMHD_ENTROPY_FIX off by default,
// Src/gpluto.hpp:877-878, Src/RMHD/rmhd.hpp:143-145).  Every failed star
// construction (no real or subluminal contact speed, a superluminal or
// non-positive star state, a non-finite value) falls back to the HLL flux.
//
// Written without data-dependent branches: both star constructions (Bx -> 0
// and general) and every test are evaluated for every face, and the flux is
// selected at the end.  Each selected value is the same expression as in the
// branching form, so the bits are the reference's; on the GPU no value has to
// survive a divergent branch.  (The branching form was miscompiled for the
// GPU: the HLL components needed by faces falling back after a superluminal
// star state shared registers with the star velocities.)
template <int N>
GPL_HD void rmhd_hllc(const double *vl, const double *vr, double gamma, FaceFlux<double> &out)
{
  constexpr int VXn = 1 + N, MXn = 1 + N, MXt = 1 + (N + 1)%3, MXb = 1 + (N + 2)%3;
  constexpr int BXn = 4 + N, BXt = 4 + (N + 1)%3, BXb = 4 + (N + 2)%3;
  constexpr int ENG = 7;
  constexpr double BX_MIN = 1.e-6;                      // Src/RMHD/hllc.cpp:32
  double ul[8], ur[8], fl[8], fr[8], pl, pr, sl_min, sl_max, sr_min, sr_max;
  rmhd_prim_to_cons<double>(vl, ul, gamma);
  rmhd_prim_to_cons<double>(vr, ur, gamma);
  const double hL = rmhd_enthalpy<double>(vl, gamma), hR = rmhd_enthalpy<double>(vr, gamma);
  const double a2L = gamma*(vl[7]/vl[0])/hL, a2R = gamma*(vr[7]/vr[0])/hR;
  rmhd_max_signal_speed<N>(vl, a2L, hL, sl_min, sl_max);
  rmhd_max_signal_speed<N>(vr, a2R, hR, sr_min, sr_max);
  const double Sl = rmin(sl_min, sr_min);
  const double Sr = rmax(sl_max, sr_max);
  out.Sl = Sl; out.Sr = Sr;
  rmhd_flux<double, N>(vl, ul, hL, fl, pl);
  rmhd_flux<double, N>(vr, ur, hR, fr, pr);
  double Uhll[8], Fhll[8], Fh[8];
  const double rdS = 1.0/(Sr - Sl);
  for (int nv = 0; nv < 8; nv++) {
    Uhll[nv] = Sr*ur[nv] - Sl*ul[nv] + fl[nv] - fr[nv];
    Uhll[nv] *= rdS;
    Fhll[nv]  = Sl*Sr*(ur[nv] - ul[nv]) + Sr*fl[nv] - Sl*fr[nv];
    Fhll[nv] *= rdS;
    Fh[nv] = Fhll[nv];             // the HLL flux, the fallback
  }
  const double vxl = vl[VXn], vxr = vr[VXn];
  const double Bx = Uhll[BXn], Bys = Uhll[BXt], Bzs = Uhll[BXb];
  Uhll[ENG] += Uhll[0];             // the star states are derived with the total energy
  Fhll[ENG] += Fhll[0];
  // coefficients of the quadratic for the contact speed v* (Bx -> 0: Eq. [47])
  const bool degenerate = rabs(Bx) < BX_MIN;
  const double BtFBt_g = Uhll[BXt]*Fhll[BXt] + Uhll[BXb]*Fhll[BXb];
  const double Bt2_g   = Uhll[BXt]*Uhll[BXt] + Uhll[BXb]*Uhll[BXb];
  const double FBt2_g  = Fhll[BXt]*Fhll[BXt] + Fhll[BXb]*Fhll[BXb];
  const double BtFBt = degenerate ? 0.0 : BtFBt_g;
  const double FBt2  = degenerate ? 0.0 : FBt2_g;
  const double a = degenerate ? Fhll[ENG] : Fhll[ENG] - BtFBt_g;
  const double b = degenerate ? -(Fhll[MXn] + Uhll[ENG]) : Bt2_g + FBt2_g - (Fhll[MXn] + Uhll[ENG]);
  const double c = degenerate ? Uhll[MXn] : Uhll[MXn] - BtFBt_g;
  const double b2 = b*b;
  // the tests combine with | (not ||): no short-circuit branch
  bool to_hll = !risfinite(a) | !risfinite(b) | !risfinite(c) | !risfinite(b2) | !(b2 > 0.0);
  const double disc = 1.0 - 4.0*a*c/b2;
  to_hll |= !risfinite(disc) | (disc < 0.0);
  const double den = b*(1.0 + sqrt(disc));
  to_hll |= !risfinite(den) | (den == 0.0);
  const double vxs = -2.0*c/den;
  to_hll |= !risfinite(vxs) | !(rabs(vxs) < 1.0) | !((Sl < vxs) & (vxs < Sr));
  const double alpha_l = (Sl - vxl)/(Sl - vxs);
  const double alpha_r = (Sr - vxr)/(Sr - vxs);
  double usl[8], usr[8];
  // Bx -> 0: (vy, vz) do not enter
  double dl[8], dr[8];
  {
    const double ps = Fhll[MXn] - Fhll[ENG]*vxs;
    dl[0] = ul[0]*alpha_l;
    dr[0] = ur[0]*alpha_r;
    dl[ENG] = (Sl*ul[ENG] - fl[ENG] + ps*vxs)/(Sl - vxs);
    dr[ENG] = (Sr*ur[ENG] - fr[ENG] + ps*vxs)/(Sr - vxs);
    dl[MXn] = (dl[ENG] + ps)*vxs;
    dr[MXn] = (dr[ENG] + ps)*vxs;
    dl[MXt] = ul[MXt]*alpha_l;
    dr[MXt] = ur[MXt]*alpha_r;
    dl[MXb] = ul[MXb]*alpha_l;
    dr[MXb] = ur[MXb]*alpha_r;
    dl[MXn] += dl[0]*vxs;            // reduced energy
    dr[MXn] += dr[0]*vxs;
    dl[BXn] = Bx;
    dr[BXn] = Bx;
    dl[BXt] = ul[BXt]*alpha_l;
    dr[BXt] = ur[BXt]*alpha_r;
    dl[BXb] = ul[BXb]*alpha_l;
    dr[BXb] = ur[BXb]*alpha_r;
  }
  // general Bx
  double gl[8], gr[8];
  const double vys = (Bys*vxs - Fhll[BXt])/Bx;
  const double vzs = (Bzs*vxs - Fhll[BXb])/Bx;
  double gammas_2 = vxs*vxs + vys*vys + vzs*vzs;
  gammas_2 = 1.0 - gammas_2;
  const bool superluminal = !risfinite(gammas_2) | (gammas_2 <= 0.0);
  {
    const double vBs = vxs*Bx + vys*Bys + vzs*Bzs;
    const double ps = (BtFBt_g - Fhll[ENG])*vxs + Bx*Bx + Fhll[MXn] - FBt2_g;
    gl[0] = ul[0]*alpha_l;
    gr[0] = ur[0]*alpha_r;
    gl[ENG] = (Sl*ul[ENG] - fl[ENG] + ps*vxs - vBs*Bx)/(Sl - vxs);
    gr[ENG] = (Sr*ur[ENG] - fr[ENG] + ps*vxs - vBs*Bx)/(Sr - vxs);
    gl[MXn] = (gl[ENG] + ps)*vxs - vBs*Bx;
    gr[MXn] = (gr[ENG] + ps)*vxs - vBs*Bx;
    gl[MXt] = (Sl*ul[MXt] - fl[MXt] - Bx*(Bys*gammas_2 + vBs*vys))/(Sl - vxs);
    gr[MXt] = (Sr*ur[MXt] - fr[MXt] - Bx*(Bys*gammas_2 + vBs*vys))/(Sr - vxs);
    gl[MXb] = (Sl*ul[MXb] - fl[MXb] - Bx*(Bzs*gammas_2 + vBs*vzs))/(Sl - vxs);
    gr[MXb] = (Sr*ur[MXb] - fr[MXb] - Bx*(Bzs*gammas_2 + vBs*vzs))/(Sr - vxs);
    gl[MXn] += gl[0]*vxs;            // reduced energy
    gr[MXn] += gr[0]*vxs;
    gl[BXn] = gr[BXn] = Bx;
    gl[BXt] = gr[BXt] = Bys;
    gl[BXb] = gr[BXb] = Bzs;
  }
  to_hll |= !degenerate & superluminal;
  for (int nv = 0; nv < 8; nv++) {
    usl[nv] = degenerate ? dl[nv] : gl[nv];
    usr[nv] = degenerate ? dr[nv] : gr[nv];
  }
  to_hll |= !(usl[0] > 0.0) | !(usr[0] > 0.0);
  for (int nv = 0; nv < 8; nv++) to_hll |= !risfinite(usl[nv]) | !risfinite(usr[nv]);
  const bool left = vxs > 0.0;
  double fs[8];
  for (int nv = 0; nv < 8; nv++)
    fs[nv] = left ? fl[nv] + Sl*(usl[nv] - ul[nv]) : fr[nv] + Sr*(usr[nv] - ur[nv]);
  for (int nv = 0; nv < 8; nv++) to_hll |= !risfinite(fs[nv]);
  to_hll |= !risfinite(left ? pl : pr);
  double *flux = out.f;
  for (int nv = 0; nv < 8; nv++)
    flux[nv] = Sl >= 0.0 ? fl[nv] : (Sr <= 0.0 ? fr[nv] : (to_hll ? Fh[nv] : fs[nv]));
  uint32_t err = 0;
  for (int nv = 0; nv < 8; nv++) if (!risfinite(flux[nv])) err |= ERR_NAN | ERR_QUIT;
  out.err = err;
}

constexpr uint32_t ERR_PRESSURE_FIX          = 1u << 5;
constexpr uint32_t ERR_ENG_SOLVE_NO_SOLUTION = 1u << 7;
constexpr uint32_t ERR_SUPERLUMINAL_SPEED    = 1u << 8;
constexpr uint32_t ERR_TOO_MANY_ITERATIONS   = 1u << 9;

// Initial guess of RMHD_EnergySolve: positive root of Eq. (A27)
// (Src/RMHD/rmhd_energy_solve.cpp:86-96), bounded below by D.
GPL_HD double rmhd_guess_a27(const double *u)
{
  const double D  = u[0];
  const double E  = u[7];
  const double m2 = u[1]*u[1] + u[2]*u[2] + u[3]*u[3];
  const double B2 = u[4]*u[4] + u[5]*u[5] + u[6]*u[6];
  const double Y1 = -4.0*(E + D - B2);
  const double Y2 = m2 - 2.0*(E + D)*B2 + B2*B2;
  double chi = Y1*Y1 - 12.0*Y2;
  chi = rmax(0.0, chi);
  const double W = ( - Y1 + sqrt(chi))/6.0;
  return rmax(D, W);
}

// RMHD_EnergySolve, Src/RMHD/rmhd_energy_solve.cpp:35-209: Newton-Raphson on
// W' = W - D, tolerance 1e-11, at most 49 passes, from the Eq. (A27) guess
// (the reference) or, when W0 > 0, from max(D, W0).  u = D, m1..3, B1..3,
// E - D.  On success writes rho, v, prs of v[]; *iters = passes made.
GPL_HD uint32_t rmhd_energy_solve(const double *u, double *v, double gamma, int *iters = nullptr,
                                  double W0 = -1.0)
{
  const double acc = 1.e-11;
  const double D  = u[0];
  const double E  = u[7];
  const double S  = u[1]*u[4] + u[2]*u[5] + u[3]*u[6];
  const double m2 = u[1]*u[1] + u[2]*u[2] + u[3]*u[3];
  const double B2 = u[4]*u[4] + u[5]*u[5] + u[6]*u[6];
  const double S2 = S*S;
  if (E < 0.0) return ERR_NEGATIVE_ENG;
  double Y1, Y2, chi;
  double W = W0 > 0.0 ? rmax(D, W0) : rmhd_guess_a27(u);
  double Wp = W - D;
  int k, done = 0;
  double p = -1.0, lor = 1.0;
  for (k = 1; k < 50; k++) {
    W = Wp + D;
    const double W2    = W*W;
    const double S2_W2 = S2/W2;
    Y1 = 1.0/(W + B2);
    Y2 = Y1*Y1;
    const double vel2 = S2_W2*Y1*(Y1*W + 1.0) + m2*Y2;
    const double one_m_v2 = 1.0 - vel2;
    const double lor2 = 1.0/one_m_v2;
    lor = sqrt(lor2);
    if (vel2 > 1.0) { if (iters) *iters = k; return ERR_ENG_SOLVE_NO_SOLUTION; }
    chi = Wp/lor2 - D*vel2/(lor + 1.0);
    const double dv2_dW  = -2.0*Y2*(3.0*S2_W2 + Y1*(S2_W2*B2*B2/W + m2));
    const double dchi_dW =  one_m_v2 - 0.5*lor*(D + 2.0*chi*lor)*dv2_dW;
    const double drho_dW = -0.5*D*lor*dv2_dW;
    const double dp_dchi = (gamma - 1.0)/gamma;
    const double dp_drho = 0.0;
    p = chi*dp_dchi;
    if (done) break;
    const double dp  = dp_dchi*dchi_dW + dp_drho*drho_dW;
    const double fW  = Wp + 0.5*(B2 + (B2*m2 - S2)*Y2) - (E + p);
    const double dfW = 1.0 - dp - (B2*m2 - S2)*Y2*Y1;
    const double dW  = fW/dfW;
    Wp -= dW;
    if (fabs(dW) < acc*Wp || fabs(fW) < acc) done = 1;
  }
  if (iters) *iters = k;
  if (p < 0.0)  return ERR_NEGATIVE_PRS;
  if (k == 50)  return ERR_TOO_MANY_ITERATIONS;
  if (p != p)   return ERR_NAN;
  W = Wp + D;
  v[0] = u[0]/lor;
  v[7] = p;
  Y1 = 1.0/(W + B2);
  const double scrh = S/W;
  v[1] = Y1*(u[1] + scrh*u[4]);
  v[2] = Y1*(u[2] + scrh*u[5]);
  v[3] = Y1*(u[3] + scrh*u[6]);
  return 0;
}

// RMHD_PressureFix, Src/RMHD/rmhd_pressure_fix.cpp:22-179: p = small_prs,
// bisection on v^2 of Eq. (A3), then E recomputed into u[7].
GPL_HD uint32_t rmhd_pressure_fix(double *u, double *v, double gamma, double small_prs)
{
  const double D  = u[0];
  const double S  = u[1]*u[4] + u[2]*u[5] + u[3]*u[6];
  const double m2 = u[1]*u[1] + u[2]*u[2] + u[3]*u[3];
  const double B2 = u[4]*u[4] + u[5]*u[5] + u[6]*u[6];
  const double S2 = S*S;
  const double p = small_prs;
  double v2min, v2max, v2, lor, plor, W, W2, f, fmin;
  v2min = v2 = 0.0;
  lor  = 1.0/sqrt(1.0 - v2);
  plor = p*lor;
  W = (D + plor*gamma/(gamma - 1.0))*lor;
  W2 = W*W;
  f  = (S2*(2.0*W + B2) + m2*W2)/((W + B2)*(W + B2)*W2);
  f -= v2;
  fmin = f;
  v2max = v2 = 1.0-1.e-9;
  int k;
  for (k = 1; k < 50; k++) {
    v2 = 0.5*(v2min + v2max);
    lor  = 1.0/sqrt(1.0 - v2);
    plor = p*lor;
    W = (D + plor*gamma/(gamma - 1.0))*lor;
    W2 = W*W;
    f  = (S2*(2.0*W + B2) + m2*W2)/((W + B2)*(W + B2)*W2);
    f -= v2;
    if (f*fmin > 0.0) { v2min = v2; fmin = f; }
    else              { v2max = v2; }
    if (fabs(f) < 1.e-9) break;
  }
  if (v2 >= 1.0 && k == 50) return ERR_PRESSURE_FIX;
  v[0] = D/lor;
  v[7] = p;
  const double w_1  = 1.0/(W + B2);
  const double scrh = S/W;
  v[1] = w_1*(u[1] + scrh*u[4]);
  v[2] = w_1*(u[2] + scrh*u[5]);
  v[3] = w_1*(u[3] + scrh*u[6]);
  const double S2_W2 = S2/W2;
  u[7] = W - D - p + 0.5*(1.0 + v2)*B2 - 0.5*S2_W2;
  return 0;
}

// ConsToPrim RMHD (Src/RMHD/mappers.cpp:33-157) inside ConsToPrim3D
// (Src/mappers3D.cpp:104-194): u[8] = D, m1..3, B1..3, E - D, rewritten by the
// recovery branches; returns the mapper's bits plus the NaN check's.
GPL_HD uint32_t rmhd_cons_to_prim(double *u, double *v, double gamma, double small_rho,
                                  double small_prs, uint32_t &local, bool guided = false,
                                  double guess_scale = 1.0)
{
  uint32_t nan_err = 0;
  for (int nv = 0; nv < 8; nv++) if (u[nv] != u[nv]) nan_err |= ERR_NAN | ERR_QUIT;
  uint32_t errm = 0;
  bool fix = false;
  v[4] = u[4]; v[5] = u[5]; v[6] = u[6];
  if (u[0] < 0.0) { u[0] = small_rho; errm |= ERR_NEGATIVE_RHO; }
  if (u[7] < 0.0) { fix = true; errm |= ERR_NEGATIVE_ENG; }
  if (!fix) {
    // guided: Newton from guess_scale * W_A27 first, the A27 start as the fallback
    uint32_t e = guided ? rmhd_energy_solve(u, v, gamma, nullptr, guess_scale*rmhd_guess_a27(u)) : 1u;
    if (e) e = rmhd_energy_solve(u, v, gamma);
    if (e) fix = true;
    errm |= e;
  }
  if (fix) {
    const uint32_t e = rmhd_pressure_fix(u, v, gamma, small_prs);
    if (e) { errm |= e; errm |= ERR_CONS2PRIM_FAIL; }
  }
  const double scrh = v[1]*v[1] + v[2]*v[2] + v[3]*v[3];
  if (scrh >= 1.0) { errm |= ERR_SUPERLUMINAL_SPEED; errm |= ERR_CONS2PRIM_FAIL; }
  local = errm;
  uint32_t err = nan_err | errm;
  if (errm & ERR_CONS2PRIM_FAIL) err |= ERR_QUIT;
  return err;
}

// HLLD of Mignone, Ugliano & Bodo (2009), Src/RMHD/hlld.cpp:95-872 (ideal
// EOS, reduced energy, no GLM; HLL_SWITCH and RMHD_ENTROPY_FIX off by
// default).  The state behind each fast wave is a function of the total
// pressure p; a secant iteration on p closes the normal-velocity jump across
// the contact; any failure gives the HLL flux.
struct RhlldState {          // Riemann_State (Src/structs.hpp)
  double S, R[8], u[8];
  double vx, vy, vz, Bx, By, Bz, w, sw, Kx, Ky, Kz, K2, Sa;
  bool fail;
};

// HLLD_GetRiemannState, Src/RMHD/hlld.cpp:639-731: false when w < 0 (the
// other members then keep their values)
template <int N>
GPL_HD bool rhlld_riemann_state(RhlldState &P, double p, double Bx, int side)
{
  constexpr int MXn = 1 + N, MXt = 1 + (N + 1)%3, MXb = 1 + (N + 2)%3;
  constexpr int BXn = 4 + N, BXt = 4 + (N + 1)%3, BXb = 4 + (N + 2)%3, ENG = 7;
  const double S = P.S;
  const double *R = P.R;
  const double A = R[MXn] + p*(1.0 - S*S) - S*R[ENG];           // Eq. [26]
  const double G = R[BXt]*R[BXt] + R[BXb]*R[BXb];               // Eq. [27]
  const double C = R[BXt]*R[MXt] + R[BXb]*R[MXb];               // Eq. [28]
  const double X = Bx*(A*S*Bx + C) - (A + G)*(S*p + R[ENG]);    // Eq. [30]
  const double vx = ( Bx*(A*Bx + C*S) - (R[MXn] + p)*(G + A) );
  const double vy = ( - (A + G - Bx*Bx*(1.0 - S*S))*R[MXt]
                      + R[BXt]*(C + Bx*(S*R[MXn] - R[ENG])) );
  const double vz = ( - (A + G - Bx*Bx*(1.0 - S*S))*R[MXb]
                      + R[BXb]*(C + Bx*(S*R[MXn] - R[ENG])) );
  double scrh = vx*R[MXn] + vy*R[MXt] + vz*R[MXb];
  scrh = X*R[ENG] - scrh;
  P.w = p + scrh/(X*S - vx);                                    // Eq. [31]
  if (P.w < 0.0) return false;
  P.vx = vx/X;
  P.vy = vy/X;
  P.vz = vz/X;
  P.Bx = Bx;
  P.By = -(R[BXt]*(S*p + R[ENG]) - Bx*R[MXt])/A;
  P.Bz = -(R[BXb]*(S*p + R[ENG]) - Bx*R[MXb])/A;
  double s = Bx > 0.0 ? 1.0 : -1.0;
  if (side < 0) s *= -1.0;
  P.sw = s*sqrt(P.w);
  scrh = 1.0/(S*p +  R[ENG] + Bx*P.sw);
  P.Kx = scrh*(R[MXn] + p + R[BXn]*P.sw);
  P.Ky = scrh*(R[MXt]     + R[BXt]*P.sw);
  P.Kz = scrh*(R[MXb]     + R[BXb]*P.sw);
  P.K2 = P.Kx*P.Kx + P.Ky*P.Ky + P.Kz*P.Kz;
  return true;
}

// HLLD_Fstar, Src/RMHD/hlld.cpp:560-636: the normal velocity jump across
// the contact at total pressure p (Eq. [48]); L.fail flags an unphysical fan
template <int N>
GPL_HD double rhlld_fstar(RhlldState &L, RhlldState &Rs, double p, double Bx, double &Sc)
{
  rhlld_riemann_state<N>(L, p, Bx, -1);     // (their own results are overwritten below,
  rhlld_riemann_state<N>(Rs, p, Bx,  1);    //  as in the reference)
  const double dK  = Rs.Kx - L.Kx + 1.e-12;
  const double Bxc = Bx*dK;
  const double Byc =   Rs.By*(Rs.Kx - Rs.vx)
                     - L.By*(L.Kx - L.vx)
                     + Bx*(Rs.vy - L.vy);
  const double Bzc =   Rs.Bz*(Rs.Kx - Rs.vx)
                     - L.Bz*(L.Kx - L.vx)
                     + Bx*(Rs.vz - L.vz);
  const double KLBc = L.Kx*Bxc + L.Ky*Byc + L.Kz*Bzc;
  const double KRBc = Rs.Kx*Bxc + Rs.Ky*Byc + Rs.Kz*Bzc;
  const double vxcL = L.Kx - dK*Bx*(1.0 - L.K2)/(L.sw*dK - KLBc);
  const double vxcR = Rs.Kx - dK*Bx*(1.0 - Rs.K2)/(Rs.sw*dK - KRBc);
  L.Sa  = L.Kx;
  Rs.Sa = Rs.Kx;
  Sc    = 0.5*(vxcL + vxcR);
  bool ok = (vxcL - L.Kx) > -1.e-6;
  ok = ok && (Rs.Kx - vxcR) > -1.e-6;
  ok = ok && (L.S - L.vx) < 0.0;
  ok = ok && (Rs.S - Rs.vx) > 0.0;
  ok = ok && (Rs.w - p) > 0.0;
  ok = ok && (L.w - p) > 0.0;
  ok = ok && (L.Sa - L.S) > -1.e-6;
  ok = ok && (Rs.S - Rs.Sa) > -1.e-6;
  L.fail = !ok;
  return vxcL - vxcR;
}

// HLLD_GetAState, Src/RMHD/hlld.cpp:734-768: the state behind a fast wave
template <int N>
GPL_HD void rhlld_astate(RhlldState &P, double p, double Bx)
{
  constexpr int MXn = 1 + N, MXt = 1 + (N + 1)%3, MXb = 1 + (N + 2)%3;
  constexpr int BXn = 4 + N, BXt = 4 + (N + 1)%3, BXb = 4 + (N + 2)%3, ENG = 7;
  double *ua = P.u;
  const double *R = P.R;
  const double scrh = 1.0/(P.S - P.vx);
  ua[0]   =  R[0]*scrh;
  ua[BXn] =  Bx;
  ua[BXt] = (R[BXt] - Bx*P.vy)*scrh;
  ua[BXb] = (R[BXb] - Bx*P.vz)*scrh;
  const double vB = P.vx*ua[BXn] + P.vy*ua[BXt] + P.vz*ua[BXb];
  ua[ENG] = (R[ENG] + p*P.vx - vB*Bx)*scrh;
  ua[MXn] = (ua[ENG] + p)*P.vx - vB*ua[BXn];
  ua[MXt] = (ua[ENG] + p)*P.vy - vB*ua[BXt];
  ua[MXb] = (ua[ENG] + p)*P.vz - vB*ua[BXb];
}

// HLLD_GetCState, Src/RMHD/hlld.cpp:771-855: the state across the contact
template <int N>
GPL_HD void rhlld_cstate(RhlldState &L, RhlldState &Rs, double p, double Bx, double *Uc)
{
  constexpr int MXn = 1 + N, MXt = 1 + (N + 1)%3, MXb = 1 + (N + 2)%3;
  constexpr int BXn = 4 + N, BXt = 4 + (N + 1)%3, BXb = 4 + (N + 2)%3, ENG = 7;
  rhlld_astate<N>(L, p, Bx);
  rhlld_astate<N>(Rs, p, Bx);
  const double dK = (Rs.Kx - L.Kx) + 1.e-12;
  double Byc =   Rs.By*(Rs.Kx - Rs.vx)                   // Eq. [45]
               - L.By*(L.Kx - L.vx)
               + Bx*(Rs.vy - L.vy);
  double Bzc =   Rs.Bz*(Rs.Kx - Rs.vx)
               - L.Bz*(L.Kx - L.vx)
               + Bx*(Rs.vz - L.vz);
  const double Bxc = Bx;
  Byc /= dK;
  Bzc /= dK;
  Uc[BXn] = Bxc;
  Uc[BXt] = Byc;
  Uc[BXb] = Bzc;
  const double KLBc = L.Kx*Bxc + L.Ky*Byc + L.Kz*Bzc;
  const double KRBc = Rs.Kx*Bxc + Rs.Ky*Byc + Rs.Kz*Bzc;
  const double scrhL = (1.0 - L.K2)/(L.sw - KLBc);
  const double scrhR = (1.0 - Rs.K2)/(Rs.sw - KRBc);
  const double vxcL = L.Kx - Uc[BXn]*scrhL;              // Eq. [47]
  const double vxcR = Rs.Kx - Uc[BXn]*scrhR;
  const double vycL = L.Ky - Uc[BXt]*scrhL;
  const double vycR = Rs.Ky - Uc[BXt]*scrhR;
  const double vzcL = L.Kz - Uc[BXb]*scrhL;
  const double vzcR = Rs.Kz - Uc[BXb]*scrhR;
  const double vxc = 0.5*(vxcL + vxcR);
  const double vyc = 0.5*(vycL + vycR);
  const double vzc = 0.5*(vzcL + vzcR);
  RhlldState &A = vxc > 0.0 ? L : Rs;
  rhlld_astate<N>(A, p, Bx);
  const double *ua = A.u;
  const double Sa = A.Sa, vxa = A.vx;
  const double vBc = vxc*Uc[BXn] + vyc*Uc[BXt] + vzc*Uc[BXb];
  Uc[0]   = ua[0]*(Sa - vxa)/(Sa - vxc);                        // Eq. [32]
  Uc[ENG] = (Sa*ua[ENG] - ua[MXn] + p*vxc - vBc*Bx)/(Sa - vxc); // Eq. [33]
  Uc[MXn] = (Uc[ENG] + p)*vxc - vBc*Bx;                         // Eq. [34]
  Uc[MXt] = (Uc[ENG] + p)*vyc - vBc*Uc[BXt];
  Uc[MXb] = (Uc[ENG] + p)*vzc - vBc*Uc[BXb];
}

template <int N>
GPL_HD void rmhd_hlld(const double *vl, const double *vr, double gamma, double small_rho,
                      double small_prs, FaceFlux<double> &out)
{
  constexpr int MXn = 1 + N, BXn = 4 + N, ENG = 7;
  double ul[8], ur[8], fl[8], fr[8], pl, pr, sl_min, sl_max, sr_min, sr_max;
  uint32_t err = 0;
  rmhd_prim_to_cons<double>(vl, ul, gamma);
  rmhd_prim_to_cons<double>(vr, ur, gamma);
  const double hL = rmhd_enthalpy<double>(vl, gamma), hR = rmhd_enthalpy<double>(vr, gamma);
  const double a2L = gamma*(vl[7]/vl[0])/hL, a2R = gamma*(vr[7]/vr[0])/hR;
  rmhd_flux<double, N>(vl, ul, hL, fl, pl);
  rmhd_flux<double, N>(vr, ur, hR, fr, pr);
  rmhd_max_signal_speed<N>(vl, a2L, hL, sl_min, sl_max);
  rmhd_max_signal_speed<N>(vr, a2R, hR, sr_min, sr_max);
  const double Sl = rmin(sl_min, sr_min);
  const double Sr = rmax(sl_max, sr_max);
  out.Sl = Sl; out.Sr = Sr;
  double *flux = out.f;
  if (Sl >= 0.0) {
    for (int nv = 0; nv < 8; nv++) flux[nv] = fl[nv];
  } else if (Sr <= 0.0) {
    for (int nv = 0; nv < 8; nv++) flux[nv] = fr[nv];
  } else {
    bool to_hll = false;
    double Uhll[8], Fhll[8];
    const double dS_1 = 1.0/(Sr - Sl);
    for (int nv = 0; nv < 8; nv++) {
      Uhll[nv]  = Sr*ur[nv] - Sl*ul[nv] + fl[nv] - fr[nv];
      Uhll[nv] *= dS_1;
      Fhll[nv]  = Sr*fl[nv] - Sl*fr[nv] + Sl*Sr*(ur[nv] - ul[nv]);
      Fhll[nv] *= dS_1;
    }
    RhlldState L = {}, Rs = {};
    L.S = Sl;
    Rs.S = Sr;
    const double Bx = Uhll[BXn];
    for (int nv = 0; nv < 8; nv++) {
      L.R[nv]  = Sl*ul[nv] - fl[nv];
      Rs.R[nv] = Sr*ur[nv] - fr[nv];
    }
    L.R[ENG]  += L.R[0];                 // reduced energy
    Rs.R[ENG] += Rs.R[0];
    // initial guess: the B -> 0 limit, or the total pressure of the HLL state
    double p0, scrh = rmax(pl, pr);
    if (Bx*Bx/scrh < 0.01) {
      const double a = Sr - Sl;
      const double b = Rs.R[ENG] - L.R[ENG] + Sr*L.R[MXn] - Sl*Rs.R[MXn];
      const double c = L.R[MXn]*Rs.R[ENG] - Rs.R[MXn]*L.R[ENG];
      scrh = b*b - 4.0*a*c;
      scrh = rmax(scrh, 0.0);
      p0 = 0.5*(- b + sqrt(scrh))*dS_1;
    } else {
      double Vhll[8];
      uint32_t local = 0;
      rmhd_cons_to_prim(Uhll, Vhll, gamma, small_rho, small_prs, local);   // HLLD_Solver :343-351
      err |= local;
      if (local) to_hll = true;
      // HLLD_TotalPressure, Src/RMHD/hlld.cpp:857-872
      const double Bmag2 = Vhll[4]*Vhll[4] + Vhll[5]*Vhll[5] + Vhll[6]*Vhll[6];
      const double vel2  = Vhll[1]*Vhll[1] + Vhll[2]*Vhll[2] + Vhll[3]*Vhll[3];
      const double vB    = Vhll[1]*Vhll[4] + Vhll[2]*Vhll[5] + Vhll[3]*Vhll[6];
      p0 = Vhll[7] + 0.5*(Bmag2*(1.0 - vel2) + vB*vB);
    }
    double Sc = 0.0, p;
    double f0 = rhlld_fstar<N>(L, Rs, p0, Bx, Sc);
    if (risnan(f0) || L.fail) to_hll = true;
    if (rabs(f0) > 1.e-12 && !to_hll) {
      p = 1.025*p0;
      for (int k1 = 1; k1 < 20; k1++) {           // MAX_ITER, Src/RMHD/hlld.cpp:45
        const double f = rhlld_fstar<N>(L, Rs, p, Bx, Sc);
        if (f != f || L.fail || (k1 > 7) || (rabs(f) > rabs(f0) && k1 > 4)) {
          to_hll = true;
          break;
        }
        const double dp = (p - p0)/(f - f0)*f;
        p0 = p; f0 = f;
        p -= dp;
        if (p < 0.0) p = 1.e-6;
        if (rabs(dp) < 1.e-5*p || rabs(f) < 1.e-6) break;
      }
    } else {
      p = p0;
    }
    if (L.fail) to_hll = true;
    if (!to_hll) {
      if (L.Sa >= -1.e-6) {
        rhlld_astate<N>(L, p, Bx);
        L.u[ENG] -= L.u[0];
        for (int nv = 0; nv < 8; nv++) flux[nv] = fl[nv] + Sl*(L.u[nv] - ul[nv]);
      } else if (Rs.Sa <= 1.e-6) {
        rhlld_astate<N>(Rs, p, Bx);
        Rs.u[ENG] -= Rs.u[0];
        for (int nv = 0; nv < 8; nv++) flux[nv] = fr[nv] + Sr*(Rs.u[nv] - ur[nv]);
      } else {
        double Uc[8];
        rhlld_cstate<N>(L, Rs, p, Bx, Uc);
        if (Sc > 0.0) {
          L.u[ENG] -= L.u[0];
          Uc[ENG]  -= Uc[0];
          for (int nv = 0; nv < 8; nv++)
            flux[nv] = fl[nv] + Sl*(L.u[nv] - ul[nv]) + L.Sa*(Uc[nv] - L.u[nv]);
        } else {
          Rs.u[ENG] -= Rs.u[0];
          Uc[ENG]   -= Uc[0];
          for (int nv = 0; nv < 8; nv++)
            flux[nv] = fr[nv] + Sr*(Rs.u[nv] - ur[nv]) + Rs.Sa*(Uc[nv] - Rs.u[nv]);
        }
      }
    } else {
      for (int nv = 0; nv < 8; nv++) flux[nv] = Fhll[nv];
    }
  }
  for (int nv = 0; nv < 8; nv++) if (risnan(flux[nv])) err |= ERR_NAN | ERR_QUIT;
  out.err = err;
}

}  // namespace gpl
