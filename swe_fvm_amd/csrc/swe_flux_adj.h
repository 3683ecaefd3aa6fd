// Reverse mode of the Riemann fluxes of the step (flux_hll / flux_hllc in
// swe_device.h: Fluxes::HLL<W> and Fluxes::HLLC<W>, include/Fluxes.h:14-111,
// wave speeds src/Fluxes.cpp:5-26).  Given the adjoint Fb of the edge flux,
// each function returns g = Fb^T dF / d(Lw, Lu, Lv, Rw, Ru, Rv) by one
// reverse sweep over the forward's operations, recomputed in double.  It is
// the transpose of the dual-number derivative of the same code, with its
// conventions: branches and min / max selections are taken on the values
// (a tie keeps the first argument), sqrt has derivative 0 at 0, |x| has
// derivative +1 at 0, and the constant 0 of dmin(0, a) / dmax(0, a) carries
// no derivative.  The edge bathymetry enters only through hl = Lw - z and
// hr = Rw - z, so its adjoint is -(g[0] + g[3]).
#pragma once

#include "swe_device.h"

namespace swe {

// ElemFlux (include/SpaceDisc.h:4-18): adds the (h, hu, hv) adjoints
// from the flux adjoint fb
__host__ __device__ inline void elem_flux_vjp(double nx, double ny, double h, double hu, double hv,
                                              const double fb[3], double& hb, double& hub, double& hvb) {
  if (!is_wet(h)) return;
  const double hveln = hu * nx + hv * ny;
  const double t = hveln / h;
  const double tb = fb[1] * hu + fb[2] * hv;
  hub += fb[1] * t;
  hvb += fb[2] * t;
  hb += (fb[1] * nx + fb[2] * ny) * h;  // d(h^2 / 2) = h dh
  const double hvelnb = fb[0] + tb / h;
  hb -= t * tb / h;
  hub += hvelnb * nx;
  hvb += hvelnb * ny;
}

// Wavespeeds (src/Fluxes.cpp:5-26): adds the (ul, hl, ur, hr) adjoints from
// the adjoints of the two speeds a0, a1
template <int WS>
__host__ __device__ inline void wavespeed_vjp(double ul, double hl, double ur, double hr, double a0b,
                                              double a1b, double& ulb, double& hlb, double& urb,
                                              double& hrb) {
  const double cl = sqrt(hl), cr = sqrt(hr);
  double clb = 0., crb = 0.;
  if (WS == 0) {  // a1 = -a0 = max(|ul| + cl, |ur| + cr)
    const double x1 = fabs(ul) + cl, x2 = fabs(ur) + cr;
    const double apb = a1b - a0b;
    if (x1 < x2) {
      urb += ur < 0. ? -apb : apb;
      crb += apb;
    } else {
      ulb += ul < 0. ? -apb : apb;
      clb += apb;
    }
  } else if (WS == 1) {  // a0 = min(ul - cl, ur - cr), a1 = max(ul + cl, ur + cr)
    if ((ur - cr) < (ul - cl)) {
      urb += a0b;
      crb -= a0b;
    } else {
      ulb += a0b;
      clb -= a0b;
    }
    if ((ul + cl) < (ur + cr)) {
      urb += a1b;
      crb += a1b;
    } else {
      ulb += a1b;
      clb += a1b;
    }
  } else {  // Einfeldt: a0 = min(ul - cl, uRoe - cRoe), a1 = max(ur + cr, uRoe + cRoe)
    const double s = cl + cr;
    const double uRoe = (cl * ul + cl * ur) / s;
    const double cRoe = sqrt(0.5 * (hl + hr));
    double uRoeb = 0., cRoeb = 0.;
    if ((uRoe - cRoe) < (ul - cl)) {
      uRoeb += a0b;
      cRoeb -= a0b;
    } else {
      ulb += a0b;
      clb -= a0b;
    }
    if ((ur + cr) < (uRoe + cRoe)) {
      uRoeb += a1b;
      cRoeb += a1b;
    } else {
      urb += a1b;
      crb += a1b;
    }
    // uRoe = (cl ul + cl ur) / (cl + cr), the reference's Roe velocity
    const double nb = uRoeb / s, sb = -uRoe * uRoeb / s;
    clb += nb * ul + nb * ur + sb;
    crb += sb;
    ulb += nb * cl;
    urb += nb * cl;
    const double qb = cRoe > 0. ? cRoeb * (0.5 / cRoe) : 0.;
    hlb += 0.5 * qb;
    hrb += 0.5 * qb;
  }
  hlb += cl > 0. ? clb * (0.5 / cl) : 0.;
  hrb += cr > 0. ? crb * (0.5 / cr) : 0.;
}

// Fluxes::HLL<W> (flux_hll)
template <int WS>
__host__ __device__ inline void flux_hll_vjp(double nx, double ny, double be, const Slot& L, const Slot& R,
                                             const double Fb[3], double g[6]) {
  for (int k = 0; k < 6; ++k) g[k] = 0.;
  const double ul = L.u * nx + L.v * ny, hl = L.w - be;
  const double ur = R.u * nx + R.v * ny, hr = R.w - be;
  if (hl + hr <= 1e-10) return;
  double a0, a1;
  wavespeed<WS>(ul, hl, ur, hr, a0, a1);
  const bool lsel = a0 < 0., rsel = 0. < a1;  // al = min(0, a0), ar = max(0, a1)
  const double al = lsel ? a0 : 0., ar = rsel ? a1 : 0.;
  if (ar - al <= 1e-10) return;
  const double Ul[3] = {hl, hl * L.u, hl * L.v}, Ur[3] = {hr, hr * R.u, hr * R.v};
  double Fl[3], Fr[3];
  elem_flux(nx, ny, Ul[0], Ul[1], Ul[2], Fl[0], Fl[1], Fl[2]);
  elem_flux(nx, ny, Ur[0], Ur[1], Ur[2], Fr[0], Fr[1], Fr[2]);
  const double aa = al * ar, den = ar - al;
  double alb = 0., arb = 0., aab = 0., denb = 0.;
  double Flb[3], Frb[3], Ulb[3], Urb[3];
  for (int q = 0; q < 3; ++q) {
    // F[q] = (ar Fl - al Fr + aa (Ur - Ul)) / den
    const double Fq = (ar * Fl[q] - al * Fr[q] + aa * (Ur[q] - Ul[q])) / den;
    const double Nb = Fb[q] / den;
    denb -= Fq * Nb;
    arb += Nb * Fl[q];
    alb -= Nb * Fr[q];
    aab += Nb * (Ur[q] - Ul[q]);
    Flb[q] = Nb * ar;
    Frb[q] = -Nb * al;
    Urb[q] = Nb * aa;
    Ulb[q] = -Nb * aa;
  }
  alb += aab * ar - denb;
  arb += aab * al + denb;
  elem_flux_vjp(nx, ny, Ul[0], Ul[1], Ul[2], Flb, Ulb[0], Ulb[1], Ulb[2]);
  elem_flux_vjp(nx, ny, Ur[0], Ur[1], Ur[2], Frb, Urb[0], Urb[1], Urb[2]);
  double hlb = Ulb[0] + Ulb[1] * L.u + Ulb[2] * L.v;
  double hrb = Urb[0] + Urb[1] * R.u + Urb[2] * R.v;
  double ulb = 0., urb = 0.;
  wavespeed_vjp<WS>(ul, hl, ur, hr, lsel ? alb : 0., rsel ? arb : 0., ulb, hlb, urb, hrb);
  g[0] = hlb;
  g[1] = Ulb[1] * hl + ulb * nx;
  g[2] = Ulb[2] * hl + ulb * ny;
  g[3] = hrb;
  g[4] = Urb[1] * hr + urb * nx;
  g[5] = Urb[2] * hr + urb * ny;
}

// Fluxes::HLLC<W> (flux_hllc)
template <int WS>
__host__ __device__ inline void flux_hllc_vjp(double tx, double ty, double be, const Slot& L, const Slot& R,
                                              const double Fb[3], double g[6]) {
  for (int k = 0; k < 6; ++k) g[k] = 0.;
  const double nx = ty, ny = -tx;
  const double ul = L.u * nx + L.v * ny, vl = L.u * tx + L.v * ty, hl = L.w - be;
  const double ur = R.u * nx + R.v * ny, vr = R.u * tx + R.v * ty, hr = R.w - be;
  if (hl + hr <= 1e-10) return;
  double al, ar;
  wavespeed<WS>(ul, hl, ur, hr, al, ar);
  const double num = (ar - ur) * hr * ur - (al - ul) * hl * ul + 0.5 * (hl * hl - hr * hr);
  const double den = hr * (ar - ur) - hl * (al - ul);
  const double ustar = num / den;
  // the upwind side: R where u* <= 0, else L
  const bool right = ustar <= 0;
  const double h = right ? hr : hl, vt = right ? vr : vl, a = right ? ar : al, un = right ? ur : ul;
  const double pu = right ? R.u : L.u, pv = right ? R.v : L.v;
  const double us = vt * tx + ustar * ty, vs = vt * nx + ustar * ny;
  const double d2 = a - ustar;
  const double hs = h * (a - un) / d2;
  const double U[3] = {h, h * pu, h * pv};
  const bool ssel = right ? (0. < a) : (a < 0.);  // s = max(0, ar) or min(0, al)
  const double s = ssel ? a : 0.;
  // F = ElemFlux(U) + s (hs (1, us, vs) - U)
  double Ub[3] = {-s * Fb[0], -s * Fb[1], -s * Fb[2]};
  elem_flux_vjp(nx, ny, U[0], U[1], U[2], Fb, Ub[0], Ub[1], Ub[2]);
  const double hsb = s * (Fb[0] + Fb[1] * us + Fb[2] * vs);
  const double usb = s * Fb[1] * hs, vsb = s * Fb[2] * hs;
  double ab = ssel ? Fb[0] * (hs - U[0]) + Fb[1] * (hs * us - U[1]) + Fb[2] * (hs * vs - U[2]) : 0.;
  // hs = h (a - un) / (a - u*)
  const double Pb = hsb / d2, d2b = -hs * hsb / d2;
  const double hb = Pb * (a - un) + Ub[0] + Ub[1] * pu + Ub[2] * pv;
  ab += Pb * h + d2b;
  const double unb = -Pb * h;
  const double vtb = usb * tx + vsb * nx;
  double ustarb = -d2b + usb * ty + vsb * ny;
  double alb = 0., arb = 0., ulb = 0., urb = 0., hlb = 0., hrb = 0.;
  double vlb = 0., vrb = 0., Lub = 0., Lvb = 0., Rub = 0., Rvb = 0.;
  if (right) {
    hrb += hb; arb += ab; urb += unb; vrb += vtb; Rub += Ub[1] * h; Rvb += Ub[2] * h;
  } else {
    hlb += hb; alb += ab; ulb += unb; vlb += vtb; Lub += Ub[1] * h; Lvb += Ub[2] * h;
  }
  // u* = num / den
  const double numb = ustarb / den, denb = -ustar * ustarb / den;
  // num = (ar - ur) hr ur - (al - ul) hl ul + (hl^2 - hr^2) / 2
  arb += numb * hr * ur;
  urb += numb * ((ar - ur) * hr - hr * ur);
  hrb += numb * ((ar - ur) * ur - hr);
  alb -= numb * hl * ul;
  ulb -= numb * ((al - ul) * hl - hl * ul);
  hlb += numb * (hl - (al - ul) * ul);
  // den = hr (ar - ur) - hl (al - ul)
  hrb += denb * (ar - ur);
  arb += denb * hr;
  urb -= denb * hr;
  hlb -= denb * (al - ul);
  alb -= denb * hl;
  ulb += denb * hl;
  wavespeed_vjp<WS>(ul, hl, ur, hr, alb, arb, ulb, hlb, urb, hrb);
  g[0] = hlb;
  g[1] = Lub + ulb * nx + vlb * tx;
  g[2] = Lvb + ulb * ny + vlb * ty;
  g[3] = hrb;
  g[4] = Rub + urb * nx + vrb * tx;
  g[5] = Rvb + urb * ny + vrb * ty;
}

}  // namespace swe
