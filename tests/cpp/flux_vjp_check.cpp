// Host check of the reverse-mode Riemann fluxes (swe_flux_adj.h) against the
// dual-number derivative of the forward flux templates (swe_device.h): for
// random edge states, g = Fb^T dF/d(Lw, Lu, Lv, Rw, Ru, Rv) from the reverse
// sweep must equal the Dual<6> Jacobian contracted with Fb.  The states mix
// dry sides (h = 0 exactly), films at the wet threshold, thin and deep water,
// sub- and supersonic velocities, and mirrored or identical sides (ties of
// the wave-speed min / max).  Agreement: 1e-12 of the largest term
// Fb_q dF_q/dx_k, 1e-6 where a side holds a film (0 < h < 1e-3).  Prints
// one line per flux and wave speed and exits nonzero on a mismatch.
#include <cmath>
#include <cstdio>
#include <random>

#include "swe_flux_adj.h"

using namespace swe;

struct Rng {
  std::mt19937_64 g;
  double uni(double a, double b) { return std::uniform_real_distribution<double>(a, b)(g); }
  double nrm(double s) { return std::normal_distribution<double>(0., s)(g); }
  int pick(int n) { return int(g() % uint64_t(n)); }
};

static double depth(Rng& r) {
  switch (r.pick(6)) {
    case 0: return 0.;
    case 1: return r.uni(0.5e-12, 2e-12);  // around IsWet's threshold
    case 2: return r.uni(1e-9, 1e-6);
    case 3: return r.uni(1e-4, 1e-2);
    default: return r.uni(0.05, 3.);
  }
}

template <int FLUX, int WS>
static int check(const char* name, uint64_t seed, int n) {
  Rng r{std::mt19937_64(seed)};
  double worst = 0., worst_film = 0.;
  int bad = 0, nonfinite = 0, zero = 0;
  for (int it = 0; it < n; ++it) {
    const double be = r.uni(-1., 1.);
    const double sp = r.pick(3) == 0 ? 0.02 : (r.pick(2) ? 0.7 : 4.);
    Slot L{be + depth(r), r.nrm(sp), r.nrm(sp)}, R{be + depth(r), r.nrm(sp), r.nrm(sp)};
    const int tie = r.pick(8);
    if (tie == 0) R = L;                                 // identical sides
    else if (tie == 1) R = Slot{L.w, -L.u, -L.v};        // mirrored
    else if (tie == 2) R = Slot{L.w, r.nrm(sp), r.nrm(sp)};  // equal depths
    const double th = r.uni(0., 6.283185307179586);
    const double tx = std::cos(th), ty = std::sin(th), nx = ty, ny = -tx;
    const double Fb[3] = {r.nrm(1.), r.nrm(1.), r.nrm(1.)};
    using D = Dual<6>;
    const SlotT<D> Ld{dvar<6>(L.w, 0), dvar<6>(L.u, 1), dvar<6>(L.v, 2)};
    const SlotT<D> Rd{dvar<6>(R.w, 3), dvar<6>(R.u, 4), dvar<6>(R.v, 5)};
    D Fd[3];
    double lw = 1., g[6];
    if (FLUX == 0) {
      flux_hll<WS>(nx, ny, dconst<6>(be), Ld, Rd, 1., 0., lw, Fd);
      flux_hll_vjp<WS>(nx, ny, be, L, R, Fb, g);
    } else {
      flux_hllc<WS>(tx, ty, dconst<6>(be), Ld, Rd, 1., 0., lw, Fd);
      flux_hllc_vjp<WS>(tx, ty, be, L, R, Fb, g);
    }
    bool fin = true, allzero = true;
    double scale = 0., err = 0.;
    for (int k = 0; k < 6; ++k) {
      double gd = 0., s = 0.;
      for (int q = 0; q < 3; ++q) {
        gd += Fb[q] * Fd[q].d[k];
        s += std::fabs(Fb[q] * Fd[q].d[k]);
      }
      if (!std::isfinite(gd)) fin = false;
      if (gd != 0.) allzero = false;
      scale = std::fmax(scale, s);
      err = std::fmax(err, std::fabs(g[k] - gd));
    }
    if (!fin) {  // both sides must then be non-finite somewhere
      ++nonfinite;
      bool gfin = true;
      for (int k = 0; k < 6; ++k) gfin = gfin && std::isfinite(g[k]);
      if (gfin) ++bad;
      continue;
    }
    if (allzero) ++zero;
    const double rel = err / (scale > 0. ? scale : 1.);
    // a film (0 < h < 1e-3) makes intermediate terms (1 / 2 sqrt(h), 1 / h)
    // far larger than the Jacobian entries: forward and reverse mode then
    // round differently, by more than 1e-12 of the result but far less than
    // a wrong term would show
    const double hl = L.w - be, hr = R.w - be;
    const bool film = (hl > 0. && hl < 1e-3) || (hr > 0. && hr < 1e-3);
    (film ? worst_film : worst) = std::fmax(film ? worst_film : worst, rel);
    if (!(rel <= (film ? 1e-6 : 1e-12))) {
      if (bad < 5)
        std::printf("  mismatch %s: L=(%.17g,%.17g,%.17g) R=(%.17g,%.17g,%.17g) be=%.17g rel=%g\n", name, L.w,
                    L.u, L.v, R.w, R.u, R.v, be, rel);
      ++bad;
    }
  }
  std::printf("%s cases=%d max_rel=%.3g (with a film: %.3g) zero_jacobian=%d nonfinite=%d bad=%d\n", name, n,
              worst, worst_film, zero, nonfinite, bad);
  return bad;
}

int main(int argc, char** argv) {
  const int n = argc > 1 ? std::atoi(argv[1]) : 200000;
  int bad = 0;
  bad += check<0, 0>("HLL<Rusanov>", 11, n);
  bad += check<0, 1>("HLL<Davis>", 12, n);
  bad += check<0, 2>("HLL<Einfeldt>", 13, n);
  bad += check<1, 0>("HLLC<Rusanov>", 14, n);
  bad += check<1, 1>("HLLC<Davis>", 15, n);
  bad += check<1, 2>("HLLC<Einfeldt>", 16, n);
  return bad ? 1 : 0;
}
