// imcmg_kernels.hpp -- per-segment physics of the multigroup 2-D tracker and
// the Planck group integral.
//
// FP64 restates MG_IMC/MG_IMC2D.py:580-771 (_move_particle_xy,
// _distance_to_radial_boundary_rz, _move_particle_rz) operation for
// operation: no face snapping, events tested census -> scatter -> face by
// exact equality, the xy tiny-step rule s < 1e-14 -> max(s, sx, sy, d_census)
// and the rz floor s >= 1e-14, no opacity cap, dep = w - w e^{-sigma_a s}.
// FP32 tracking reuses the gray FP32 geometry with the cap switched off.
#pragma once
#include "imc2d_kernels.hpp"
#include "imc_device.hpp"

namespace imc {

struct MoveMG {
  double w, dep, path, d1, d2, p1, p2, s;
  int crossed, evt;
};

// new direction after an effective scatter: slots 4e+2, 4e+3 (one Philox block)
__device__ __forceinline__ void mg_isotropic(int geom, Pair u, double& d1, double& d2) {
  if (geom == 0) {
    double uz = 2.0 * u.u0 - 1.0;
    double phi = 2.0 * kPi * u.u1;
    double r = sqrt(fmax(0.0, 1.0 - uz * uz));
    double sp, cp;
    sincos(phi, &sp, &cp);
    d1 = r * cp;
    d2 = r * sp;
  } else {
    d2 = 2.0 * u.u0 - 1.0;
    d1 = cos(2.0 * kPi * u.u1);
  }
}

__device__ __forceinline__ void mg_absorb(MoveMG& o, double w, double sa, double s) {
  if (sa > 1e-12) {
    o.w = w * exp(-sa * s);
    o.dep = w - o.w;
    o.path = 0.0;
  } else {
    o.w = w;
    o.dep = 0.0;
    o.path = w * s;
  }
}

__device__ __forceinline__ void mg_move_xy(Ctr ctr, Key key, uint32_t ev, double w, double ux, double uy, double x,
                                           double y, double xl, double xr, double yl, double yr, double sa,
                                           double ss, double dcen, MoveMG& o) {
  double sx = 1e30, sy = 1e30;
  if (ux > 0.0)
    sx = (xr - x) / (ux + 1e-300);
  else if (ux < 0.0)
    sx = (xl - x) / (ux - 1e-300);
  if (uy > 0.0)
    sy = (yr - y) / (uy + 1e-300);
  else if (uy < 0.0)
    sy = (yl - y) / (uy - 1e-300);
  double s_scat = 1e30;
  if (ss > 1e-12) s_scat = -log(1.0 - block_d(ctr, key, 2u * ev).u0) / ss;
  double s = sx;
  if (sy < s) s = sy;
  if (s_scat < s) s = s_scat;
  if (dcen < s) s = dcen;
  if (s < 1e-14) {
    double m = s;
    if (sx > m) m = sx;
    if (sy > m) m = sy;
    if (dcen > m) m = dcen;
    s = m;
  }
  o.p1 = x + ux * s;
  o.p2 = y + uy * s;
  o.d1 = ux;
  o.d2 = uy;
  o.s = s;
  o.crossed = kCrossNone;
  if (s == dcen) {
    o.evt = kEvtCensus;
  } else if (s == s_scat) {
    o.evt = kEvtScatter;
    mg_isotropic(0, block_d(ctr, key, 2u * ev + 1u), o.d1, o.d2);
  } else if (s == sx) {
    o.evt = kEvtBoundary;
    o.crossed = ux > 0.0 ? kCrossIPlus : kCrossIMinus;
  } else if (s == sy) {
    o.evt = kEvtBoundary;
    o.crossed = uy > 0.0 ? kCrossJPlus : kCrossJMinus;
  } else {
    o.evt = kEvtCensus;
  }
  mg_absorb(o, w, sa, s);
}

__device__ __forceinline__ double mg_dist_radial(double r, double mu, double eta, double R) {
  double ome2 = 1.0 - eta * eta;
  if (ome2 < 1e-15) return 1e30;
  double b2 = r * r * (1.0 - mu * mu);
  double disc = R * R - b2;
  if (disc <= 0.0) return 1e30;
  double root = sqrt(disc);
  double lp = -r * mu + root, lm = -r * mu - root;
  double den = sqrt(ome2);
  double s = 1e30;
  if (lp > 1e-15 && lp / den < s) s = lp / den;
  if (lm > 1e-15 && lm / den < s) s = lm / den;
  return s;
}

__device__ __forceinline__ void mg_move_rz(Ctr ctr, Key key, uint32_t ev, double w, double mu, double eta, double r,
                                           double z, double rl, double rr, double zl, double zr, double sa,
                                           double ss, double dcen, MoveMG& o) {
  double srl = rl > 0.0 ? mg_dist_radial(r, mu, eta, rl) : 1e30;
  double srr = mg_dist_radial(r, mu, eta, rr);
  double sr = srl < srr ? srl : srr;
  double sz = 1e30;
  if (fabs(eta) > 1e-12) sz = eta > 0.0 ? (zr - z) / (eta + 1e-300) : (zl - z) / (eta - 1e-300);
  double s_scat = 1e30;
  if (ss > 1e-12) s_scat = -log(1.0 - block_d(ctr, key, 2u * ev).u0) / ss;
  double s = sr;
  if (sz < s) s = sz;
  if (s_scat < s) s = s_scat;
  if (dcen < s) s = dcen;
  if (s < 1e-14) s = 1e-14;
  double ome2 = fmax(0.0, 1.0 - eta * eta);
  double l = s * sqrt(ome2);
  double rn2 = r * r + 2.0 * r * mu * l + l * l;
  if (rn2 < 0.0) rn2 = 0.0;
  double rn = sqrt(rn2);
  double mun = mu;
  if (rn > 1e-15 && l > 0.0) {
    mun = (r * mu + l) / rn;
    if (mun > 1.0)
      mun = 1.0;
    else if (mun < -1.0)
      mun = -1.0;
  }
  o.p1 = rn;
  o.p2 = z + eta * s;
  o.d1 = mun;
  o.d2 = eta;
  o.s = s;
  o.crossed = kCrossNone;
  if (s == dcen) {
    o.evt = kEvtCensus;
  } else if (s == s_scat) {
    o.evt = kEvtScatter;
    mg_isotropic(1, block_d(ctr, key, 2u * ev + 1u), o.d1, o.d2);
  } else if (s == sr) {
    o.evt = kEvtBoundary;
    o.crossed = fabs(srl - sr) < 1e-14 ? kCrossIMinus : kCrossIPlus;
  } else if (s == sz) {
    o.evt = kEvtBoundary;
    o.crossed = eta > 0.0 ? kCrossJPlus : kCrossJMinus;
  } else {
    o.evt = kEvtCensus;
  }
  mg_absorb(o, w, sa, s);
}

// ---------------------------------------------------------------------------
// Planck group integral: (15/pi^4) int_{xlo}^{xhi} t^3/(e^t - 1) dt by the
// Bernoulli series below t = 1 and the exponential series above.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double planck_head(double x) {
#pragma clang fp contract(off)
  const double c[10] = {
      (1.0 / 6.0) / (2.0 * 5.0),
      (-1.0 / 30.0) / (24.0 * 7.0),
      (1.0 / 42.0) / (720.0 * 9.0),
      (-1.0 / 30.0) / (40320.0 * 11.0),
      (5.0 / 66.0) / (3628800.0 * 13.0),
      (-691.0 / 2730.0) / (479001600.0 * 15.0),
      (7.0 / 6.0) / (87178291200.0 * 17.0),
      (-3617.0 / 510.0) / (20922789888000.0 * 19.0),
      (43867.0 / 798.0) / (6402373705728000.0 * 21.0),
      (-174611.0 / 330.0) / (2432902008176640000.0 * 23.0),
  };
  double x2 = x * x, x3 = x2 * x;
  double s = 0.0, p = x3 * x2;
  for (int m = 0; m < 10; ++m) {
    s += c[m] * p;
    p *= x2;
  }
  return x3 / 3.0 - x3 * x / 8.0 + s;
}

__device__ __forceinline__ double planck_tail(double x) {
#pragma clang fp contract(off)
  int N = (int)(45.0 / x) + 2;
  if (N > 60) N = 60;
  double s = 0.0;
  double e1 = exp(-x), en = e1;
  for (int n = 1; n <= N; ++n) {
    double dn = (double)n;
    s += en * (x * x * x / dn + 3.0 * x * x / (dn * dn) + 6.0 * x / (dn * dn * dn) + 6.0 / (dn * dn * dn * dn));
    en *= e1;
  }
  return s;
}

__device__ __forceinline__ double planck_fraction(double xlo, double xhi) {
#pragma clang fp contract(off)
  const double full = kPi * kPi * kPi * kPi / 15.0;
  double v;
  if (xlo >= 1.0)
    v = planck_tail(xlo) - planck_tail(xhi);
  else if (xhi < 1.0)
    v = planck_head(xhi) - planck_head(xlo);
  else
    v = (full - planck_tail(xhi)) - planck_head(xlo);
  return v / full;
}

// B_g = a c T^4 / (4 pi) * fraction (T <= 0: 0)
__device__ __forceinline__ double planck_bg(double T, double elo, double ehi) {
#pragma clang fp contract(off)
  if (!(T > 0.0)) return 0.0;
  double t4 = (T * T) * (T * T);
  return kA * kC / (4.0 * kPi) * t4 * planck_fraction(elo / T, ehi / T);
}

// Group of a photon from the Planck spectrum at T (mixture of gammas,
// MG_IMC2D.py:202-254).  Slots slot0 .. slot0+4 of the particle's stream.
__device__ __forceinline__ int planck_group(Ctr ctr, Key key, uint32_t slot0, double T, int G,
                                            const double* __restrict__ edges, const double* __restrict__ pcdf) {
#pragma clang fp contract(off)
  if (!(T > 0.0)) return 0;
  double xi = slot_d(ctr, key, slot0);
  int lo = 0, hi = 199;
  if (!(xi <= pcdf[199])) {
    lo = 199;
  } else {
    while (lo < hi) {
      int mid = (lo + hi) >> 1;
      if (xi <= pcdf[mid])
        hi = mid;
      else
        lo = mid + 1;
    }
  }
  double ns = (double)(lo + 1);
  double r1 = -log(slot_d(ctr, key, slot0 + 1) + 1e-300);
  double r2 = -log(slot_d(ctr, key, slot0 + 2) + 1e-300);
  double r3 = -log(slot_d(ctr, key, slot0 + 3) + 1e-300);
  double r4 = -log(slot_d(ctr, key, slot0 + 4) + 1e-300);
  double x = (r1 + r2 + r3 + r4) / ns;
  double nu = T * x;
  if (nu < edges[0]) return 0;
  if (nu >= edges[G]) return G - 1;
  int a = 0, b = G - 1;
  while (a < b) {
    int mid = (a + b + 1) >> 1;
    if (edges[mid] <= nu)
      a = mid;
    else
      b = mid - 1;
  }
  return a;
}

}  // namespace imc
