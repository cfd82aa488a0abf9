// imc2d_cf_kernels.hpp -- per-segment physics of the Carter-Forest gray 2-D
// tracker (IMC/IMC2D_CarterForest.py:103-343), FP64.
//
// Analog absorption at the true sigma_a; an absorbed particle is re-emitted
// isotropically after a delay -ln(xi)/(c sigma_a beta) if that ends before
// census, otherwise it is captured and its weight deposited.  Every segment
// adds w s to the scalar-intensity tally.  A face crossing moves the particle
// max(1e-14, 1e-12 h) past the face (h the cell width).  The operations are
// the reference's, in its order, with contraction off.
//
// Draw slots for event e (0-based) of a history's transport stream: 6e
// absorption distance, 6e+1 capped redraw, 6e+2 re-emission delay, 6e+3
// fast-path capture test, 6e+4 and 6e+5 the new direction.
#pragma once
#include "imc2d_kernels.hpp"
#include "imc_device.hpp"

namespace imc {

enum : int { kCfCensus = 0, kCfBoundary = 1, kCfAbsorb = 2, kCfCapture = 4 };

struct CfOut {
  double w, d1, d2, p1, p2, t, dep_w, dep_i, dcen;
  int crossed, evt;
};

// Emission time from uniform xi (IMC/IMC2D_CarterForest.py:51-56).
__device__ __forceinline__ double cf_time(double rate, double dt, double xi) {
#pragma clang fp contract(off)
  if (rate <= 1e-12) return xi * dt;
  return -log(1.0 - xi * (1.0 - exp(-rate * dt))) / rate;
}

__device__ __forceinline__ double cf_exp_dist(Ctr ctr, Key key, uint32_t slot, double rate) {
#pragma clang fp contract(off)
  double u = 1.0 - slot_d(ctr, key, slot);
  return -log(u > 1e-16 ? u : 1e-16) / rate;
}

// Absorption branch (IMC/IMC2D_CarterForest.py:183-213, 317-343); o holds
// the moved particle with dcen already reduced by the segment.
template <int GEOM>
__device__ __forceinline__ void cf_absorb(Ctr ctr, Key key, uint32_t ev, double w, double sa, double beta, double se,
                                          double tl, double eps_s, double fastpath, CfOut& o) {
#pragma clang fp contract(off)
  if (beta > 1e-15 && sa > 1e-15) {
    double rate = kC * sa * beta;
    double trem = o.dcen / kC;
    bool iso = false;
    if (fastpath > 0.0 && trem > 1e-14 && rate * trem > fastpath) {
      double pcap = exp(-rate * trem);
      if (slot_d(ctr, key, 6u * ev + 3u) < pcap) {
        o.w = 0.0; o.dep_w = w; o.dcen = 0.0; o.t = tl; o.evt = kCfCapture;
        return;
      }
      iso = true;
      o.t = tl;
      o.evt = kCfAbsorb;
    } else {
      double xi = slot_d(ctr, key, 6u * ev + 2u);
      if (xi < 1e-15) xi = 1e-15;
      double delay = -log(xi) / rate;
      if (delay < trem - 1e-15) {
        iso = true;
        o.t = tl + se / kC + delay;
        double dc = (trem - delay) * kC;
        o.dcen = dc > 0.0 ? dc : 0.0;
        if (o.dcen <= eps_s) {
          o.dcen = 0.0;
          o.evt = kCfCensus;
        } else {
          o.evt = kCfAbsorb;
        }
      }
    }
    if (iso) {
      o.w = w;
      o.dep_w = 0.0;
      Pair u = block_d(ctr, key, 3u * ev + 2u);  // slots 6e+4, 6e+5
      if (GEOM == 0) {
        double uz = 2.0 * u.u0 - 1.0;
        double phi = 2.0 * kPi * u.u1;
        double rxy = sqrt(fmax(0.0, 1.0 - uz * uz));
        double sp, cp;
        sincos(phi, &sp, &cp);
        o.d1 = rxy * cp;
        o.d2 = rxy * sp;
      } else {
        o.d2 = 2.0 * u.u0 - 1.0;
        o.d1 = cos(2.0 * kPi * u.u1);
      }
      return;
    }
  }
  o.w = 0.0; o.dep_w = w; o.dcen = 0.0; o.t = tl; o.evt = kCfCapture;
}

__device__ __forceinline__ void cf_move_xy(Ctr ctr, Key key, uint32_t ev, double w, double ux, double uy, double x,
                                           double y, double xl, double xr, double yl, double yr, double sa,
                                           double beta, double dcen, double tl, double fastpath, CfOut& o) {
#pragma clang fp contract(off)
  double sx = 1e30, sy = 1e30;
  if (ux > 0.0)
    sx = (xr - x) / ux;
  else if (ux < 0.0)
    sx = (xl - x) / ux;
  if (uy > 0.0)
    sy = (yr - y) / uy;
  else if (uy < 0.0)
    sy = (yl - y) / uy;
  double sb = sx < sy ? sx : sy;
  double s_abs = sa > 1e-12 ? cf_exp_dist(ctr, key, 6u * ev, sa) : 1e30;
  double cl = (xr - xl) < (yr - yl) ? (xr - xl) : (yr - yl);
  double eps_s = fmax(1e-14, 1e-12 * cl), eps_x = fmax(1e-14, 1e-12 * (xr - xl));
  double eps_y = fmax(1e-14, 1e-12 * (yr - yl));
  if (sa * cl > 10000.0) s_abs = cf_exp_dist(ctr, key, 6u * ev + 1u, 10000.0 / cl);
  double se = sb;
  if (s_abs < se) se = s_abs;
  if (dcen < se) se = dcen;
  if (fabs(se - sb) < 1e-11 && se < eps_s && dcen > eps_s) se = eps_s;
  o.w = w; o.d1 = ux; o.d2 = uy; o.p1 = x; o.p2 = y; o.t = tl; o.dep_w = 0.0; o.dep_i = 0.0;
  o.crossed = kCrossNone;
  if (se <= 0.0) {
    o.dcen = 0.0; o.evt = kCfCensus;
    return;
  }
  o.p1 = x + ux * se;
  o.p2 = y + uy * se;
  o.dep_i = w * se;
  o.dcen = dcen - se;
  if (fabs(se - sb) < 1e-11) {
    if (sx <= sy) {
      if (ux > 0.0) { o.p1 = xr + eps_x; o.crossed = kCrossIPlus; } else { o.p1 = xl - eps_x; o.crossed = kCrossIMinus; }
    } else {
      if (uy > 0.0) { o.p2 = yr + eps_y; o.crossed = kCrossJPlus; } else { o.p2 = yl - eps_y; o.crossed = kCrossJMinus; }
    }
    o.evt = kCfBoundary;
    return;
  }
  if (fabs(se - s_abs) < 1e-11) {
    cf_absorb<0>(ctr, key, ev, w, sa, beta, se, tl, eps_s, fastpath, o);
    return;
  }
  o.dcen = 0.0; o.evt = kCfCensus;
}

__device__ __forceinline__ double cf_dist_radial(double r, double mu, double eta, double R) {
#pragma clang fp contract(off)
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

__device__ __forceinline__ void cf_move_rz(Ctr ctr, Key key, uint32_t ev, double w, double mu, double eta, double r,
                                           double z, double rin, double rout, double zlo, double zhi, double sa,
                                           double beta, double dcen, double tl, double fastpath, CfOut& o) {
#pragma clang fp contract(off)
  double s_rin = rin > 0.0 ? cf_dist_radial(r, mu, eta, rin) : 1e30;
  double s_rout = cf_dist_radial(r, mu, eta, rout);
  double s_zlo, s_zhi;
  if (eta > 1e-15) {
    s_zlo = 1e30; s_zhi = (zhi - z) / eta;
  } else if (eta < -1e-15) {
    s_zlo = (zlo - z) / eta; s_zhi = 1e30;
  } else {
    s_zlo = 1e30; s_zhi = 1e30;
  }
  double sb = s_rin;
  if (s_rout < sb) sb = s_rout;
  if (s_zlo < sb) sb = s_zlo;
  if (s_zhi < sb) sb = s_zhi;
  double s_abs = sa > 1e-12 ? cf_exp_dist(ctr, key, 6u * ev, sa) : 1e30;
  double cl = (rout - rin) < (zhi - zlo) ? (rout - rin) : (zhi - zlo);
  double eps_s = fmax(1e-14, 1e-12 * cl), eps_r = fmax(1e-14, 1e-12 * (rout - rin));
  double eps_z = fmax(1e-14, 1e-12 * (zhi - zlo));
  if (sa * cl > 10000.0) s_abs = cf_exp_dist(ctr, key, 6u * ev + 1u, 10000.0 / cl);
  double se = sb;
  if (s_abs < se) se = s_abs;
  if (dcen < se) se = dcen;
  if (fabs(se - sb) < 1e-11 && se < eps_s && dcen > eps_s) se = eps_s;
  o.w = w; o.d1 = mu; o.d2 = eta; o.p1 = r; o.p2 = z; o.t = tl; o.dep_w = 0.0; o.dep_i = 0.0;
  o.crossed = kCrossNone;
  if (se <= 0.0) {
    o.dcen = 0.0; o.evt = kCfCensus;
    return;
  }
  double ome2 = 1.0 - eta * eta;
  if (ome2 < 0.0) ome2 = 0.0;
  double l = se * sqrt(ome2);
  double rn2 = r * r + 2.0 * r * mu * l + l * l;
  double rn = sqrt(rn2 > 0.0 ? rn2 : 0.0);
  double zn = z + se * eta;
  double mun = mu;
  if (rn > 1e-15 && l > 0.0) {
    mun = (r * mu + l) / rn;
    if (mun > 1.0) mun = 1.0; else if (mun < -1.0) mun = -1.0;
  }
  o.d1 = mun; o.p1 = rn; o.p2 = zn;
  o.dep_i = w * se;
  o.dcen = dcen - se;
  if (fabs(se - sb) < 1e-11) {
    if (sb == s_rin) {
      double rr = rin - eps_r;
      o.p1 = rin > 0.0 ? (rr > 1e-12 ? rr : 1e-12) : 1e-12;
      o.crossed = kCrossIMinus;
    } else if (sb == s_rout) {
      o.p1 = rout + eps_r; o.crossed = kCrossIPlus;
    } else if (sb == s_zlo) {
      o.p2 = zlo - eps_z; o.crossed = kCrossJMinus;
    } else {
      o.p2 = zhi + eps_z; o.crossed = kCrossJPlus;
    }
    o.evt = kCfBoundary;
    return;
  }
  if (fabs(se - s_abs) < 1e-11) {
    cf_absorb<1>(ctr, key, ev, w, sa, beta, se, tl, eps_s, fastpath, o);
    return;
  }
  o.dcen = 0.0; o.evt = kCfCensus;
}

}  // namespace imc
