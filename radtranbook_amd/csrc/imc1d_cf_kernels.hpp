// imc1d_cf_kernels.hpp -- per-segment physics of the Carter-Forest gray 1-D
// tracker, slab (IMC/IMC1D_CarterForest.py:64-207) and sphere (615-720), FP64.
//
// Analog absorption at the true sigma_a; an absorbed particle is re-emitted
// isotropically after a delay -ln(xi)/(c sigma_a beta) when that ends before
// census, otherwise it is captured and its weight deposited.  Every segment
// adds w s to the path tally of its cell.  Faces are snapped to exactly, with
// the reference's 1e-10 tie tolerances.  The operations are the reference's,
// in its order, with contraction off.
//
// Draw slots for event e (0-based) of a history's transport stream: 6e
// absorption distance, 6e+1 capped redraw, 6e+2 re-emission delay, 6e+3
// fast-path capture test, 6e+4 new direction.
#pragma once
#include "imc_device.hpp"

namespace imc {

enum : int { kCf1Boundary = 1, kCf1Absorb = 2, kCf1Census = 3, kCf1Capture = 4 };

struct Cf1Out {
  double w, mu, x, t, dep_w, path, dcen;
  int newloc, code;
};

// Emission time from uniform xi (IMC/IMC1D_CarterForest.py:531-539, 954-959).
__device__ __forceinline__ double cf1_time(double rate, double dt, double xi) {
#pragma clang fp contract(off)
  if (rate > 1e-10) return -log(1.0 - xi * (1.0 - exp(-rate * dt))) / rate;
  return xi * dt;
}

__device__ __forceinline__ double cf1_exp_dist(Ctr ctr, Key key, uint32_t slot, double rate) {
#pragma clang fp contract(off)
  return -log(1.0 - slot_d(ctr, key, slot)) / rate;
}

__device__ __forceinline__ double cf1_iso(Ctr ctr, Key key, uint32_t ev) {
#pragma clang fp contract(off)
  return -1.0 + 2.0 * slot_d(ctr, key, 6u * ev + 4u);
}

// Absorption event (IMC/IMC1D_CarterForest.py:145-196, 681-710).  On entry
// o.dcen is the distance to census before the segment of length s and o.mu
// the direction after it.
__device__ __forceinline__ void cf1_absorb(Ctr ctr, Key key, uint32_t ev, double sa, double beta, double s,
                                           double fastpath, Cf1Out& o) {
#pragma clang fp contract(off)
  o.code = kCf1Absorb;
  if (beta > 1e-15 && sa > 1e-15) {
    const double rate = kC * sa * beta;
    const double trem = (o.dcen - s) / kC;
    bool capture;
    if (fastpath > 0.0 && trem > 1e-14 && rate * trem > fastpath) {
      const double pcap = exp(-rate * trem);
      capture = slot_d(ctr, key, 6u * ev + 3u) < pcap;
      if (!capture) {
        o.mu = cf1_iso(ctr, key, ev);
        o.dcen = o.dcen - s;
      }
    } else {
      double xi = slot_d(ctr, key, 6u * ev + 2u);
      if (xi < 1e-15) xi = 1e-15;
      const double t_abs = o.t + s / kC;
      const double delay = -log(xi) / rate;
      capture = !(trem > 1e-14 && delay < trem);
      if (!capture) {
        o.mu = cf1_iso(ctr, key, ev);
        o.t = t_abs + delay;
        o.dcen = (trem - delay) * kC;
      }
    }
    if (capture) {
      o.dep_w = o.w;
      o.w = 0.0;
      o.code = kCf1Capture;
    }
  } else {
    o.mu = cf1_iso(ctr, key, ev);
    o.dcen = o.dcen - s;
  }
}

// Slab segment (IMC/IMC1D_CarterForest.py:64-207).  A direction cosine of
// exactly zero never reaches a face (the reference divides by it).
__device__ __forceinline__ void cf1_move_slab(Ctr ctr, Key key, uint32_t ev, double w, double mu, double x, double t,
                                              double xl, double xr, double sa, double beta, double dcen,
                                              double fastpath, Cf1Out& o) {
#pragma clang fp contract(off)
  o.w = w; o.mu = mu; o.x = x; o.t = t; o.dep_w = 0.0; o.path = 0.0; o.newloc = 0; o.code = kCf1Census;
  if (dcen < 1e-14) {
    o.dcen = 0.0;
    return;
  }
  const double dx = xr - xl;
  double sb = 1e30;
  if (mu > 0.0)
    sb = fmax(0.0, (xr - x) / mu);
  else if (mu < 0.0)
    sb = fmax(0.0, (xl - x) / mu);
  double sabs = sa > 1e-10 ? cf1_exp_dist(ctr, key, 6u * ev, sa) : 1e30;
  if (sa * dx > 10000.0) sabs = cf1_exp_dist(ctr, key, 6u * ev + 1u, 10000.0 / dx);
  double s = sb;
  if (sabs < s) s = sabs;
  if (dcen < s) s = dcen;
  if (s <= 0.0) {
    o.dcen = 0.0;
    return;
  }
  o.x = x + mu * s;
  o.path = w * s;
  if (fabs(sb - s) < 1e-10) {
    o.code = kCf1Boundary;
    if (mu > 0.0) {
      o.x = xr;
      o.newloc = 1;
    } else {
      o.x = xl;
      o.newloc = -1;
    }
    o.dcen = dcen - s;
  } else if (fabs(sabs - s) < 1e-10) {
    o.dcen = dcen;
    cf1_absorb(ctr, key, ev, sa, beta, s, fastpath, o);
  } else {
    o.dcen = 0.0;
  }
}

// Spherical-shell segment (IMC/IMC1D_CarterForest.py:615-720).
__device__ __forceinline__ void cf1_move_sph(Ctr ctr, Key key, uint32_t ev, double w, double mu, double r, double t,
                                             double ri, double ro, double sa, double beta, double dcen,
                                             double fastpath, Cf1Out& o) {
#pragma clang fp contract(off)
  o.w = w; o.mu = mu; o.x = r; o.t = t; o.dep_w = 0.0; o.path = 0.0; o.newloc = 0; o.code = kCf1Census;
  if (dcen < 1e-14) {
    o.dcen = 0.0;
    return;
  }
  const double b2 = r * r * (1.0 - mu * mu);
  const double dout = -r * mu + sqrt(fmax(ro * ro - b2, 0.0));
  const double disc_in = ri * ri - b2;
  const double din = (mu < 0.0 && disc_in > 0.0) ? -r * mu - sqrt(disc_in) : 1e30;
  const double sb = din < dout ? din : dout;
  const bool hit_outer = dout <= din;
  const double dr = ro - ri;
  double sabs = sa > 1e-10 ? cf1_exp_dist(ctr, key, 6u * ev, sa) : 1e30;
  if (sa * dr > 10000.0) sabs = cf1_exp_dist(ctr, key, 6u * ev + 1u, 10000.0 / dr);
  double s = sb;
  if (sabs < s) s = sabs;
  if (dcen < s) s = dcen;
  if (s <= 0.0) {
    o.dcen = 0.0;
    return;
  }
  const double rn2 = r * r + 2.0 * r * s * mu + s * s;
  double rn = rn2 > 0.0 ? sqrt(rn2) : 0.0;
  double mun = rn > 1e-15 ? (r * mu + s) / rn : 1.0;
  mun = fmax(-1.0, fmin(1.0, mun));
  o.path = w * s;
  if (fabs(sb - s) < 1e-10) {
    o.code = kCf1Boundary;
    if (hit_outer) {
      rn = ro;
      mun = ro > 1e-15 ? (r * mu + s) / ro : 1.0;
      o.newloc = 1;
    } else {
      rn = ri;
      mun = ri > 1e-15 ? (r * mu + s) / ri : -1.0;
      o.newloc = -1;
    }
    mun = fmax(-1.0, fmin(1.0, mun));
    o.x = rn;
    o.mu = mun;
    o.dcen = dcen - s;
  } else if (fabs(sabs - s) < 1e-10) {
    o.x = rn;
    o.mu = mun;
    o.dcen = dcen;
    cf1_absorb(ctr, key, ev, sa, beta, s, fastpath, o);
  } else {
    o.x = rn;
    o.mu = mun;
    o.dcen = 0.0;
  }
}

}  // namespace imc
