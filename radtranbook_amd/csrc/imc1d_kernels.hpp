// imc1d_kernels.hpp -- per-segment physics of the gray 1-D IMC tracker
// (slab and sphere).
//
//  * double: restates IMC/IMC1D.py:75-130 (move_particle) and 336-436
//    (move_particle_spherical) operation for operation: 1e-10 face
//    tolerance, 1e-10 opacity thresholds, opacity cap sigma_s dx > 1e4,
//    non-positive distances reported as failures (the reference asserts).
//  * float: FP32 geometry with FP64 weights; argmin event choice (face,
//    then scatter, then census), moves clamped into the cell,
//    cancellation-free sphere chords and radius updates.  The position is an
//    FP32 offset from the cell's inner edge, carried by the tracker between
//    segments (cell tables (r0, r1 - r0) in FP32), so that moves of ~1e-4 of
//    a cell (the capped scatter rate) keep their precision, which against an
//    absolute FP32 coordinate they would not.
// Outputs are raw energies: dep = w (1 - e^{-sigma_a s}), path = w s; the
// caller divides by the cell volume (dx for slab) once per cell.
#pragma once
#include "imc2d_kernels.hpp"
#include "imc_device.hpp"

namespace imc {

// x: the new position (FP64) or the new offset from the cell's inner edge
// (FP32; a face for a crossing, in the coordinates of the cell moved in).
template <typename R>
struct Move1D {
  double w, dep, path;
  R x, mu, dcen;
  int newloc;  // -1, 0, +1
  int code;    // 1 boundary, 2 scatter, 3 census
  bool bad;    // non-positive distance (reference AssertionError)
};

// ---------------------------------------------------------------------------
// FP64: reference restatement.  Slots of event e: 4e scatter distance,
// 4e+1 capped redraw, 4e+2 new direction cosine.
// ---------------------------------------------------------------------------
template <int SPH>
__device__ __forceinline__ void move_1d(Ctr ctr, Key key, uint32_t ev, double w, double mu, double x, double r0,
                                        double r1, double sa, double ss, double dcen, Move1D<double>& o) {
  double dxc = r1 - r0;
  double db;
  int hit_outer = 1;
  if (!SPH) {
    db = (mu > 0) ? (r1 - x) / mu : (r0 - x) / mu;
  } else {
    double b2 = x * x * (1.0 - mu * mu);
    double dout = -x * mu + sqrt(r1 * r1 - b2);
    double din = 1e30;
    double discin = r0 * r0 - b2;
    if (mu < 0.0 && discin > 0.0) din = -x * mu - sqrt(discin);
    db = dout < din ? dout : din;
    hit_outer = (dout <= din);
  }
  double dsc = 1e30;
  Pair u0{0.0, 0.0};
  bool drew = false;
  if (ss > 1e-10) {
    u0 = block_d(ctr, key, 2u * ev);
    drew = true;
    dsc = -log(1.0 - u0.u0) / ss;
  }
  if (ss * dxc > 10000) {
    if (!drew) u0 = block_d(ctr, key, 2u * ev);
    dsc = -log(1.0 - u0.u1) / (10000.0 / dxc);
  }
  double s = db;
  if (dsc < s) s = dsc;
  if (dcen < s) s = dcen;
  o.bad = !(s > 0.0);
  o.newloc = 0;
  o.mu = mu;
  o.x = x;
  if (o.bad) {
    o.w = w; o.dep = 0.0; o.path = 0.0; o.dcen = 0.0; o.code = 3;
    return;
  }
  double rn, mn = mu;
  int code;
  if (!SPH) {
    rn = x + mu * s;
    if (fabs(db - s) < 1e-10) {
      if (mu > 0) { rn = r1; o.newloc = 1; } else { rn = r0; o.newloc = -1; }
      code = 1;
    } else if (s == dsc) {
      mn = 2.0 * block_d(ctr, key, 2u * ev + 1u).u0 - 1.0;
      code = 2;
    } else {
      code = 3;
    }
  } else {
    double rn2 = x * x + 2.0 * x * s * mu + s * s;
    rn = rn2 > 0.0 ? sqrt(rn2) : 0.0;
    mn = rn > 1e-15 ? (x * mu + s) / rn : 1.0;
    mn = fmax(-1.0, fmin(1.0, mn));
    if (fabs(db - s) < 1e-10) {
      if (hit_outer) {
        rn = r1;
        mn = r1 > 1e-15 ? (x * mu + s) / r1 : 1.0;
        o.newloc = 1;
      } else {
        rn = r0;
        mn = r0 > 1e-15 ? (x * mu + s) / r0 : -1.0;
        o.newloc = -1;
      }
      mn = fmax(-1.0, fmin(1.0, mn));
      code = 1;
    } else if (s == dsc) {
      mn = 2.0 * block_d(ctr, key, 2u * ev + 1u).u0 - 1.0;
      code = 2;
    } else {
      code = 3;
    }
  }
  o.code = code;
  o.dcen = dcen - s;
  double wf = exp(-sa * s);
  o.dep = w * (1.0 - wf);
  o.path = w * s;
  o.w = w * wf;
  o.mu = mn;
  o.x = rn;
}

// ---------------------------------------------------------------------------
// FP32 geometry.  One Philox block per event: u0 distance, u1 capped redraw,
// u2 new direction cosine.
// ---------------------------------------------------------------------------
template <int SPH>
__device__ __forceinline__ void move_1d(Ctr ctr, Key key, uint32_t ev, double w, float mu, float lx, float r0,
                                        float dxc, float sa, float ss, float dcen, Move1D<float>& o) {
  const float lo = lx, hi = dxc - lx;  // gaps to the cell's faces
  const float x = r0 + lx, r1 = r0 + dxc;
  float db;
  int hit_outer = 1;
  if (!SPH) {
    db = (mu != 0.0f) ? fmaxf(fdiv_fast(mu > 0.0f ? hi : -lo, mu), 0.0f) : 1e30f;
  } else {
    // r1^2 - x^2 (1 - mu^2) = hi (r1 + x) + (x mu)^2 and
    // r0^2 - x^2 (1 - mu^2) = (x mu)^2 - lo (x + r0): no cancellation next to a shell
    const float xm = x * mu;
    float rootout = __builtin_amdgcn_sqrtf(fmaxf(hi * (r1 + x) + xm * xm, 0.0f));
    float dout = (mu > 0.0f) ? fdiv_fast(fmaxf(hi * (r1 + x), 0.0f), rootout + xm) : rootout - xm;
    if (!(dout >= 0.0f)) dout = 0.0f;
    float din = 1e30f;
    float discin = xm * xm - lo * (x + r0);
    if (mu < 0.0f && r0 > 0.0f && discin > 0.0f) {
      float rootin = __builtin_amdgcn_sqrtf(discin);
      din = fdiv_fast(fmaxf(lo * (x + r0), 0.0f), rootin - xm);
    }
    db = fminf(dout, din);
    hit_outer = (dout <= din);
  }
  QuadF u = (ss > 1e-10f || ss * dxc > 10000.0f) ? block4_f(ctr, key, ev) : QuadF{0.0f, 0.0f, 0.0f, 0.0f};
  float dsc = 1e30f;
  if (ss > 1e-10f) dsc = fdiv_fast(neg_ln_fast(1.0f - u.u0), ss);
  if (ss * dxc > 10000.0f) dsc = neg_ln_fast(1.0f - u.u1) * (dxc * 1.0e-4f);
  o.bad = false;
  o.newloc = 0;
  float s, mn = mu, rn;
  if (db <= dsc && db <= dcen) {
    s = db;
    o.code = 1;
    if (SPH) {
      float a = x + mu * s;
      float rr = __builtin_amdgcn_sqrtf(a * a + s * s * fmaxf((1.0f - mu) * (1.0f + mu), 0.0f));
      mn = rr > 0.0f ? fminf(1.0f, fmaxf(-1.0f, fdiv_fast(x * mu + s, rr))) : mu;
      if (hit_outer) { rn = dxc; o.newloc = 1; } else { rn = 0.0f; o.newloc = -1; }
    } else {
      if (mu > 0.0f) { rn = dxc; o.newloc = 1; } else { rn = 0.0f; o.newloc = -1; }
    }
  } else {
    s = fminf(dsc, dcen);
    if (!SPH) {
      rn = lo + mu * s;
    } else {
      // r' - r = s (2 r mu + s) / (r' + r), added to the offset
      float a = x + mu * s;
      float rr = __builtin_amdgcn_sqrtf(a * a + s * s * fmaxf((1.0f - mu) * (1.0f + mu), 0.0f));
      mn = rr > 0.0f ? fminf(1.0f, fmaxf(-1.0f, fdiv_fast(x * mu + s, rr))) : 1.0f;
      float sum = rr + x;
      rn = lo + (sum > 0.0f ? fdiv_fast(s * (2.0f * x * mu + s), sum) : 0.0f);
    }
    rn = fminf(fmaxf(rn, 0.0f), dxc);
    if (dsc <= dcen) {
      o.code = 2;
      mn = 2.0f * u.u2 - 1.0f;
    } else {
      o.code = 3;
    }
  }
  o.dcen = dcen - s;
  float xs = sa * s;
  double dw, wn;
  if (xs < 0.5f) {
    dw = w * (double)one_minus_exp_small(xs);
    wn = w - dw;
  } else {
    wn = w * (double)__expf(-xs);
    dw = w - wn;
  }
  o.dep = dw;
  o.path = w * (double)s;
  o.w = wn;
  o.mu = mn;
  o.x = rn;
}

}  // namespace imc
