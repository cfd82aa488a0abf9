// imc2d_kernels.hpp -- per-segment physics of the gray 2-D IMC tracker.
//
// One call = one segment ("event") of one history: distance to the cell faces,
// to the next effective scatter and to census; the particle moves by the
// smallest; continuous absorption deposits w(1 - e^{-sigma_a s}).
//
// Two instantiations:
//  * double: restates IMC/IMC2D.py:711-935 (_distance_to_radial_boundary_rz_jit,
//    _move_particle_xy_jit, _move_particle_rz_jit) operation for operation,
//    including the 1e-11 face-tie tolerance and the 1e-12 radial nudges.
//  * float: FP32 geometry for the same physics.  Face ties are decided by
//    argmin (boundary wins ties, then scatter, as in the reference), non-face
//    moves are clamped into the cell, the cylinder roots use the
//    cancellation-free forms, and the absorption split keeps
//    dep + w_new == w in FP64 so energy bookkeeping stays exact.  The
//    tracker carries the history's position as FP32 offsets from its cell's
//    lower edges: with absolute FP32 coordinates a move shorter than a few
//    ulps of the coordinate (the capped scatter rate 1e4 / h gives ~1e-4 h;
//    5e-7 cm against an ulp of 6e-8 cm at x = 0.5 on the 200x200 mesh) is
//    rounded away or biased, while an offset keeps ~1e-7 of the cell width.
#pragma once
#include <type_traits>

#include "imc_device.hpp"

namespace imc {

enum Crossing : int { kCrossNone = 0, kCrossIPlus = 1, kCrossIMinus = 2, kCrossJPlus = 3, kCrossJMinus = 4 };
enum Event : int { kEvtCensus = 0, kEvtBoundary = 1, kEvtScatter = 2 };

// dep_w: energy absorbed on the segment.  dep_i: w(1-e^{-sigma_a s})/sigma_a
// (FP64) or, in FP32, the path weight w*s used only where sigma_a <= 1e-12.
// s: the segment length (the adjoint replay needs it).
// p1, p2: the new position (FP64) or the new offsets from the cell's lower
// edges (FP32).
template <typename R>
struct MoveOut {
  double w, dep_w, dep_i;
  R p1, p2, d1, d2, dcen, s;
  int crossed, evt;
};

// ---------------------------------------------------------------------------
// FP64: reference restatement
// ---------------------------------------------------------------------------
__device__ __forceinline__ void absorb_f64(MoveOut<double>& o, double w, double sa, double s) {
  double wf = exp(-sa * s);
  double di = (sa > 1e-12) ? w * (1.0 - wf) / sa : w * s;
  o.dep_i = di;
  o.dep_w = di * sa;
  o.w = w * wf;
}

// s_scatter with the opacity cap of IMC/IMC2D.py:756-765 (slots 4ev, 4ev+1)
__device__ __forceinline__ double scatter_distance_f64(Ctr c, Key k, uint32_t ev, double ss, double cl) {
  double s = 1e30;
  if (ss > 1e-12) {
    Pair u = block_d(c, k, 2u * ev);
    s = -log(1.0 - u.u0) / ss;
    if (ss * cl > 10000.0) s = -log(1.0 - u.u1) / (10000.0 / cl);
  }
  return s;
}

__device__ __forceinline__ void move_xy(Ctr ctr, Key key, uint32_t ev, double w, double ux, double uy, double x,
                                        double y, double xl, double xr, double yl, double yr, double sa, double ss,
                                        double dcen, MoveOut<double>& o) {
  double sx = 1e30, sy = 1e30;
  if (ux > 0.0)
    sx = (xr - x) / ux;
  else if (ux < 0.0)
    sx = (xl - x) / ux;
  if (uy > 0.0)
    sy = (yr - y) / uy;
  else if (uy < 0.0)
    sy = (yl - y) / uy;
  double cl = fmin(xr - xl, yr - yl);
  double s_scat = scatter_distance_f64(ctr, key, ev, ss, cl);
  double sb = sx < sy ? sx : sy;
  double se = sb;
  if (s_scat < se) se = s_scat;
  if (dcen < se) se = dcen;
  o.d1 = ux; o.d2 = uy; o.p1 = x; o.p2 = y; o.crossed = kCrossNone; o.s = 0.0;
  if (se <= 0.0) {
    o.w = w; o.dep_w = 0.0; o.dep_i = 0.0; o.dcen = 0.0; o.evt = kEvtCensus;
    return;
  }
  o.s = se;
  double xn = x + ux * se, yn = y + uy * se;
  int evt = kEvtCensus;
  if (fabs(se - sb) < 1e-11) {
    evt = kEvtBoundary;
    if (sx <= sy) {
      if (ux > 0.0) { xn = xr; o.crossed = kCrossIPlus; } else { xn = xl; o.crossed = kCrossIMinus; }
    } else {
      if (uy > 0.0) { yn = yr; o.crossed = kCrossJPlus; } else { yn = yl; o.crossed = kCrossJMinus; }
    }
  } else if (se == s_scat) {
    evt = kEvtScatter;
    Pair u = block_d(ctr, key, 2u * ev + 1u);
    double uz = 2.0 * u.u0 - 1.0;
    double phi = 2.0 * kPi * u.u1;
    double rxy = sqrt(fmax(0.0, 1.0 - uz * uz));
    double sp, cp;
    sincos(phi, &sp, &cp);
    o.d1 = rxy * cp;
    o.d2 = rxy * sp;
  }
  o.p1 = xn; o.p2 = yn; o.evt = evt;
  o.dcen = dcen - se;
  absorb_f64(o, w, sa, se);
}

__device__ __forceinline__ double dist_radial_f64(double r, double mu, double eta, double R) {
  double ome2 = 1.0 - eta * eta;
  if (ome2 < 1e-15) return 1e30;
  double b2 = r * r * (1.0 - mu * mu);
  double disc = R * R - b2;
  if (disc <= 0.0) return 1e30;
  double root = sqrt(disc);
  double lp = -r * mu + root, lm = -r * mu - root;
  double den = sqrt(ome2);
  double smin = 1e30;
  if (lp > 1e-15) smin = fmin(smin, lp / den);
  if (lm > 1e-15) smin = fmin(smin, lm / den);
  return smin;
}

__device__ __forceinline__ void move_rz(Ctr ctr, Key key, uint32_t ev, double w, double mu, double eta, double r,
                                        double z, double rin, double rout, double zlo, double zhi, double sa,
                                        double ss, double dcen, MoveOut<double>& o) {
  double s_rin = rin > 0.0 ? dist_radial_f64(r, mu, eta, rin) : 1e30;
  double s_rout = dist_radial_f64(r, mu, eta, rout);
  double s_zlo, s_zhi;
  if (eta > 1e-15) {
    s_zlo = 1e30; s_zhi = (zhi - z) / eta;
  } else if (eta < -1e-15) {
    s_zlo = (zlo - z) / eta; s_zhi = 1e30;
  } else {
    s_zlo = 1e30; s_zhi = 1e30;
  }
  double cl = fmin(rout - rin, zhi - zlo);
  double s_scat = scatter_distance_f64(ctr, key, ev, ss, cl);
  double sb = s_rin;
  if (s_rout < sb) sb = s_rout;
  if (s_zlo < sb) sb = s_zlo;
  if (s_zhi < sb) sb = s_zhi;
  double se = sb;
  if (s_scat < se) se = s_scat;
  if (dcen < se) se = dcen;
  o.d1 = mu; o.d2 = eta; o.p1 = r; o.p2 = z; o.crossed = kCrossNone; o.s = 0.0;
  if (se <= 0.0) {
    o.w = w; o.dep_w = 0.0; o.dep_i = 0.0; o.dcen = 0.0; o.evt = kEvtCensus;
    return;
  }
  o.s = se;
  double ome2 = 1.0 - eta * eta;
  if (ome2 < 0.0) ome2 = 0.0;
  double l = se * sqrt(ome2);
  double rn2 = r * r + 2.0 * r * mu * l + l * l;
  if (rn2 < 0.0) rn2 = 0.0;
  double rn = sqrt(rn2);
  double zn = z + se * eta;
  double mun = mu;
  if (rn > 1e-15 && l > 0.0) mun = fmin(1.0, fmax(-1.0, (r * mu + l) / rn));
  double etan = eta;
  int evt = kEvtCensus;
  if (fabs(se - sb) < 1e-11) {
    evt = kEvtBoundary;
    if (sb == s_rin) {
      rn = rin > 0.0 ? rin * (1.0 - 1e-12) : 1e-12;
      o.crossed = kCrossIMinus;
    } else if (sb == s_rout) {
      rn = rout * (1.0 + 1e-12);
      o.crossed = kCrossIPlus;
    } else if (sb == s_zlo) {
      zn = zlo;
      o.crossed = kCrossJMinus;
    } else {
      zn = zhi;
      o.crossed = kCrossJPlus;
    }
  } else if (se == s_scat) {
    evt = kEvtScatter;
    Pair u = block_d(ctr, key, 2u * ev + 1u);
    etan = 2.0 * u.u0 - 1.0;
    mun = cos(2.0 * kPi * u.u1);
  }
  o.d1 = mun; o.d2 = etan; o.p1 = rn; o.p2 = zn; o.evt = evt;
  o.dcen = dcen - se;
  absorb_f64(o, w, sa, se);
}

// ---------------------------------------------------------------------------
// FP32 geometry, FP64 weights
// ---------------------------------------------------------------------------
// 1 - e^{-x} for 0 <= x < 0.5 by its Taylor series to x^8/8! (remainder
// below 1e-8 relative): no cancellation for small x.
__device__ __forceinline__ float one_minus_exp_small(float x) {
  float p = 1.0f / 40320.0f;
  p = 1.0f / 5040.0f - x * p;
  p = 1.0f / 720.0f - x * p;
  p = 1.0f / 120.0f - x * p;
  p = 1.0f / 24.0f - x * p;
  p = 1.0f / 6.0f - x * p;
  p = 0.5f - x * p;
  p = 1.0f - x * p;
  return x * p;
}

__device__ __forceinline__ void absorb_f32(MoveOut<float>& o, double w, float sa, float s) {
  float x = sa * s;
  double dw, wn;
  if (x < 0.5f) {
    float om = one_minus_exp_small(x);
    dw = w * (double)om;
    wn = w - dw;
  } else {
    float wf = __expf(-x);
    wn = w * (double)wf;
    dw = w - wn;
  }
  o.dep_w = dw;
  // path-length tally w*s is needed only where sigma_a <= 1e-12; elsewhere the
  // cell total is sum(dep_w) / sigma_a
  o.dep_i = (sa > 1e-12f) ? 0.0 : w * (double)s;
  o.w = wn;
}

// FP32 tracking draws four 24-bit uniforms from one Philox block per event:
// u0 scatter distance, u1 capped redraw, u2 and u3 the new direction.
__device__ __forceinline__ float scatter_distance_f32(const QuadF& u, float ss, float cl) {
  float s = 1e30f;
  if (ss > 1e-12f) {
    s = fdiv_fast(neg_ln_fast(1.0f - u.u0), ss);
    if (ss * cl > 10000.0f) s = neg_ln_fast(1.0f - u.u1) * (cl * 1.0e-4f);
  }
  return s;
}

// Cell tables of the FP32 trackers hold (lower edge, width) per index, both
// FP32, the width the FP32 difference of the FP32-rounded edges (so that
// lower edge + width gives the next lower edge).  A history's position is an
// FP32 offset from its cell's lower edges, carried from segment to segment;
// kAtUpperFace marks the upper face of the cell (a history that entered
// through it; the width is known once the cell's entry is read).
constexpr float kAtUpperFace = -1.0f;

__device__ __forceinline__ float resolve_offset(float l, float width) { return l < 0.0f ? width : l; }

// CAP: apply the gray opacity cap (sigma_s * min cell width > 1e4); the
// multigroup tracker has none.  lx, ly: offsets in [0, wx] x [0, wy]; the
// new offsets (same cell; a face for a crossing) are returned in o.p1, o.p2.
template <bool CAP = true>
__device__ __forceinline__ void move_xy(Ctr ctr, Key key, uint32_t ev, double w, float ux, float uy, float lx,
                                        float ly, float wx, float wy, float sa, float ss, float dcen,
                                        MoveOut<float>& o) {
  // face distances by reciprocal-multiply (about 2 ulp; clamped at 0)
  float sx = (ux != 0.0f) ? fmaxf(fdiv_fast((ux > 0.0f ? wx : 0.0f) - lx, ux), 0.0f) : 1e30f;
  float sy = (uy != 0.0f) ? fmaxf(fdiv_fast((uy > 0.0f ? wy : 0.0f) - ly, uy), 0.0f) : 1e30f;
  float cl = CAP ? fminf(wx, wy) : 0.0f;
  QuadF u = block4_f(ctr, key, ev);  // independent of the material gathers: overlaps their latency
  float s_scat = scatter_distance_f32(u, ss, cl);
  float sb = fminf(sx, sy);
  o.crossed = kCrossNone;
  o.d1 = ux; o.d2 = uy;
  float se;
  if (sb <= s_scat && sb <= dcen) {
    se = sb;
    o.evt = kEvtBoundary;
    if (sx <= sy) {
      if (ux > 0.0f) { o.p1 = wx; o.crossed = kCrossIPlus; } else { o.p1 = 0.0f; o.crossed = kCrossIMinus; }
      o.p2 = fminf(fmaxf(ly + uy * se, 0.0f), wy);
    } else {
      if (uy > 0.0f) { o.p2 = wy; o.crossed = kCrossJPlus; } else { o.p2 = 0.0f; o.crossed = kCrossJMinus; }
      o.p1 = fminf(fmaxf(lx + ux * se, 0.0f), wx);
    }
  } else {
    se = fminf(s_scat, dcen);
    o.p1 = fminf(fmaxf(lx + ux * se, 0.0f), wx);
    o.p2 = fminf(fmaxf(ly + uy * se, 0.0f), wy);
    if (s_scat <= dcen) {
      o.evt = kEvtScatter;
      float uz = 2.0f * u.u2 - 1.0f;
      float rxy = __builtin_amdgcn_sqrtf(fmaxf(0.0f, 1.0f - uz * uz));
      o.d1 = rxy * __builtin_amdgcn_cosf(u.u3);
      o.d2 = rxy * __builtin_amdgcn_sinf(u.u3);
    } else {
      o.evt = kEvtCensus;
    }
  }
  o.dcen = dcen - se;
  o.s = se;
  absorb_f32(o, w, sa, se);
}

// Distance to the cylinder of radius R along the transverse chord, divided by
// sqrt(1 - eta^2).  outer: the particle is inside (r <= R); inner: outside.
// gap = |R - r| from the cell offset.  The outer discriminant is written
// R^2 - r^2 (1 - mu^2) = gap (R + r) + (r mu)^2 and the chord
// (R - r)(R + r) / (root + r mu): both free of cancellation for r next to R.
__device__ __forceinline__ float dist_outer_f32(float r, float mu, float R, float gap, float den) {
  const float rm = r * mu;
  float disc = fmaxf(gap * (R + r) + rm * rm, 0.0f);
  float root = __builtin_amdgcn_sqrtf(disc);
  float l = (mu > 0.0f) ? fdiv_fast(fmaxf(gap * (R + r), 0.0f), root + rm) : root - rm;
  if (!(l >= 0.0f)) l = 0.0f;  // r == R, mu == 1 degenerate: 0/0
  return fdiv_fast(l, den);
}
__device__ __forceinline__ float dist_inner_f32(float r, float mu, float R, float gap, float den) {
  if (!(mu < 0.0f)) return 1e30f;
  const float rm = r * mu;
  float disc = rm * rm - gap * (r + R);
  if (disc <= 0.0f) return 1e30f;
  float root = __builtin_amdgcn_sqrtf(disc);
  float l = fdiv_fast(fmaxf(gap * (r + R), 0.0f), root - rm);
  return fdiv_fast(l, den);
}

// lr, lz: offsets from the cell's inner radius rin and lower z; wr, wz: its
// widths.  The new offsets (same cell; a face for a crossing) are returned in
// o.p1, o.p2.
template <bool CAP = true>
__device__ __forceinline__ void move_rz(Ctr ctr, Key key, uint32_t ev, double w, float mu, float eta, float lr,
                                        float lz, float rin, float wr, float wz, float sa, float ss, float dcen,
                                        MoveOut<float>& o) {
  const float r = rin + lr, rout = rin + wr;
  float ome2 = fmaxf((1.0f - eta) * (1.0f + eta), 0.0f);
  float den = __builtin_amdgcn_sqrtf(ome2);
  float s_rin = 1e30f, s_rout = 1e30f;
  if (den > 0.0f) {
    if (rin > 0.0f) s_rin = dist_inner_f32(r, mu, rin, lr, den);
    s_rout = dist_outer_f32(r, mu, rout, wr - lr, den);
  }
  float s_zlo = 1e30f, s_zhi = 1e30f;
  if (eta > 0.0f)
    s_zhi = fmaxf(fdiv_fast(wz - lz, eta), 0.0f);
  else if (eta < 0.0f)
    s_zlo = fmaxf(fdiv_fast(-lz, eta), 0.0f);
  float cl = CAP ? fminf(wr, wz) : 0.0f;
  QuadF u = block4_f(ctr, key, ev);  // independent of the material gathers: overlaps their latency
  float s_scat = scatter_distance_f32(u, ss, cl);
  float sr = fminf(s_rin, s_rout), sz = fminf(s_zlo, s_zhi);
  float sb = fminf(sr, sz);
  o.crossed = kCrossNone;
  float se;
  bool bnd = (sb <= s_scat && sb <= dcen);
  if (bnd)
    se = sb;
  else
    se = fminf(s_scat, dcen);
  // stream along the chord: r'^2 = (r + mu l)^2 + l^2 (1 - mu^2), and the new
  // radial cosine is mu' = (r mu + l) / r'; the radius moves by
  // r' - r = (2 r mu l + l^2) / (r' + r), added to the offset
  float l = se * den;
  float a = r + mu * l;
  float rn = __builtin_amdgcn_sqrtf(a * a + l * l * fmaxf((1.0f - mu) * (1.0f + mu), 0.0f));
  float sum = rn + r;
  float dr = sum > 0.0f ? fdiv_fast(l * (2.0f * r * mu + l), sum) : 0.0f;
  float lrn = fminf(fmaxf(lr + dr, 0.0f), wr);
  float lzn = fminf(fmaxf(lz + se * eta, 0.0f), wz);
  float mun = mu;
  if (rn > 0.0f && l > 0.0f) mun = fminf(1.0f, fmaxf(-1.0f, fdiv_fast(r * mu + l, rn)));
  float etan = eta;
  if (bnd) {
    o.evt = kEvtBoundary;
    if (sr <= sz) {
      if (s_rin <= s_rout) { lrn = 0.0f; o.crossed = kCrossIMinus; } else { lrn = wr; o.crossed = kCrossIPlus; }
    } else {
      if (s_zlo <= s_zhi) { lzn = 0.0f; o.crossed = kCrossJMinus; } else { lzn = wz; o.crossed = kCrossJPlus; }
    }
  } else {
    if (s_scat <= dcen) {
      o.evt = kEvtScatter;
      etan = 2.0f * u.u2 - 1.0f;
      mun = __builtin_amdgcn_cosf(u.u3);
    } else {
      o.evt = kEvtCensus;
    }
  }
  o.d1 = mun; o.d2 = etan;
  o.p1 = lrn;
  o.p2 = lzn;
  o.dcen = dcen - se;
  o.s = se;
  absorb_f32(o, w, sa, se);
}

// ---------------------------------------------------------------------------
// position state of the trackers
// ---------------------------------------------------------------------------
template <typename R>
constexpr bool kF32 = std::is_same<R, float>::value;

__device__ __forceinline__ double span_width(double2 e) { return e.y - e.x; }
__device__ __forceinline__ float span_width(float2 e) { return e.y; }

// One segment in cell tables ex, ey.  FP32: resolves the carried offsets
// (the pool's position on a history's first segment, the upper-face marker)
// against the cell first.
template <int GEOM, bool CAP = true>
__device__ __forceinline__ void segment_move(Ctr ctr, Key key, uint32_t ev, double w, double d1, double d2, double p1,
                                             double p2, bool&, double2 ex, double2 ey, double2 mat, double dcen,
                                             MoveOut<double>& o) {
  if (GEOM == 0)
    move_xy(ctr, key, ev, w, d1, d2, p1, p2, ex.x, ex.y, ey.x, ey.y, mat.x, mat.y, dcen, o);
  else
    move_rz(ctr, key, ev, w, d1, d2, p1, p2, ex.x, ex.y, ey.x, ey.y, mat.x, mat.y, dcen, o);
}
template <int GEOM, bool CAP = true>
__device__ __forceinline__ void segment_move(Ctr ctr, Key key, uint32_t ev, double w, float d1, float d2, float& p1,
                                             float& p2, bool& fresh, float2 ex, float2 ey, float2 mat, float dcen,
                                             MoveOut<float>& o) {
  p1 = fresh ? fminf(fmaxf(p1 - ex.x, 0.0f), ex.y) : resolve_offset(p1, ex.y);
  p2 = fresh ? fminf(fmaxf(p2 - ey.x, 0.0f), ey.y) : resolve_offset(p2, ey.y);
  fresh = false;
  if (GEOM == 0)
    move_xy<CAP>(ctr, key, ev, w, d1, d2, p1, p2, ex.y, ey.y, mat.x, mat.y, dcen, o);
  else
    move_rz<CAP>(ctr, key, ev, w, d1, d2, p1, p2, ex.x, ex.y, ey.y, mat.x, mat.y, dcen, o);
}

// FP32 offsets after a face crossing: the new cell's lower / upper face
__device__ __forceinline__ void enter_lower(double&) {}
__device__ __forceinline__ void enter_upper(double&) {}
__device__ __forceinline__ void enter_lower(float& l) { l = 0.0f; }
__device__ __forceinline__ void enter_upper(float& l) { l = kAtUpperFace; }

// Specular reflection at domain side s (0: i < 0, 1: i >= nx, 2: j < 0,
// 3: j >= ny) after the crossing moved the cell index out of the mesh: the
// history stays on that face of its last cell.
template <typename R>
__device__ __forceinline__ void reflect_at(int side, int nx, int ny, int& i, int& j, R& d1, R& d2, R& p1, R& p2) {
  if (side < 2) {
    d1 = -d1;
    i = side == 0 ? 0 : nx - 1;
    if constexpr (kF32<R>) p1 = side == 0 ? 0.0f : kAtUpperFace;
  } else {
    d2 = -d2;
    j = side == 2 ? 0 : ny - 1;
    if constexpr (kF32<R>) p2 = side == 2 ? 0.0f : kAtUpperFace;
  }
}

}  // namespace imc
