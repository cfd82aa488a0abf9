// imcmg.hip -- multigroup 2-D (xy / rz) Fleck-Cummings IMC step on one MI355X.
//
// Particles carry a group index; opacities, tallies and the comb are per
// (group, cell) bin b = g * ncell + c with c = i * ny + j.  The Planck-mean
// Fleck factor, the group fractions b*_g and the emission energies are one
// kernel over cells; sources are drawn per particle by inverse CDF (the
// reference's multinomial np.random.choice); the persistent tracker and the
// comb follow the gray engine's structure.  Follows MG_IMC/MG_IMC2D.py.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/imc_hip.h"
#include "imc_device.hpp"
#include "imc_util.hpp"
#include "imc1d_kernels.hpp"
#include "imcmg_kernels.hpp"

namespace imc {

constexpr int kMaxGroups = 256;

template <typename R>
struct MgView {
  double* w;
  R* d1;
  R* d2;
  R* t;
  R* p1;
  R* p2;
  int32_t* ci;
  int32_t* cj;
  int32_t* g;
};

struct MgPool {
  void* a[9] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  int64_t cap = 0;
  int rsz = 8;
  size_t esz(int k) const { return k == 0 ? 8 : (k >= 6 ? 4 : (size_t)rsz); }
  template <typename R>
  MgView<R> view() const {
    return MgView<R>{(double*)a[0], (R*)a[1], (R*)a[2], (R*)a[3], (R*)a[4], (R*)a[5],
                     (int32_t*)a[6], (int32_t*)a[7], (int32_t*)a[8]};
  }
  int grow(int64_t n, int64_t keep, hipStream_t s) {
    if (n <= cap) return kOk;
    int64_t nc = std::max<int64_t>(n, cap + cap / 4);
    void* nw[9];
    for (int k = 0; k < 9; ++k) {
      IMC_HIP(hipMalloc(&nw[k], esz(k) * (size_t)nc));
      if (a[k] && keep > 0) IMC_HIP(hipMemcpyAsync(nw[k], a[k], esz(k) * (size_t)keep, hipMemcpyDeviceToDevice, s));
    }
    IMC_HIP(hipStreamSynchronize(s));
    for (int k = 0; k < 9; ++k) {
      if (a[k]) IMC_HIP(hipFree(a[k]));
      a[k] = nw[k];
    }
    cap = nc;
    return kOk;
  }
  void release() {
    for (int k = 0; k < 9; ++k) {
      if (a[k]) (void)hipFree(a[k]);
      a[k] = nullptr;
    }
    cap = 0;
  }
};

// ---------------------------------------------------------------------------
// per-cell kernels (FMA contraction off: energies, CDFs and sourcing match
// the oracle up to device-libm ulps)
// ---------------------------------------------------------------------------
// Planck-mean Fleck factor (MG_IMC2D.py:1693-1719), b*_g (1879-1885) or the
// caller's emission fractions, and emission energies sigma_a,g b*_g a c T^4
// dt V (1888-1894).  Eplanck: per-cell emission with the Planck b*_g, no
// floor (the particle-budget estimate of 1741-1755).
__global__ void k_mg_fleck(int64_t ncell, int G, const double* __restrict__ edges, const double* __restrict__ T,
                           const double* __restrict__ sig, const double* __restrict__ cv, double theta, double dt,
                           const double* __restrict__ frac, double T_emit_floor, const double* __restrict__ vol,
                           double* __restrict__ sa, double* __restrict__ ss, double2* __restrict__ sas64,
                           float2* __restrict__ sas32, double* __restrict__ bstar, double* __restrict__ E,
                           double* __restrict__ Ecell, double* __restrict__ Eplanck) {
#pragma clang fp contract(off)
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  double t = T[c];
  double bs = 0.0, num = 0.0;
  for (int g = 0; g < G; ++g) {
    double B = planck_bg(t, edges[g], edges[g + 1]);
    bs += B;
    num += sig[(int64_t)g * ncell + c] * B;
  }
  double sP = num / (bs + 1e-300);
  double beta = 4.0 * kA * (t * t * t) / cv[c];
  double f = 1.0 / (1.0 + theta * beta * sP * kC * dt);
  if (f < 0.0) f = 0.0;
  if (f > 1.0) f = 1.0;
  double t4 = (t * t) * (t * t);
  bool cold = T_emit_floor > 0.0 && t < T_emit_floor;
  double tot = 0.0, totp = 0.0;
  for (int g = 0; g < G; ++g) {
    int64_t q = (int64_t)g * ncell + c;
    double s = sig[q];
    double s_s = s * (1.0 - f), s_a = s * f;
    double bp = planck_bg(t, edges[g], edges[g + 1]) / (bs + 1e-300);
    double b = frac ? frac[g] : bp;
    sa[q] = s_a;
    ss[q] = s_s;
    sas64[q] = make_double2(s_a, s_s);
    sas32[q] = make_float2((float)s_a, (float)s_s);
    bstar[q] = b;
    double e = s_a * b * kA * kC * t4 * dt * vol[c];
    if (cold) e = 0.0;
    E[q] = e;
    tot += e;
    totp += s_a * bp * kA * kC * t4 * dt * vol[c];
  }
  Ecell[c] = tot;
  Eplanck[c] = totp;
}

// Canonical inclusive prefix: sequential within chunks of 1024, chunk offsets
// summed sequentially (one thread), then offset + local prefix.
// One workgroup per chunk, staged through LDS (coalesced loads and stores);
// one lane forms the running sum in order.
__global__ __launch_bounds__(256) void k_prefix_chunks(const double* __restrict__ x, int64_t n,
                                                       double* __restrict__ local, double* __restrict__ chunk_tot) {
#pragma clang fp contract(off)
  __shared__ double v[1024];
  const int64_t c0 = (int64_t)blockIdx.x * 1024;
  const int m = (int)((n - c0) < 1024 ? (n - c0) : 1024);
  for (int i = threadIdx.x; i < m; i += 256) v[i] = x[c0 + i];
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < m; ++i) {
      s += v[i];
      v[i] = s;
    }
    chunk_tot[blockIdx.x] = s;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < m; i += 256) local[c0 + i] = v[i];
}

__global__ void k_prefix_offsets(const double* __restrict__ chunk_tot, int64_t nch, double* __restrict__ off) {
  double o = 0.0;
  for (int64_t k = 0; k < nch; ++k) {
    off[k] = o;
    o += chunk_tot[k];
  }
}

__global__ void k_prefix_finish(const double* __restrict__ local, const double* __restrict__ off, int64_t n,
                                double* __restrict__ cdf) {
#pragma clang fp contract(off)
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) cdf[i] = off[i >> 10] + local[i];
}

// Running sum for long arrays in three levels, each left to right: the
// 1024-element chunks of k_prefix_chunks, chunk offsets within each group of
// 1024 chunks (one lane per group), then the group offsets (one lane);
// S[i] = (group offset + chunk offset) + local prefix.  The oracle's
// om_prefix3 restates the same order.
__global__ void k_prefix_groups(const double* __restrict__ chunk_tot, int64_t nch, double* __restrict__ choff,
                                double* __restrict__ group_tot) {
#pragma clang fp contract(off)
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t b0 = q * 1024;
  if (b0 >= nch) return;
  const int64_t b1 = b0 + 1024 < nch ? b0 + 1024 : nch;
  double o = 0.0;
  for (int64_t b = b0; b < b1; ++b) {
    choff[b] = o;
    o += chunk_tot[b];
  }
  group_tot[q] = o;
}

__global__ void k_prefix_finish3(const double* __restrict__ local, const double* __restrict__ choff,
                                 const double* __restrict__ group_tot, int64_t ngroup, int64_t n,
                                 double* __restrict__ S) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t b = i >> 10, q = b >> 10;
  double go = 0.0;  // the group offset, summed left to right (few groups)
  for (int64_t r = 0; r < q; ++r) go += group_tot[r];
  S[i] = (go + choff[b]) + local[i];
}

__device__ __forceinline__ int64_t mg_pick(const double* __restrict__ cdf, int64_t n, double u) {
  int64_t lo = 0, hi = n - 1;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (cdf[mid] > u)
      hi = mid;
    else
      lo = mid + 1;
  }
  return lo;
}

// Guide table for mg_pick over a running sum cdf[0..n-1] of total E =
// cdf[n-1]: guide[j] = mg_pick(cdf, n, (j / kGuide) * E), j = 0..kGuide.  For
// u = x * E with x in [0, 1), j = floor(x * kGuide) is exact and rounding is
// monotone, so mg_pick(cdf, n, u) lies in [guide[j], guide[j + 1]] and the
// search over that bracket with the same predicate returns the same index.
constexpr int kGuide = 65536;

__global__ void k_mg_guide(const double* __restrict__ cdf, int64_t n, int32_t* __restrict__ guide) {
#pragma clang fp contract(off)
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j > kGuide) return;
  guide[j] = (int32_t)mg_pick(cdf, n, ((double)j * (1.0 / kGuide)) * cdf[n - 1]);
}

__device__ __forceinline__ int64_t mg_pick_guided(const double* __restrict__ cdf, const int32_t* __restrict__ guide,
                                                  double x, double u) {
  int j = (int)(x * (double)kGuide);
  int64_t lo = guide[j], hi = guide[j + 1];
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (cdf[mid] > u)
      hi = mid;
    else
      lo = mid + 1;
  }
  return lo;
}

__device__ __forceinline__ void mg_pos(int geom, const double* __restrict__ e1, const double* __restrict__ e2, int i,
                                       int j, double u1, double u2, double& p1, double& p2) {
#pragma clang fp contract(off)
  if (geom == 0) {
    p1 = e1[i] + u1 * (e1[i + 1] - e1[i]);
  } else {
    double rl = e1[i], rr = e1[i + 1];
    p1 = sqrt(rl * rl + u1 * (rr * rr - rl * rl));
  }
  p2 = e2[j] + u2 * (e2[j + 1] - e2[j]);
}

template <typename R>
__device__ __forceinline__ void mg_store(MgView<R> P, int64_t o, double w, double d1, double d2, double t, double p1,
                                         double p2, int i, int j, int g) {
  P.w[o] = w;
  P.d1[o] = (R)d1;
  P.d2[o] = (R)d2;
  P.t[o] = (R)t;
  P.p1[o] = (R)p1;
  P.p2[o] = (R)p2;
  P.ci[o] = i;
  P.cj[o] = j;
  P.g[o] = g;
}

// Group energies cell-major, Ecg[c G + g] = E[g ncell + c]: the group draw
// of an emission particle reads its cell's G values from one or two cache
// lines instead of G lines ncell apart.
__global__ void k_mg_cell_major(const double* __restrict__ E, int64_t ncell, int G, double* __restrict__ Ecg) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ncell * G) return;
  const int64_t c = q / G;
  const int g = (int)(q - c * G);
  Ecg[q] = E[(int64_t)g * ncell + c];
}

// Cell and group of emission particle k (slot 0: the cell by the guided
// inverse CDF, then the group within the cell by its group energies, Ecg
// cell-major).
__device__ __forceinline__ void mg_emission_bin(Pair b0, int64_t ncell, int G, const double* __restrict__ Ecg,
                                                const double* __restrict__ cdf, const int32_t* __restrict__ guide,
                                                int64_t* cell, int* grp) {
#pragma clang fp contract(off)
  const double Etot = cdf[ncell - 1];
  int64_t c = mg_pick_guided(cdf, guide, b0.u0, b0.u0 * Etot);
  const double* e = Ecg + c * G;
  double tot = 0.0;
  for (int g = 0; g < G; ++g) tot += e[g];
  int gs = 0;
  if (tot > 0.0) {
    double xi = b0.u1 * tot, cum = 0.0;
    for (int g = 0; g < G; ++g) {
      cum += e[g];
      if (xi <= cum) {
        gs = g;
        break;
      }
    }
  }
  *cell = c;
  *grp = gs;
}

// ln x for x in (0, 1] in plain FP64 arithmetic (no library call, no FMA
// contraction), so that the oracle's identical restatement gives the same
// bits: x = m 2^e with m in [1/sqrt 2, sqrt 2), ln m = 2 atanh(s) with s =
// (m - 1) / (m + 1), |s| < 0.172, summed to s^23 (truncation below 1e-19).
__device__ __forceinline__ double mg_ln01(double x) {
#pragma clang fp contract(off)
  int e;
  double m = frexp(x, &e);
  if (m < 0.70710678118654752440) {
    m = m * 2.0;
    e -= 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double s2 = s * s;
  double p = 1.0 / 23.0;
  p = 1.0 / 21.0 + s2 * p;
  p = 1.0 / 19.0 + s2 * p;
  p = 1.0 / 17.0 + s2 * p;
  p = 1.0 / 15.0 + s2 * p;
  p = 1.0 / 13.0 + s2 * p;
  p = 1.0 / 11.0 + s2 * p;
  p = 1.0 / 9.0 + s2 * p;
  p = 1.0 / 7.0 + s2 * p;
  p = 1.0 / 5.0 + s2 * p;
  p = 1.0 / 3.0 + s2 * p;
  p = 1.0 + s2 * p;
  return (double)e * 0.69314718055994530942 + 2.0 * s * p;
}

// The N emission cells are the order statistics of N i.i.d. uniforms, taken
// as u_k = S_k / S_N from the running sums S of N + 1 exponential spacings
// x_k = -ln(1 - u0_k) (slot 0 of particle k's stream; k = N is the closing
// spacing): cells then come out in non-decreasing order, which keeps a
// wave's histories, and so its tally adds, in few cells, and the multiset of
// (cell, group) draws has the reference's distribution (N i.i.d. draws,
// MG_IMC2D.py:1896-1943).
__global__ void k_mg_emission_spacings(int64_t n1, Key key, uint32_t step, double* __restrict__ x) {
#pragma clang fp contract(off)
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n1) return;
  const Pair b0 = block_d(make_ctr(kTagEmission, (uint64_t)k, step), key, 0);
  x[k] = -mg_ln01(1.0 - b0.u0);
}

// Material emission (MG_IMC2D.py:1896-1943).  Slots: 0 cell (the sorted
// uniform u_k = S_k / S_N of k_mg_emission_spacings) and group, 1 position,
// 2 direction, 3 time.
template <typename R>
__global__ void k_mg_sample_emission(MgView<R> P, int64_t base, int64_t N, int geom, int nx, int ny, int G,
                                     const double* __restrict__ e1, const double* __restrict__ e2,
                                     const double* __restrict__ Ecg, const double* __restrict__ cdf,
                                     const int32_t* __restrict__ guide, double dt, Key key, uint32_t step,
                                     const double* __restrict__ S) {
#pragma clang fp contract(off)
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N) return;
  const int64_t ncell = (int64_t)nx * ny;
  Ctr ctr = make_ctr(kTagEmission, (uint64_t)k, step);
  Pair b0 = block_d(ctr, key, 0), b1 = block_d(ctr, key, 1), b2 = block_d(ctr, key, 2), b3 = block_d(ctr, key, 3);
  double u = S[k] / S[N];
  if (!(u < 1.0)) u = 0x1.fffffffffffffp-1;  // the largest double below 1
  b0.u0 = u;
  int64_t c;
  int gs;
  mg_emission_bin(b0, ncell, G, Ecg, cdf, guide, &c, &gs);
  const double Etot = cdf[ncell - 1];
  int i = (int)(c / ny), j = (int)(c % ny);
  double p1, p2, d1, d2;
  mg_pos(geom, e1, e2, i, j, b1.u0, b1.u1, p1, p2);
  mg_isotropic(geom, b2, d1, d2);
  mg_store(P, base + k, Etot / (double)N, d1, d2, dt * b3.u0, p1, p2, i, j, gs);
}

// Equilibrium census (MG_IMC2D.py:992-1095).  Slots: 0 cell, 1 pos1, 2 pos2,
// 3/4 direction, 5..9 Planck group at the cell's Tr.
template <typename R>
__global__ void k_mg_sample_ic(MgView<R> P, int64_t N, int geom, int nx, int ny, int G, const double* __restrict__ e1,
                               const double* __restrict__ e2, const double* __restrict__ edges,
                               const double* __restrict__ Tr, const double* __restrict__ cdf,
                               const double* __restrict__ pcdf, Key key) {
#pragma clang fp contract(off)
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N) return;
  const int64_t ncell = (int64_t)nx * ny;
  const double Etot = cdf[ncell - 1];
  Ctr ctr = make_ctr(kTagIC, (uint64_t)k, 0);
  Pair b0 = block_d(ctr, key, 0), b1 = block_d(ctr, key, 1), b2 = block_d(ctr, key, 2);
  int64_t c = mg_pick(cdf, ncell, b0.u0 * Etot);
  int i = (int)(c / ny), j = (int)(c % ny);
  double p1, p2, d1, d2;
  mg_pos(geom, e1, e2, i, j, b0.u1, b1.u0, p1, p2);
  mg_isotropic(geom, Pair{b1.u1, b2.u0}, d1, d2);
  int g = planck_group(ctr, key, 5, Tr[c], G, edges, pcdf);
  mg_store(P, k, Etot / (double)N, d1, d2, 0.0, p1, p2, i, j, g);
}

__device__ __forceinline__ int mg_locate(const double* __restrict__ e, int n, double x) {
  int lo = 0, hi = n + 1;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (e[mid] > x)
      hi = mid;
    else
      lo = mid + 1;
  }
  int idx = lo - 1;
  return idx < 0 ? 0 : (idx > n - 1 ? n - 1 : idx);
}

// Face source (MG_IMC2D.py:392-577).  Slots: 0 time, 1 mu_n, 2 phi / eta, 3
// along the face, 4 face cell (profile), 5..9 Planck group.  Returns the group.
template <typename R>
__device__ __forceinline__ int mg_face_particle(MgView<R> P, int64_t base, int64_t k, int64_t n, int geom, int side,
                                                double E_total, double dt, double Tb, int nx, int ny,
                                                const double* __restrict__ e1, const double* __restrict__ e2, int G,
                                                const double* __restrict__ edges, const double* __restrict__ pcdf,
                                                const double* __restrict__ prof, const double* __restrict__ Tface,
                                                Key key, uint32_t step) {
#pragma clang fp contract(off)
  const int nface = side < 2 ? ny : nx;
  const double* fe = side < 2 ? e2 : e1;
  double x0 = e1[0], x1 = e1[nx], y0 = e2[0], y1 = e2[ny];
  Ctr ctr = make_ctr(kTagBoundary0 + side, (uint64_t)k, step);
  Pair b0 = block_d(ctr, key, 0), b1 = block_d(ctr, key, 1);
  double a = side < 2 ? y0 : x0, b = side < 2 ? y1 : x1, Tk = Tb;
  if (prof) {
    double u = slot_d(ctr, key, 4) * prof[nface - 1];
    int m = first_above(prof, nface, u);
    while (m > 0 && prof[m] == prof[m - 1]) --m;
    a = fe[m];
    b = fe[m + 1];
    Tk = Tface[m];
  }
  double t = dt * b0.u0;
  double along = a + (b - a) * b1.u1;
  double mun = sqrt(b0.u1);
  double q1, q2, u1, u2;
  if (geom == 0) {
    double mut = sqrt(fmax(0.0, 1.0 - mun * mun));
    double ct = mut * cos(2.0 * kPi * b1.u0);
    if (side == 0) { u1 = mun; u2 = ct; q1 = x0; q2 = along; }
    else if (side == 1) { u1 = -mun; u2 = ct; q1 = x1; q2 = along; }
    else if (side == 2) { u2 = mun; u1 = ct; q2 = y0; q1 = along; }
    else { u2 = -mun; u1 = ct; q2 = y1; q1 = along; }
  } else if (side < 2) {
    u2 = 2.0 * b1.u0 - 1.0;
    u1 = side == 0 ? mun : -mun;
    q1 = side == 0 ? x0 : x1;
    q2 = along;
    if (prof) q1 = side == 0 ? x0 + 1e-12 : x1 - 1e-12;
  } else {
    u1 = cos(2.0 * kPi * b1.u0);
    u2 = side == 2 ? mun : -mun;
    q1 = sqrt(a * a + b1.u1 * (b * b - a * a));
    q2 = side == 2 ? y0 : y1;
    if (prof) q2 = side == 2 ? y0 + 1e-12 : y1 - 1e-12;
  }
  int g = planck_group(ctr, key, 5, Tk, G, edges, pcdf);
  mg_store(P, base + k, E_total / (double)n, u1, u2, t, q1, q2, mg_locate(e1, nx, q1), mg_locate(e2, ny, q2), g);
  return g;
}

// gcount (nullable): particles per group, counted per workgroup in LDS (G
// 32-bit counters of dynamic shared memory) and added once per group.
template <typename R>
__global__ void k_mg_sample_boundary(MgView<R> P, int64_t base, int64_t n, int geom, int side, double E_total,
                                     double dt, double Tb, int nx, int ny, const double* __restrict__ e1,
                                     const double* __restrict__ e2, int G, const double* __restrict__ edges,
                                     const double* __restrict__ pcdf, const double* __restrict__ prof,
                                     const double* __restrict__ Tface, Key key, uint32_t step,
                                     unsigned long long* __restrict__ gcount) {
  extern __shared__ unsigned int ghist[];
  if (gcount) {
    for (int q = threadIdx.x; q < G; q += blockDim.x) ghist[q] = 0u;
    __syncthreads();
  }
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) {
    int g = mg_face_particle<R>(P, base, k, n, geom, side, E_total, dt, Tb, nx, ny, e1, e2, G, edges, pcdf, prof,
                                Tface, key, step);
    if (gcount) atomicAdd(ghist + g, 1u);
  }
  if (gcount) {
    __syncthreads();
    for (int q = threadIdx.x; q < G; q += blockDim.x)
      if (ghist[q]) atomicAdd(gcount + q, (unsigned long long)ghist[q]);
  }
}

// Fixed source (MG_IMC2D.py:1098-1234).  cdfg: (G, ncell) canonical prefixes
// of the per-cell source energy; gcdf: cumulative group probabilities.
// Slots: 0 group, 1 cell, 2 pos1, 3 pos2, 4/5 direction, 6 time.
template <typename R>
__global__ void k_mg_sample_source(MgView<R> P, int64_t base, int64_t N, int geom, int nx, int ny, int G,
                                   const double* __restrict__ e1, const double* __restrict__ e2,
                                   const double* __restrict__ cdfg, const double* __restrict__ gcdf, double total,
                                   double dt, Key key, uint32_t step) {
#pragma clang fp contract(off)
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N) return;
  const int64_t ncell = (int64_t)nx * ny;
  Ctr ctr = make_ctr(kTagSource, (uint64_t)k, step);
  Pair b0 = block_d(ctr, key, 0), b1 = block_d(ctr, key, 1), b2 = block_d(ctr, key, 2), b3 = block_d(ctr, key, 3);
  int g = G - 1;
  for (int q = 0; q < G; ++q)
    if (b0.u0 <= gcdf[q]) {
      g = q;
      break;
    }
  const double* cg = cdfg + (int64_t)g * ncell;
  int64_t c = 0;
  if (cg[ncell - 1] > 0.0) c = mg_pick(cg, ncell, b0.u1 * cg[ncell - 1]);
  int i = (int)(c / ny), j = (int)(c % ny);
  double p1, p2, d1, d2;
  mg_pos(geom, e1, e2, i, j, b1.u0, b1.u1, p1, p2);
  mg_isotropic(geom, b2, d1, d2);
  mg_store(P, base + k, total / (double)N, d1, d2, dt * b3.u0, p1, p2, i, j, g);
}

// ---------------------------------------------------------------------------
// transport (MG_IMC2D.py:774-989)
// ---------------------------------------------------------------------------
template <typename R>
struct MgArgs {
  using R2 = typename Vec2<R>::type;
  MgView<R> P;
  int64_t n;
  const R2* e1p;  // FP64: (e1[i], e1[i+1]); FP32: (e1[i], e1[i+1] - e1[i])
  const R2* e2p;  // the same for e2
  int nx, ny, G;
  const R2* sas;  // (sigma_a, sigma_s) per bin
  double dt;
  int reflect;
  int64_t max_events;
  const double* floor_ptr;
  Key key;
  uint32_t step;
  unsigned long long* next;
  int batch;
  double* dep_raw;   // copies x (nb + G + 4): bins, loss by group, loss by side
  double* path_raw;  // copies x nb: w s where sigma_a <= 1e-12
  int64_t nb;
  int copy_mask;
  unsigned long long* stats;
  QueueMap Q;
  int refill_min;  // a wave with active lanes refills once this many lanes are idle (0 or 1: any)
};

constexpr int kMgWaveSlots = 512;  // tally cache slots per wave of the 2-D tracker
// Private tally copies of the multigroup engine: the (group, cell) adds that
// leave the per-wave caches are FP64 atomics on far more bins than the gray
// engine has cells, and same-bin contention falls with the copy count
// (crooked pipe 240x240x10, mean tracker time over 20 steps: 1 copy 13.1 ms,
// 8 copies 8.3 ms, 32 copies 7.3 ms).
constexpr int kMgTallyCopies = 32;

template <typename R, int GEOM>
__global__ __launch_bounds__(256) void k_mg_transport(MgArgs<R> A) {
  using R2 = typename Vec2<R>::type;
  const double wfloor = *A.floor_ptr;
  const int64_t stride = A.nb + A.G + 4;
  const int copy = blockIdx.x & A.copy_mask;
  double* const dep_t = A.dep_raw + (int64_t)copy * stride;
  double* const path_t = A.path_raw + (int64_t)copy * A.nb;
  using Tally = WaveTally<kMgWaveSlots>;
  __shared__ __align__(8) char tally_lds[4 * Tally::kBytes];
  Tally tally(tally_lds + (threadIdx.x >> 6) * Tally::kBytes);
  const int nx = A.nx, ny = A.ny;
  const int64_t ncell = (int64_t)nx * ny;
  WaveQueue wq;
  bool active = false;
  int64_t k = 0;
  double w = 0.0, t = 0.0;
  R d1 = 0, d2 = 0, dcen = 0;
  // FP64: the position.  FP32: offsets from the cell's lower edges
  // (kAtUpperFace: on the upper face), or, while `fresh`, the position read
  // from the pool; q1, q2 hold the position to write back.
  R p1 = 0, p2 = 0, q1 = 0, q2 = 0;
  bool fresh = false;
  int i = 0, j = 0, g = 0;
  uint32_t ev = 0;
  double acc_dep = 0.0, acc_path = 0.0;
  int acc_bin = -1;
  uint32_t st_ev = 0, st_bnd = 0, st_sc = 0, st_cen = 0, st_floor = 0, st_refl = 0, st_cap = 0;
  while (true) {
    uint64_t need = __ballot(!active);
    if (need && (__popcll(need) >= A.refill_min || need == ~0ull)) {
      const int64_t kk = wave_take(wq, !active, need, A.next, A.batch, A.Q, A.n);
      if (kk >= 0) {
        k = kk;
        w = A.P.w[k];
        d1 = A.P.d1[k];
        d2 = A.P.d2[k];
        p1 = A.P.p1[k];
        p2 = A.P.p2[k];
        q1 = p1;
        q2 = p2;
        fresh = true;
        t = (double)A.P.t[k];
        i = A.P.ci[k];
        j = A.P.cj[k];
        g = A.P.g[k];
        // FP32 times are stored rounded to nearest: a time sampled just below
        // dt can read back above it, and a negative distance to census would
        // grow the weight by e^{sigma_a |dcen|}.  Such a particle is at census.
        dcen = (R)fmax((A.dt - t) * kC, 0.0);
        ev = 0;
        active = w > wfloor;  // below the floor: skipped, unchanged (MG_IMC2D.py:842-843)
      }
    }
    if (__ballot(active) == 0) {
      if (wq.exhausted) break;
      continue;
    }
    if (!active) continue;
    bool finish = false;
    if (i < 0 || i >= nx || j < 0 || j >= ny) {
      int side = i < 0 ? 0 : (i >= nx ? 1 : (j < 0 ? 2 : 3));
      int kg = (int)(A.nb + g), ks = (int)(A.nb + A.G + side);
      tally.add(kg, w, dep_t);
      tally.add(ks, w, dep_t);
      ++st_bnd;
      w = 0.0;
      finish = true;
    } else if (w <= wfloor) {
      ++st_floor;  // weight kept in census (MG_IMC2D.py:854-856)
      finish = true;
    } else {
      const int64_t c = (int64_t)i * ny + j;
      const int bin = (int)((int64_t)g * ncell + c);
      R2 mat = A.sas[bin];
      const R2 ex = A.e1p[i];
      const R2 ey = A.e2p[j];
      Ctr ctr = make_ctr(kTagTransport, (uint64_t)k, A.step);
      double dep, path;
      int evt, crossed;
      if constexpr (sizeof(R) == 8) {
        MoveMG o;
        double dc = kC * (A.dt - t);
        if (GEOM == 0)
          mg_move_xy(ctr, A.key, ev, w, d1, d2, p1, p2, ex.x, ex.y, ey.x, ey.y, mat.x, mat.y, dc, o);
        else
          mg_move_rz(ctr, A.key, ev, w, d1, d2, p1, p2, ex.x, ex.y, ey.x, ey.y, mat.x, mat.y, dc, o);
        dep = o.dep;
        path = o.path;
        evt = o.evt;
        crossed = o.crossed;
        w = o.w;
        d1 = o.d1; d2 = o.d2; p1 = o.p1; p2 = o.p2;
        t = t + o.s / kC;
      } else {
        MoveOut<float> o;
        segment_move<GEOM, false>(ctr, A.key, ev, w, d1, d2, p1, p2, fresh, ex, ey, mat, dcen, o);
        dep = o.dep_w;
        path = o.dep_i;
        evt = o.evt;
        crossed = o.crossed;
        w = o.w;
        d1 = o.d1; d2 = o.d2; p1 = o.p1; p2 = o.p2;
        q1 = ex.x + p1;
        q2 = ey.x + p2;
        dcen = o.dcen;
      }
      ++ev;
      ++st_ev;
      if (bin != acc_bin) {
        if (acc_bin >= 0) {
          tally.add(acc_bin, acc_dep, dep_t);
          if (acc_path != 0.0) atomic_add_f64(path_t + acc_bin, acc_path);
        }
        acc_bin = bin;
        acc_dep = 0.0;
        acc_path = 0.0;
      }
      acc_dep += dep;
      if (!(mat.x > (R)1e-12)) acc_path += path;
      if (evt == kEvtCensus) {
        ++st_cen;
        finish = true;
      } else if (evt == kEvtScatter) {
        ++st_sc;
      } else {
        ++st_bnd;
        if (crossed == kCrossIPlus) { ++i; enter_lower(p1); }
        else if (crossed == kCrossIMinus) { --i; enter_upper(p1); }
        else if (crossed == kCrossJPlus) { ++j; enter_lower(p2); }
        else if (crossed == kCrossJMinus) { --j; enter_upper(p2); }
        int side = -1;
        if (i < 0) side = 0;
        else if (i >= nx) side = 1;
        else if (j < 0) side = 2;
        else if (j >= ny) side = 3;
        if (side >= 0) {
          if ((A.reflect >> side) & 1) {
            reflect_at(side, nx, ny, i, j, d1, d2, p1, p2);
            ++st_refl;
          } else {
            int kg = (int)(A.nb + g), ks = (int)(A.nb + A.G + side);
            tally.add(kg, w, dep_t);
            tally.add(ks, w, dep_t);
            w = 0.0;
            finish = true;
          }
        }
      }
      if (!finish && (int64_t)ev >= A.max_events) {
        ++st_cap;
        finish = true;
      }
    }
    if (finish) {
      A.P.w[k] = w;
      A.P.d1[k] = d1;
      A.P.d2[k] = d2;
      A.P.p1[k] = kF32<R> ? q1 : p1;
      A.P.p2[k] = kF32<R> ? q2 : p2;
      A.P.t[k] = (R)(sizeof(R) == 8 ? t : A.dt - (double)dcen / kC);
      A.P.ci[k] = i;
      A.P.cj[k] = j;
      active = false;
    }
  }
  if (acc_bin >= 0) {
    tally.add(acc_bin, acc_dep, dep_t);
    if (acc_path != 0.0) atomic_add_f64(path_t + acc_bin, acc_path);
  }
  tally.flush(dep_t);
  unsigned long long s0 = wave_sum<unsigned long long>(st_ev);
  unsigned long long s1 = wave_sum<unsigned long long>(st_bnd);
  unsigned long long s2 = wave_sum<unsigned long long>(st_sc);
  unsigned long long s3 = wave_sum<unsigned long long>(st_cen);
  unsigned long long s5 = wave_sum<unsigned long long>(st_floor);
  unsigned long long s6 = wave_sum<unsigned long long>(st_refl);
  unsigned long long s7 = wave_sum<unsigned long long>(st_cap);
  if (__lane_id() == 0) {
    atomicAdd(A.stats + 0, s0);
    atomicAdd(A.stats + 1, s1);
    atomicAdd(A.stats + 2, s2);
    atomicAdd(A.stats + 3, s3);
    atomicAdd(A.stats + 5, s5);
    atomicAdd(A.stats + 6, s6);
    atomicAdd(A.stats + 7, s7);
  }
}

__global__ void k_mg_floor(const double* __restrict__ sumw, int64_t n, double* __restrict__ out) {
#pragma clang fp contract(off)
  *out = 1e-10 * (*sumw) / (double)(n > 1 ? n : 1);
}

// dep = sum dep / V; si = sum dep / sigma_a / (dt V) or sum w s / (dt V)
// (MG_IMC2D.py:872-885); losses: [total, by group, by side].
__global__ void k_mg_finalize(int64_t nb, int64_t ncell, int G, int copies, const double* __restrict__ dep_raw,
                              const double* __restrict__ path_raw, const double* __restrict__ sa64,
                              const float2* __restrict__ sas32, int f32, const double* __restrict__ vol, double dt,
                              double* __restrict__ dep, double* __restrict__ si, double* __restrict__ bl) {
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = nb + G + 4;
  if (b < nb) {
    double a = 0.0, p = 0.0;
    for (int k = 0; k < copies; ++k) {
      a += dep_raw[(int64_t)k * stride + b];
      p += path_raw[(int64_t)k * nb + b];
    }
    int64_t c = b % ncell;
    double sa = f32 ? (double)sas32[b].x : sa64[b];
    bool thick = f32 ? (sas32[b].x > 1e-12f) : (sa64[b] > 1e-12);
    dep[b] = a / vol[c];
    si[b] = thick ? a / sa / (dt * vol[c]) : p / (dt * vol[c]);
  } else if (b < nb + G + 4) {
    double a = 0.0;
    for (int k = 0; k < copies; ++k) a += dep_raw[(int64_t)k * stride + b];
    bl[1 + (b - nb)] = a;
  }
}

__global__ void k_mg_bl_total(int G, double* __restrict__ bl) {
  double s = 0.0;
  for (int g = 0; g < G; ++g) s += bl[1 + g];
  bl[0] = s;
}

// ---------------------------------------------------------------------------
// comb over (group, cell) bins (MG_IMC2D.py:1270-1378)
// ---------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ int mg_bin(const MgView<R>& P, int64_t k, int nx, int ny) {
  int ci = P.ci[k], cj = P.cj[k];
  if (!(P.w[k] > 0.0) || ci < 0 || ci >= nx || cj < 0 || cj >= ny) return -1;
  return (int)((int64_t)P.g[k] * nx * ny + (int64_t)ci * ny + cj);
}

template <typename R>
__global__ __launch_bounds__(256) void k_mg_ecen(MgView<R> P, int64_t n, int nx, int ny, int64_t nb, int copy_mask,
                                                 double* __restrict__ part) {
  using Tally = WaveTally<kMgWaveSlots>;
  __shared__ __align__(8) char tally_lds[4 * Tally::kBytes];
  Tally tally(tally_lds + (threadIdx.x >> 6) * Tally::kBytes);
  double* const dst = part + (int64_t)(blockIdx.x & copy_mask) * nb;
  // one contiguous range of the pool per workgroup, four tiles per trip (the
  // pool comes cell by cell, so a range covers few bins and the wave caches
  // keep them; a grid-stride walk spread every wave over the whole mesh)
  constexpr int kU = 4;
  const int64_t span = (int64_t)kU * blockDim.x;
  const int64_t per = ((n + gridDim.x - 1) / gridDim.x + span - 1) / span * span;
  const int64_t lo = (int64_t)blockIdx.x * per;
  const int64_t hi = n < lo + per ? n : lo + per;
  // uniform trip count over the workgroup: the run sum needs all 64 lanes
  for (int64_t k0 = lo; k0 < hi; k0 += span) {
    int key[kU];
    double v[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int64_t k = k0 + u * (int64_t)blockDim.x + threadIdx.x;
      key[u] = -1;
      v[u] = 0.0;
      if (k < hi) {
        key[u] = mg_bin(P, k, nx, ny);
        if (key[u] >= 0) v[u] = P.w[k];
      }
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      bool head = wave_run_sum(key[u], v[u]);
      if (head && key[u] >= 0) tally.add(key[u], v[u], dst);
    }
  }
  tally.flush(dst);
}

__global__ void k_mg_ew(int64_t nb, const double* __restrict__ ecen, const double* __restrict__ Etot, int64_t Nmax,
                        double* __restrict__ ew) {
#pragma clang fp contract(off)
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  double e = ecen[b], E = *Etot;
  if (e > 0.0 && E > 0.0) {
    double d = rint((double)Nmax * e / E);
    if (d < 1.0) d = 1.0;
    ew[b] = e / d;
  } else {
    ew[b] = 0.0;
  }
}

template <typename R>
__global__ void k_mg_copies(MgView<R> P, int64_t n, int nx, int ny, const double* __restrict__ ew, int passthrough,
                            Key key, uint32_t step, int32_t* __restrict__ copies) {
#pragma clang fp contract(off)
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  int b = mg_bin(P, k, nx, ny);
  int32_t m = 0;
  if (b >= 0) {
    if (passthrough) {
      m = 1;
    } else {
      double et = ew[b];
      if (et > 0.0) m = (int32_t)(P.w[k] / et + slot_d(make_ctr(kTagComb, (uint64_t)k, step), key, 0));
    }
  }
  copies[k] = m;
}

template <typename R>
__global__ __launch_bounds__(256) void k_mg_comb_write(MgView<R> S, MgView<R> D, int64_t n, int nx, int ny,
                                                       const int64_t* __restrict__ offs, const double* __restrict__ ew,
                                                       int passthrough) {
  __shared__ int64_t so[257];
  const int t = threadIdx.x;
  const int64_t k0 = (int64_t)blockIdx.x * 256;
  const int cnt = (n - k0 < 256) ? (int)(n - k0) : 256;
  if (t < cnt) so[t] = offs[k0 + t];
  if (t == 0) so[cnt] = offs[k0 + cnt];
  __syncthreads();
  const int64_t lo = so[0], hi = so[cnt];
  const int64_t ncell = (int64_t)nx * ny;
  for (int64_t o = lo + t; o < hi; o += 256) {
    int a = 0, b = cnt - 1;
    while (a < b) {
      int mid = (a + b + 1) >> 1;
      if (so[mid] <= o)
        a = mid;
      else
        b = mid - 1;
    }
    const int64_t src = k0 + a;
    const int ci = S.ci[src], cj = S.cj[src], g = S.g[src];
    D.w[o] = passthrough ? S.w[src] : ew[(int64_t)g * ncell + (int64_t)ci * ny + cj];
    D.d1[o] = S.d1[src];
    D.d2[o] = S.d2[src];
    D.p1[o] = S.p1[src];
    D.p2[o] = S.p2[src];
    D.t[o] = (R)0;  // times reset after the comb (MG_IMC2D.py:2018)
    D.ci[o] = ci;
    D.cj[o] = cj;
    D.g[o] = g;
  }
}

__global__ void k_sub(int64_t n, const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ c) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) c[i] = a[i] - b[i];
}

// out[c] = sum over g of a[g, c] (or of a[g, c] - b[g, c]), accumulated from
// g = 0 upwards: the order of numpy's sum over axis 0 of a (G, ...) array.
__global__ void k_group_sum(int64_t ncell, int G, const double* __restrict__ a, const double* __restrict__ b,
                            double* __restrict__ out) {
#pragma clang fp contract(off)
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  double acc = b ? a[c] - b[c] : a[c];
  for (int g = 1; g < G; ++g) {
    int64_t q = (int64_t)g * ncell + c;
    acc += b ? a[q] - b[q] : a[q];
  }
  out[c] = acc;
}

template <typename R>
__global__ void k_mg_put(MgView<R> P, int64_t n, const double* __restrict__ w, const double* __restrict__ d1,
                         const double* __restrict__ d2, const double* __restrict__ t, const double* __restrict__ p1,
                         const double* __restrict__ p2, const int64_t* __restrict__ ci, const int64_t* __restrict__ cj,
                         const int32_t* __restrict__ g) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  mg_store(P, k, w[k], d1[k], d2[k], t[k], p1[k], p2[k], (int)ci[k], (int)cj[k], g[k]);
}

template <typename R>
__global__ void k_mg_get(MgView<R> P, int64_t n, double* __restrict__ w, double* __restrict__ d1,
                         double* __restrict__ d2, double* __restrict__ t, double* __restrict__ p1,
                         double* __restrict__ p2, int64_t* __restrict__ ci, int64_t* __restrict__ cj,
                         int32_t* __restrict__ g) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  w[k] = P.w[k];
  d1[k] = (double)P.d1[k];
  d2[k] = (double)P.d2[k];
  t[k] = (double)P.t[k];
  p1[k] = (double)P.p1[k];
  p2[k] = (double)P.p2[k];
  ci[k] = P.ci[k];
  cj[k] = P.cj[k];
  g[k] = P.g[k];
}

// ---------------------------------------------------------------------------
// spherical 1-D multigroup (MG_IMC/MG_IMC1D.py): pool d1 = mu, p1 = r,
// ci = cell; bins b = g I + i (the (group, cell) layout with ny = 1)
// ---------------------------------------------------------------------------
// a Tr^4 V per cell, zero below T_emit_floor (MG_IMC1D.py:526-528)
__global__ void k_m1_eq_energy(int I, const double* __restrict__ Tr, const double* __restrict__ vol,
                               double T_emit_floor, double* __restrict__ E) {
#pragma clang fp contract(off)
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= I) return;
  double t = Tr[i];
  double e = kA * ((t * t) * (t * t)) * vol[i];
  if (T_emit_floor > 0.0 && t < T_emit_floor) e = 0.0;
  E[i] = e;
}

// Counts by rounding half to even: E_b > 0 gets max(min_one, round(N E_b /
// E_total)) (MG_IMC1D.py:536-539, 626-628, 712-714).
__global__ void k_m1_counts(int64_t nb, const double* __restrict__ E, int64_t N, int min_one,
                            const double* __restrict__ Etot, int64_t* __restrict__ cnt) {
#pragma clang fp contract(off)
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  double e = E[b], et = *Etot;
  int64_t m = 0;
  if (e > 0.0 && et > 0.0) {
    double d = rint((double)N * e / et);
    if (d < (double)min_one) d = (double)min_one;
    m = (int64_t)d;
  }
  cnt[b] = m;
}

__device__ __forceinline__ int64_t m1_bin_of(const int64_t* __restrict__ offs, int64_t nb, int64_t k) {
  int64_t lo = 0, hi = nb - 1;  // last b with offs[b] <= k
  while (lo < hi) {
    int64_t mid = (lo + hi + 1) >> 1;
    if (offs[mid] <= k)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ double m1_radius(double r0, double r1, double u) {
#pragma clang fp contract(off)
  return pow(r0 * r0 * r0 + (r1 * r1 * r1 - r0 * r0 * r0) * u, 1.0 / 3.0);
}

template <typename R>
__device__ __forceinline__ void m1_store(MgView<R> P, int64_t o, double w, double mu, double t, double r, int i,
                                         int g) {
  P.w[o] = w;
  P.d1[o] = (R)mu;
  P.d2[o] = (R)0;
  P.t[o] = (R)t;
  P.p1[o] = (R)r;
  P.p2[o] = (R)0;
  P.ci[o] = i;
  P.cj[o] = 0;
  P.g[o] = g;
}

// Equilibrium census (MG_IMC1D.py:514-566).  Slots: 0 radius, 1 mu, 2..6
// Planck group at the cell's Tr.  offs: exclusive offsets over cells.
template <typename R>
__global__ void k_m1_sample_ic(MgView<R> P, int64_t total, const int64_t* __restrict__ offs, int I,
                               const double* __restrict__ e, const double* __restrict__ Tr,
                               const double* __restrict__ E, int G, const double* __restrict__ edges,
                               const double* __restrict__ pcdf, Key key) {
#pragma clang fp contract(off)
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= total) return;
  int i = (int)m1_bin_of(offs, I, k);
  int64_t n = offs[i + 1] - offs[i];
  Ctr ctr = make_ctr(kTagIC, (uint64_t)k, 0);
  Pair b0 = block_d(ctr, key, 0);
  int g = planck_group(ctr, key, 2, Tr[i], G, edges, pcdf);
  m1_store(P, k, E[i] / (double)n, 2.0 * b0.u1 - 1.0, 0.0, m1_radius(e[i], e[i + 1], b0.u0), i, g);
}

// Volume source by (group, cell) bin (emission MG_IMC1D.py:569-650, fixed
// source 679-743).  Slots: 0 radius, 1 mu, 2 time.
template <typename R>
__global__ void k_m1_sample_bins(MgView<R> P, int64_t base, int64_t total, const int64_t* __restrict__ offs,
                                 int64_t nb, int I, const double* __restrict__ e, const double* __restrict__ E,
                                 double dt, Key key, uint32_t step, uint32_t tag) {
#pragma clang fp contract(off)
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= total) return;
  int64_t b = m1_bin_of(offs, nb, k);
  int64_t n = offs[b + 1] - offs[b];
  int i = (int)(b % I), g = (int)(b / I);
  Ctr ctr = make_ctr(tag, (uint64_t)k, step);
  Pair b0 = block_d(ctr, key, 0), b1 = block_d(ctr, key, 1);
  m1_store(P, base + k, E[b] / (double)n, 2.0 * b0.u1 - 1.0, dt * b1.u0, m1_radius(e[i], e[i + 1], b0.u0), i, g);
}

// Half-Lambertian sphere source (MG_IMC1D.py:653-676, placed by step()
// 1061-1088).  Slots: 0 |mu| = sqrt(u), 1 time, 2..6 Planck group at T.
template <typename R>
__global__ void k_m1_sample_face(MgView<R> P, int64_t base, int64_t n, int side, double E, double T, double dt,
                                 double Rf, int I, int G, const double* __restrict__ edges,
                                 const double* __restrict__ pcdf, Key key, uint32_t step) {
#pragma clang fp contract(off)
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  Ctr ctr = make_ctr(kTagBoundary0 + side, (uint64_t)k, step);
  Pair b0 = block_d(ctr, key, 0);
  double m = sqrt(b0.u0);
  int g = planck_group(ctr, key, 2, T, G, edges, pcdf);
  m1_store(P, base + k, E / (double)n, side == 0 ? m : -m, dt * b0.u1, Rf, side == 0 ? 0 : I - 1, g);
}

// Transport (MG_IMC1D.py:277-461): the gray spherical move (opacity cap
// included) with group opacities; histories run while dcen > 1e-15; an
// escape books the weight by group (tally key nb + g); a weight below the
// floor is deposited where the particle stands.  stats: events, boundary,
// scatter, census, 0, floor kills, reflections, non-positive distances.
template <typename R>
__global__ __launch_bounds__(256) void k_m1_transport(MgArgs<R> A) {
  using R2 = typename Vec2<R>::type;
  const double wfloor = *A.floor_ptr;
  const int64_t stride = A.nb + A.G + 4;
  const int copy = blockIdx.x & A.copy_mask;
  double* const dep_t = A.dep_raw + (int64_t)copy * stride;
  double* const path_t = A.path_raw + (int64_t)copy * A.nb;
  __shared__ int lkey[kLdsSlots];
  __shared__ double lval[kLdsSlots];
  lds_tally_init(lkey, lval);
  __syncthreads();
  const int I = A.nx;
  WaveQueue wq;
  bool active = false;
  int64_t k = 0;
  double w = 0.0;
  R mu = 0, dcen = 0;
  // FP64: the radius.  FP32: the offset from the cell's inner radius
  // (kAtUpperFace: on the outer shell), or, while `fresh`, the radius read
  // from the pool; q holds the radius to write back.
  R x = 0, q = 0;
  bool fresh = false;
  int cell = 0, g = 0;
  uint32_t ev = 0;
  double acc_dep = 0.0, acc_path = 0.0;
  int acc_bin = -1;
  uint32_t st_ev = 0, st_bnd = 0, st_sc = 0, st_cen = 0, st_floor = 0, st_refl = 0, st_bad = 0;
  while (true) {
    uint64_t need = __ballot(!active);
    if (need) {
      const int64_t kk = wave_take(wq, !active, need, A.next, A.batch, A.Q, A.n);
      if (kk >= 0) {
        k = kk;
        w = A.P.w[k];
        mu = A.P.d1[k];
        x = A.P.p1[k];
        q = x;
        fresh = true;
        cell = A.P.ci[k];
        g = A.P.g[k];
        dcen = (R)((A.dt - (double)A.P.t[k]) * kC);
        ev = 0;
        active = dcen > (R)1e-15 && cell >= 0 && cell < I;
      }
    }
    if (__ballot(active) == 0) {
      if (wq.exhausted) break;
      continue;
    }
    if (!active) continue;
    bool finish = false;
    ++st_ev;
    const int bin = g * I + cell;
    const R2 edg = A.e1p[cell];
    R2 mat = A.sas[bin];
    Move1D<R> o;
    Ctr ctr = make_ctr(kTagTransport, (uint64_t)k, A.step);
    if constexpr (kF32<R>) {
      x = fresh ? fminf(fmaxf(x - edg.x, 0.0f), edg.y) : resolve_offset(x, edg.y);
      fresh = false;
    }
    move_1d<1>(ctr, A.key, ev, w, mu, x, edg.x, edg.y, mat.x, mat.y, dcen, o);
    ++ev;
    if (o.bad) {
      ++st_bad;
      finish = true;
    } else {
      if (bin != acc_bin) {
        if (acc_bin >= 0) {
          if (!lds_tally_add(lkey, lval, acc_bin, acc_dep)) atomic_add_f64(dep_t + acc_bin, acc_dep);
          if (acc_path != 0.0) atomic_add_f64(path_t + acc_bin, acc_path);
        }
        acc_bin = bin;
        acc_dep = 0.0;
        acc_path = 0.0;
      }
      acc_dep += o.dep;
      if (!(mat.x > (R)1e-10)) acc_path += o.path;
      w = o.w;
      mu = o.mu;
      x = o.x;
      if constexpr (kF32<R>) q = edg.x + x;
      cell += o.newloc;
      if (o.newloc > 0) enter_lower(x);
      else if (o.newloc < 0) enter_upper(x);
      dcen = o.dcen;
      st_bnd += o.code == 1;
      st_sc += o.code == 2;
      st_cen += o.code == 3;
      bool gone = false;
      if (cell >= I || cell < 0) {
        int face = cell < 0 ? 0 : 1;
        if ((A.reflect >> face) & 1) {
          mu = -mu;
          cell = face == 0 ? 0 : I - 1;
          if (face == 0) enter_lower(x);
          else enter_upper(x);
          ++st_refl;
        } else {
          int kg = (int)(A.nb + g);
          if (!lds_tally_add(lkey, lval, kg, w)) atomic_add_f64(dep_t + kg, w);
          w = 0.0;
          gone = true;
        }
      }
      if (gone) {
        finish = true;
      } else if (w < wfloor && w > 0.0) {
        ++st_floor;
        const int b2 = g * I + cell;
        if (!lds_tally_add(lkey, lval, b2, w)) atomic_add_f64(dep_t + b2, w);
        w = 0.0;
        finish = true;
      } else if (!(dcen > (R)1e-15)) {
        finish = true;
      }
    }
    if (finish) {
      A.P.w[k] = w;
      A.P.d1[k] = mu;
      A.P.p1[k] = kF32<R> ? q : x;
      A.P.ci[k] = cell;
      active = false;
    }
  }
  if (acc_bin >= 0) {
    if (!lds_tally_add(lkey, lval, acc_bin, acc_dep)) atomic_add_f64(dep_t + acc_bin, acc_dep);
    if (acc_path != 0.0) atomic_add_f64(path_t + acc_bin, acc_path);
  }
  __syncthreads();
  lds_tally_flush(lkey, lval, dep_t);
  unsigned long long s0 = wave_sum<unsigned long long>(st_ev);
  unsigned long long s1 = wave_sum<unsigned long long>(st_bnd);
  unsigned long long s2 = wave_sum<unsigned long long>(st_sc);
  unsigned long long s3 = wave_sum<unsigned long long>(st_cen);
  unsigned long long s5 = wave_sum<unsigned long long>(st_floor);
  unsigned long long s6 = wave_sum<unsigned long long>(st_refl);
  unsigned long long s7 = wave_sum<unsigned long long>(st_bad);
  if (__lane_id() == 0) {
    atomicAdd(A.stats + 0, s0);
    atomicAdd(A.stats + 1, s1);
    atomicAdd(A.stats + 2, s2);
    atomicAdd(A.stats + 3, s3);
    atomicAdd(A.stats + 5, s5);
    atomicAdd(A.stats + 6, s6);
    atomicAdd(A.stats + 7, s7);
  }
}

__global__ void k_m1_floor(const double* __restrict__ sumw, int64_t n, double* __restrict__ out) {
#pragma clang fp contract(off)
  *out = 1e-14 * (*sumw) / (double)(n > 1 ? n : 1);  // MG_IMC1D.py:1138
}

// dep = sum dep / V; si = sum dep / (sigma_a dt V) where sigma_a > 1e-10,
// else sum w s / (dt V) (MG_IMC1D.py:341-347, 413-414, 455-458); loss by group.
__global__ void k_m1_finalize(int64_t nb, int I, int G, int copies, const double* __restrict__ dep_raw,
                              const double* __restrict__ path_raw, const double* __restrict__ sa64,
                              const float2* __restrict__ sas32, int f32, const double* __restrict__ vol, double dt,
                              double* __restrict__ dep, double* __restrict__ si, double* __restrict__ bl) {
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = nb + G + 4;
  if (b < nb) {
    double a = 0.0, p = 0.0;
    for (int k = 0; k < copies; ++k) {
      a += dep_raw[(int64_t)k * stride + b];
      p += path_raw[(int64_t)k * nb + b];
    }
    int64_t c = b % I;
    double sa = f32 ? (double)sas32[b].x : sa64[b];
    bool thick = f32 ? (sas32[b].x > 1e-10f) : (sa64[b] > 1e-10);
    dep[b] = a / vol[c];
    si[b] = thick ? a / sa / (dt * vol[c]) : p / (dt * vol[c]);
  } else if (b < nb + G) {
    double a = 0.0;
    for (int k = 0; k < copies; ++k) a += dep_raw[(int64_t)k * stride + b];
    bl[b - nb] = a;
  }
}

// Energy of bins whose combed copies all vanished: the one loss the
// "radiation" mode cannot rescale away.
__global__ void k_m1_lost(int64_t nb, const double* __restrict__ ecen, const double* __restrict__ post,
                          double* __restrict__ lost) {
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nb) lost[b] = post[b] > 0.0 ? 0.0 : ecen[b];
}

// Histogram of w mu per bin over live particles (the radial flux tally,
// MG_IMC1D.py:1270-1276).
template <typename R>
__global__ __launch_bounds__(256) void k_m1_flux(MgView<R> P, int64_t n, int I, int64_t nb, int copy_mask,
                                                 double* __restrict__ part) {
  __shared__ int lkey[kLdsSlots];
  __shared__ double lval[kLdsSlots];
  lds_tally_init(lkey, lval);
  __syncthreads();
  double* const dst = part + (int64_t)(blockIdx.x & copy_mask) * nb;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t trips = (n + stride - 1) / stride;
  for (int64_t tr = 0; tr < trips; ++tr) {
    int64_t k = tr * stride + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int key = -1;
    double v = 0.0;
    if (k < n) {
      key = mg_bin(P, k, I, 1);
      if (key >= 0) v = P.w[k] * (double)P.d1[k];
    }
    bool head = wave_run_sum(key, v);
    if (head && key >= 0 && !lds_tally_add(lkey, lval, key, v)) atomic_add_f64(dst + key, v);
  }
  __syncthreads();
  lds_tally_flush(lkey, lval, dst);
}

// "radiation" comb mode: survivors of each bin rescaled to the bin's
// pre-comb energy (MG_IMC1D.py:811-823).
template <typename R>
__global__ void k_m1_rescale(MgView<R> P, int64_t n, int I, const double* __restrict__ ecen,
                             const double* __restrict__ post) {
#pragma clang fp contract(off)
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  int b = mg_bin(P, k, I, 1);
  if (b < 0) return;
  double q = post[b];
  if (q > 0.0) P.w[k] = P.w[k] * (ecen[b] / (q + 1e-300));
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
struct MgCtx {
  int geometry = 0, nx = 0, ny = 0, G = 0, precision = 0, device = 0, n_cu = 256;
  int64_t ncell = 0, nb = 0;
  uint64_t seed = 1;
  uint32_t step = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  float last_ms = 0.0f;
  std::vector<double> h_e1, h_e2, h_edges;
  DevBuf<double> e1, e2, edges, vol, pcdf, frac, T, cv, sig, sa, ss, bstar, E, Ecell, Eplanck, cdf, local, chunk,
      choff, dep_raw, path_raw, dep, si, bl, ecen_part, ecen, ew, tmp, post, cdfg, gcdf, prof, Tface, scal, sum_scr,
      tree_scr;
  DevBuf<double2> sas64, e1p64, e2p64;
  DevBuf<float2> sas32, e1p32, e2p32;
  DevBuf<int64_t> counts, scan_scr;
  DevBuf<int32_t> copies, guide;
  DevBuf<unsigned long long> gcount;
  DevBuf<double> Ecg;                       // group energies cell-major
  DevBuf<double> spacing, spacing_sum;      // emission cells: exponential spacings, running sum
  unsigned long long* stats = nullptr;
  unsigned long long* next = nullptr;
  MgPool A, B;
  int64_t n = 0;
  int grid_override = 0, batch = 256, tally_copies = kMgTallyCopies;
  int refill_min = 1;  // idle lanes before a wave of the 2-D tracker refills
  MgPool snap;         // mgimc_snapshot
  int64_t n_snap = -1;  // < 0: no snapshot saved
  bool have_material = false;
  // page-locked staging of the (G, ncell) transfers: slot 0 opacities in,
  // slot 1 scalar intensity out, slot 2 post-comb census energy out
  PinnedBuf stage;
  // device arrays that mgimc_group_sum may read
  bool e_valid = false, dep_valid = false, post_valid = false;
};

static int prefix_canonical(MgCtx* c, const double* x, int64_t n, double* out) {
  int64_t nch = (n + 1023) / 1024;
  hipStream_t s = c->stream;
  k_prefix_chunks<<<(unsigned)nch, 256, 0, s>>>(x, n, c->local.p, c->chunk.p);
  IMC_HIP(hipGetLastError());
  k_prefix_offsets<<<1, 1, 0, s>>>(c->chunk.p, nch, c->choff.p);
  IMC_HIP(hipGetLastError());
  k_prefix_finish<<<launch_blocks(n, 256), 256, 0, s>>>(c->local.p, c->choff.p, n, out);
  IMC_HIP(hipGetLastError());
  return kOk;
}

// Three-level running sum (k_prefix_groups); scratch: c->local (n), c->chunk
// and c->choff (chunks), c->tmp (groups).
static int prefix_three_level(MgCtx* c, const double* x, int64_t n, double* out) {
  const int64_t nch = (n + 1023) / 1024, ngroup = (nch + 1023) / 1024;
  int rc;
  if ((rc = c->local.reserve(n)) || (rc = c->chunk.reserve(nch)) || (rc = c->choff.reserve(nch)) ||
      (rc = c->tmp.reserve(ngroup)))
    return rc;
  hipStream_t s = c->stream;
  k_prefix_chunks<<<(unsigned)nch, 256, 0, s>>>(x, n, c->local.p, c->chunk.p);
  IMC_HIP(hipGetLastError());
  k_prefix_groups<<<launch_blocks(ngroup, 64), 64, 0, s>>>(c->chunk.p, nch, c->choff.p, c->tmp.p);
  IMC_HIP(hipGetLastError());
  k_prefix_finish3<<<launch_blocks(n, 256), 256, 0, s>>>(c->local.p, c->choff.p, c->tmp.p, ngroup, n, out);
  IMC_HIP(hipGetLastError());
  return kOk;
}

static int mg_histogram(MgCtx* c, const MgPool& P, int64_t n, double* out) {
  hipStream_t s = c->stream;
  int64_t nb = c->nb;
  IMC_HIP(hipMemsetAsync(c->ecen_part.p, 0, sizeof(double) * nb * c->tally_copies, s));
  if (n > 0) {
    int g = std::min(launch_blocks(n, 256), c->n_cu * 4);
    if (c->precision == IMC_FP64)
      k_mg_ecen<double><<<g, 256, 0, s>>>(P.view<double>(), n, c->nx, c->ny, nb, c->tally_copies - 1,
                                          c->ecen_part.p);
    else
      k_mg_ecen<float><<<g, 256, 0, s>>>(P.view<float>(), n, c->nx, c->ny, nb, c->tally_copies - 1, c->ecen_part.p);
    IMC_HIP(hipGetLastError());
  }
  k_reduce_copies<double, double><<<launch_blocks(nb, 256), 256, 0, s>>>(nb, c->tally_copies, c->ecen_part.p, out);
  IMC_HIP(hipGetLastError());
  return kOk;
}

}  // namespace imc

using namespace imc;

extern "C" {

struct mgimc_ctx : public MgCtx {};

static int mg_create(mgimc_ctx** out, int geometry, int nx, int ny, const double* edges1, const double* edges2, int G,
                     const double* energy_edges, int precision, uint64_t seed, int device) {
  IMC_REQUIRE(out && edges1 && edges2 && energy_edges, kErrArg, "mgimc_create: null argument");
  IMC_REQUIRE(geometry == IMC_GEOM_XY || geometry == IMC_GEOM_RZ || geometry == IMC_GEOM_SPHERE, kErrArg,
              "mgimc_create: bad geometry");
  IMC_REQUIRE(nx > 0 && ny > 0 && G > 0 && G <= kMaxGroups, kErrArg, "mgimc_create: bad sizes");
  IMC_REQUIRE(precision == IMC_FP64 || precision == IMC_FP32_TRACK, kErrArg, "mgimc_create: bad precision");
  for (int a = 0; a < nx; ++a) IMC_REQUIRE(edges1[a + 1] > edges1[a], kErrArg, "mgimc_create: edges1 not increasing");
  for (int a = 0; a < ny; ++a) IMC_REQUIRE(edges2[a + 1] > edges2[a], kErrArg, "mgimc_create: edges2 not increasing");
  for (int a = 0; a < G; ++a)
    IMC_REQUIRE(energy_edges[a + 1] > energy_edges[a] && energy_edges[a] >= 0.0, kErrArg,
                "mgimc_create: energy edges not increasing");
  if (geometry != IMC_GEOM_XY) IMC_REQUIRE(edges1[0] >= 0.0, kErrArg, "mgimc_create: negative radius");
  IMC_REQUIRE((int64_t)nx * ny * G + G + 4 < (int64_t)0x7fffffff, kErrArg, "mgimc_create: too many bins");
  IMC_HIP(hipSetDevice(device));
  mgimc_ctx* c = new mgimc_ctx();
  c->geometry = geometry;
  c->nx = nx;
  c->ny = ny;
  c->G = G;
  c->ncell = (int64_t)nx * ny;
  c->nb = c->ncell * G;
  c->precision = precision;
  c->seed = seed;
  c->device = device;
  hipDeviceProp_t prop;
  IMC_HIP(hipGetDeviceProperties(&prop, device));
  c->n_cu = prop.multiProcessorCount;
  IMC_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  IMC_HIP(hipEventCreate(&c->ev0));
  IMC_HIP(hipEventCreate(&c->ev1));
  c->h_e1.assign(edges1, edges1 + nx + 1);
  c->h_e2.assign(edges2, edges2 + ny + 1);
  c->h_edges.assign(energy_edges, energy_edges + G + 1);
  int64_t nc = c->ncell, nb = c->nb;
  int rc = 0;
  struct R {
    DevBuf<double>* b;
    int64_t n;
  } req[] = {{&c->e1, nx + 1},  {&c->e2, ny + 1},   {&c->edges, G + 1}, {&c->vol, nc},
             {&c->pcdf, 200},   {&c->frac, G},      {&c->T, nc},        {&c->cv, nc},
             {&c->sig, nb},     {&c->sa, nb},       {&c->ss, nb},       {&c->bstar, nb},
             {&c->E, nb},       {&c->Ecell, nc},    {&c->Eplanck, nc},  {&c->cdf, nc},
             {&c->local, nc},   {&c->chunk, nc / 1024 + 2}, {&c->choff, nc / 1024 + 2},
             {&c->dep_raw, (nb + G + 4) * kMaxTallyCopies}, {&c->path_raw, nb * kMaxTallyCopies},
             {&c->dep, nb},     {&c->si, nb},       {&c->bl, G + 5},    {&c->ecen_part, nb * kMaxTallyCopies},
             {&c->ecen, nb},    {&c->ew, nb},       {&c->tmp, nb},      {&c->post, nb},
             {&c->cdfg, nb},    {&c->gcdf, G},      {&c->prof, std::max(nx, ny) + 1},
             {&c->Tface, std::max(nx, ny) + 1},     {&c->scal, 16},     {&c->sum_scr, canonical_sum_scratch(nb) + 8},
             {&c->tree_scr, 1024}};
  for (auto& r : req)
    if ((rc = r.b->reserve(r.n))) return rc;
  if ((rc = c->sas64.reserve(nb)) || (rc = c->sas32.reserve(nb)) || (rc = c->e1p64.reserve(nx)) ||
      (rc = c->e2p64.reserve(ny)) || (rc = c->e1p32.reserve(nx)) || (rc = c->e2p32.reserve(ny)) ||
      (rc = c->gcount.reserve(G)) || (rc = c->stage.reserve(sizeof(double) * 3 * (size_t)nb)))
    return rc;
  IMC_HIP(hipMemcpy(c->e1.p, edges1, sizeof(double) * (nx + 1), hipMemcpyHostToDevice));
  IMC_HIP(hipMemcpy(c->e2.p, edges2, sizeof(double) * (ny + 1), hipMemcpyHostToDevice));
  IMC_HIP(hipMemcpy(c->edges.p, energy_edges, sizeof(double) * (G + 1), hipMemcpyHostToDevice));
  {
    // cell volumes (MG_IMC2D.py:146-157; sphere 4/3 pi (r1^3 - r0^3),
    // MG_IMC1D.py:920)
    std::vector<double> vol((size_t)nc);
    for (int a = 0; a < nx; ++a)
      for (int b = 0; b < ny; ++b) {
        double dy = edges2[b + 1] - edges2[b];
        double r0 = edges1[a], r1 = edges1[a + 1];
        if (geometry == IMC_GEOM_XY)
          vol[(size_t)a * ny + b] = (r1 - r0) * dy;
        else if (geometry == IMC_GEOM_RZ)
          vol[(size_t)a * ny + b] = M_PI * (r1 * r1 - r0 * r0) * dy;
        else
          vol[(size_t)a * ny + b] = (4.0 / 3.0) * M_PI * (r1 * r1 * r1 - r0 * r0 * r0);
      }
    IMC_HIP(hipMemcpy(c->vol.p, vol.data(), sizeof(double) * nc, hipMemcpyHostToDevice));
    // Planck mixture CDF over n = 1..200 (MG_IMC2D.py:99-101), sequential sums
    double inv[200], s = 0.0, cum = 0.0, pc[200];
    for (int q = 1; q <= 200; ++q) {
      double d = (double)q;
      inv[q - 1] = 1.0 / (d * d * d * d);
      s += inv[q - 1];
    }
    for (int q = 0; q < 200; ++q) {
      cum += inv[q] / s;
      pc[q] = cum;
    }
    IMC_HIP(hipMemcpy(c->pcdf.p, pc, sizeof(pc), hipMemcpyHostToDevice));
    std::vector<float2> a32(nx), b32(ny);
    std::vector<double2> a64(nx), b64(ny);
    for (int q = 0; q < nx; ++q) {
      a32[q] = make_float2((float)edges1[q], (float)edges1[q + 1] - (float)edges1[q]);
      a64[q] = make_double2(edges1[q], edges1[q + 1]);
    }
    for (int q = 0; q < ny; ++q) {
      b32[q] = make_float2((float)edges2[q], (float)edges2[q + 1] - (float)edges2[q]);
      b64[q] = make_double2(edges2[q], edges2[q + 1]);
    }
    IMC_HIP(hipMemcpy(c->e1p32.p, a32.data(), sizeof(float2) * nx, hipMemcpyHostToDevice));
    IMC_HIP(hipMemcpy(c->e2p32.p, b32.data(), sizeof(float2) * ny, hipMemcpyHostToDevice));
    IMC_HIP(hipMemcpy(c->e1p64.p, a64.data(), sizeof(double2) * nx, hipMemcpyHostToDevice));
    IMC_HIP(hipMemcpy(c->e2p64.p, b64.data(), sizeof(double2) * ny, hipMemcpyHostToDevice));
  }
  IMC_HIP(hipMalloc(&c->stats, sizeof(unsigned long long) * 8));
  IMC_HIP(hipMalloc(&c->next, sizeof(unsigned long long)));
  c->A.rsz = c->B.rsz = c->snap.rsz = precision == IMC_FP64 ? 8 : 4;
  *out = c;
  return kOk;
}

int mgimc_create(mgimc_ctx** out, int geometry, int nx, int ny, const double* edges1, const double* edges2, int G,
                 const double* energy_edges, int precision, uint64_t seed, int device) {
  IMC_REQUIRE(geometry == IMC_GEOM_XY || geometry == IMC_GEOM_RZ, kErrArg, "mgimc_create: bad geometry");
  return mg_create(out, geometry, nx, ny, edges1, edges2, G, energy_edges, precision, seed, device);
}

// Spherical 1-D multigroup context (MG_IMC/MG_IMC1D.py): I radial cells, the
// pool's d1 is mu, p1 the radius, ci the cell; d2, p2, cj stay unused.
int mgimc_create_sphere(mgimc_ctx** out, int I, const double* edges, int G, const double* energy_edges,
                        int precision, uint64_t seed, int device) {
  const double unit[2] = {0.0, 1.0};
  return mg_create(out, IMC_GEOM_SPHERE, I, 1, edges, unit, G, energy_edges, precision, seed, device);
}

void mgimc_destroy(mgimc_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  DevBuf<double>* d[] = {&c->e1, &c->e2, &c->edges, &c->vol, &c->pcdf, &c->frac, &c->T, &c->cv, &c->sig,
                         &c->sa, &c->ss, &c->bstar, &c->E, &c->Ecell, &c->Eplanck, &c->cdf, &c->local, &c->chunk,
                         &c->choff, &c->dep_raw, &c->path_raw, &c->dep, &c->si, &c->bl, &c->ecen_part, &c->ecen,
                         &c->ew, &c->tmp, &c->post, &c->cdfg, &c->gcdf, &c->prof, &c->Tface, &c->scal, &c->sum_scr,
                         &c->tree_scr};
  for (auto* b : d) b->release();
  c->sas64.release();
  c->sas32.release();
  c->e1p64.release();
  c->e2p64.release();
  c->e1p32.release();
  c->e2p32.release();
  c->counts.release();
  c->scan_scr.release();
  c->copies.release();
  c->guide.release();
  c->gcount.release();
  c->Ecg.release();
  c->spacing.release();
  c->spacing_sum.release();
  c->stage.release();
  c->A.release();
  c->B.release();
  c->snap.release();
  if (c->stats) (void)hipFree(c->stats);
  if (c->next) (void)hipFree(c->next);
  (void)hipEventDestroy(c->ev0);
  (void)hipEventDestroy(c->ev1);
  (void)hipStreamDestroy(c->stream);
  delete c;
}

int mgimc_set_step(mgimc_ctx* c, uint32_t step) {
  IMC_REQUIRE(c, kErrArg, "null ctx");
  c->step = step;
  return kOk;
}

int mgimc_set_seed(mgimc_ctx* c, uint64_t seed) {
  IMC_REQUIRE(c, kErrArg, "null ctx");
  c->seed = seed;
  return kOk;
}

int mgimc_set_launch(mgimc_ctx* c, int grid, int batch, int tally_copies) {
  IMC_REQUIRE(c && grid >= 0 && batch > 0, kErrArg, "mgimc_set_launch: bad arguments");
  IMC_REQUIRE(tally_copies >= 1 && tally_copies <= kMaxTallyCopies && (tally_copies & (tally_copies - 1)) == 0,
              kErrArg, "mgimc_set_launch: tally_copies must be a power of two in [1, %d]", kMaxTallyCopies);
  c->grid_override = grid;
  c->batch = batch;
  c->tally_copies = tally_copies;
  return kOk;
}

int mgimc_set_tracker(mgimc_ctx* c, int block, int refill_min) {
  IMC_REQUIRE(c, kErrArg, "null ctx");
  IMC_REQUIRE(block == 256, kErrArg, "mgimc_set_tracker: block must be 256");
  IMC_REQUIRE(refill_min >= 1 && refill_min <= 64, kErrArg, "mgimc_set_tracker: refill_min must be in [1, 64]");
  c->refill_min = refill_min;
  return kOk;
}

// Copy of the pool: save = 1 keeps the current pool, save = 0 puts the kept
// one back (it stays kept).
int mgimc_snapshot(mgimc_ctx* c, int save) {
  IMC_REQUIRE(c, kErrArg, "null ctx");
  IMC_REQUIRE(save || c->n_snap >= 0, kErrState, "mgimc_snapshot: no snapshot saved to restore");
  IMC_HIP(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  MgPool& from = save ? c->A : c->snap;
  MgPool& to = save ? c->snap : c->A;
  const int64_t n = save ? c->n : c->n_snap;
  int rc = to.grow(std::max<int64_t>(n, 1), 0, s);
  if (rc) return rc;
  for (int k = 0; k < 9; ++k)
    if (n > 0) IMC_HIP(hipMemcpyAsync(to.a[k], from.a[k], from.esz(k) * (size_t)n, hipMemcpyDeviceToDevice, s));
  IMC_HIP(hipStreamSynchronize(s));
  if (save)
    c->n_snap = n;
  else
    c->n = n;
  return kOk;
}

int64_t mgimc_census_size(mgimc_ctx* c) { return c ? c->n : -1; }

int mgimc_put_census(mgimc_ctx* c, int64_t n, const double* w, const double* d1, const double* d2, const double* t,
                     const double* p1, const double* p2, const int64_t* ci, const int64_t* cj, const int32_t* g) {
  IMC_REQUIRE(c && n >= 0, kErrArg, "mgimc_put_census: bad arguments");
  IMC_HIP(hipSetDevice(c->device));
  int rc = c->A.grow(std::max<int64_t>(n, 1), 0, c->stream);
  if (rc) return rc;
  c->n = 0;
  if (n > 0) {
    const void* src[9] = {w, d1, d2, t, p1, p2, ci, cj, g};
    size_t esz[9] = {8, 8, 8, 8, 8, 8, 8, 8, 4};
    void* tmp[9];
    for (int a = 0; a < 9; ++a) {
      IMC_REQUIRE(src[a], kErrArg, "mgimc_put_census: null array");
      IMC_HIP(hipMalloc(&tmp[a], esz[a] * (size_t)n));
      IMC_HIP(hipMemcpyAsync(tmp[a], src[a], esz[a] * (size_t)n, hipMemcpyHostToDevice, c->stream));
    }
    int gr = launch_blocks(n, 256);
    if (c->precision == IMC_FP64)
      k_mg_put<double><<<gr, 256, 0, c->stream>>>(c->A.view<double>(), n, (double*)tmp[0], (double*)tmp[1],
                                                  (double*)tmp[2], (double*)tmp[3], (double*)tmp[4], (double*)tmp[5],
                                                  (int64_t*)tmp[6], (int64_t*)tmp[7], (int32_t*)tmp[8]);
    else
      k_mg_put<float><<<gr, 256, 0, c->stream>>>(c->A.view<float>(), n, (double*)tmp[0], (double*)tmp[1],
                                                 (double*)tmp[2], (double*)tmp[3], (double*)tmp[4], (double*)tmp[5],
                                                 (int64_t*)tmp[6], (int64_t*)tmp[7], (int32_t*)tmp[8]);
    IMC_HIP(hipGetLastError());
    IMC_HIP(hipStreamSynchronize(c->stream));
    for (int a = 0; a < 9; ++a) IMC_HIP(hipFree(tmp[a]));
  }
  c->n = n;
  return kOk;
}

int mgimc_get_census(mgimc_ctx* c, int64_t n, double* w, double* d1, double* d2, double* t, double* p1, double* p2,
                     int64_t* ci, int64_t* cj, int32_t* g) {
  IMC_REQUIRE(c && n == c->n, kErrArg, "mgimc_get_census: size mismatch");
  IMC_HIP(hipSetDevice(c->device));
  if (n == 0) return kOk;
  void* dst[9] = {w, d1, d2, t, p1, p2, ci, cj, g};
  size_t esz[9] = {8, 8, 8, 8, 8, 8, 8, 8, 4};
  void* tmp[9];
  for (int a = 0; a < 9; ++a) {
    IMC_REQUIRE(dst[a], kErrArg, "mgimc_get_census: null array");
    IMC_HIP(hipMalloc(&tmp[a], esz[a] * (size_t)n));
  }
  int gr = launch_blocks(n, 256);
  if (c->precision == IMC_FP64)
    k_mg_get<double><<<gr, 256, 0, c->stream>>>(c->A.view<double>(), n, (double*)tmp[0], (double*)tmp[1],
                                                (double*)tmp[2], (double*)tmp[3], (double*)tmp[4], (double*)tmp[5],
                                                (int64_t*)tmp[6], (int64_t*)tmp[7], (int32_t*)tmp[8]);
  else
    k_mg_get<float><<<gr, 256, 0, c->stream>>>(c->A.view<float>(), n, (double*)tmp[0], (double*)tmp[1],
                                               (double*)tmp[2], (double*)tmp[3], (double*)tmp[4], (double*)tmp[5],
                                               (int64_t*)tmp[6], (int64_t*)tmp[7], (int32_t*)tmp[8]);
  IMC_HIP(hipGetLastError());
  for (int a = 0; a < 9; ++a)
    IMC_HIP(hipMemcpyAsync(dst[a], tmp[a], esz[a] * (size_t)n, hipMemcpyDeviceToHost, c->stream));
  IMC_HIP(hipStreamSynchronize(c->stream));
  for (int a = 0; a < 9; ++a) IMC_HIP(hipFree(tmp[a]));
  return kOk;
}

// Census energy per (group, cell) bin and in total.
int mgimc_census_by_bin(mgimc_ctx* c, double* out, double* total_out) {
  IMC_REQUIRE(c, kErrArg, "null ctx");
  IMC_HIP(hipSetDevice(c->device));
  int rc;
  if ((rc = mg_histogram(c, c->A, c->n, c->post.p))) return rc;
  if ((rc = tree_sum((const double*)c->A.a[0], c->n, c->scal.p + 1, c->tree_scr.p, c->stream))) return rc;
  if (out) IMC_HIP(hipMemcpyAsync(out, c->post.p, sizeof(double) * c->nb, hipMemcpyDeviceToHost, c->stream));
  double tot = 0.0;
  IMC_HIP(hipMemcpyAsync(&tot, c->scal.p + 1, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  IMC_HIP(hipStreamSynchronize(c->stream));
  if (total_out) *total_out = tot;
  return kOk;
}

// Fleck factor with the Planck-mean opacity, b*_g (or the caller's emission
// fractions), emission energies by (group, cell) and their cell CDF.
// sigma: (G, ncell) true opacities.  Outputs nullable.
int mgimc_set_material(mgimc_ctx* c, const double* T, const double* sigma, const double* cv, double theta, double dt,
                       const double* emission_fractions, double T_emit_floor, double* sa_out, double* ss_out,
                       double* E_out, double* Etot_out, double* Eplanck_out) {
  IMC_REQUIRE(c && T && sigma && cv, kErrArg, "mgimc_set_material: null argument");
  IMC_HIP(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  int64_t nc = c->ncell, nb = c->nb;
  IMC_HIP(hipMemcpyAsync(c->T.p, T, sizeof(double) * nc, hipMemcpyHostToDevice, s));
  if (host_device_alias(sigma, sizeof(double) * nb)) {
    // page-locked (imc_host_alloc): one DMA copy, no host memcpy
    IMC_HIP(hipMemcpyAsync(c->sig.p, sigma, sizeof(double) * nb, hipMemcpyHostToDevice, s));
  } else {
    double* h = c->stage.slot(nb, 0);
    std::memcpy(h, sigma, sizeof(double) * nb);
    IMC_HIP(hipMemcpyAsync(c->sig.p, h, sizeof(double) * nb, hipMemcpyHostToDevice, s));
  }
  IMC_HIP(hipMemcpyAsync(c->cv.p, cv, sizeof(double) * nc, hipMemcpyHostToDevice, s));
  if (emission_fractions)
    IMC_HIP(hipMemcpyAsync(c->frac.p, emission_fractions, sizeof(double) * c->G, hipMemcpyHostToDevice, s));
  k_mg_fleck<<<launch_blocks(nc, 128), 128, 0, s>>>(nc, c->G, c->edges.p, c->T.p, c->sig.p, c->cv.p, theta, dt,
                                                    emission_fractions ? c->frac.p : nullptr, T_emit_floor, c->vol.p,
                                                    c->sa.p, c->ss.p, c->sas64.p, c->sas32.p, c->bstar.p, c->E.p,
                                                    c->Ecell.p, c->Eplanck.p);
  IMC_HIP(hipGetLastError());
  int rc = prefix_canonical(c, c->Ecell.p, nc, c->cdf.p);
  if (rc) return rc;
  if (sa_out) IMC_HIP(hipMemcpyAsync(sa_out, c->sa.p, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
  if (ss_out) IMC_HIP(hipMemcpyAsync(ss_out, c->ss.p, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
  if (E_out) IMC_HIP(hipMemcpyAsync(E_out, c->E.p, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
  if (Eplanck_out) IMC_HIP(hipMemcpyAsync(Eplanck_out, c->Eplanck.p, sizeof(double) * nc, hipMemcpyDeviceToHost, s));
  double Et = 0.0;
  IMC_HIP(hipMemcpyAsync(&Et, c->cdf.p + nc - 1, sizeof(double), hipMemcpyDeviceToHost, s));
  IMC_HIP(hipStreamSynchronize(s));
  if (Etot_out) *Etot_out = Et;
  c->have_material = true;
  c->e_valid = true;
  return kOk;
}

// Equilibrium census at t = 0 replacing the pool (MG_IMC2D.py:992-1095).
int mgimc_equilibrium_sample(mgimc_ctx* c, int64_t N, const double* Tr, double T_emit_floor) {
  IMC_REQUIRE(c && Tr && N >= 0, kErrArg, "mgimc_equilibrium_sample: bad arguments");
  IMC_HIP(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  int64_t nc = c->ncell;
  std::vector<double> tr(Tr, Tr + nc), e(nc);
  {
    // a Tr^4 V per cell, zero below the floor; volumes as on the device
    std::vector<double> vol(nc);
    IMC_HIP(hipMemcpy(vol.data(), c->vol.p, sizeof(double) * nc, hipMemcpyDeviceToHost));
    for (int64_t q = 0; q < nc; ++q) {
      double t = tr[q];
      e[q] = kA * ((t * t) * (t * t)) * vol[q];
      if (T_emit_floor > 0.0 && t < T_emit_floor) e[q] = 0.0;
    }
  }
  IMC_HIP(hipMemcpyAsync(c->T.p, tr.data(), sizeof(double) * nc, hipMemcpyHostToDevice, s));
  IMC_HIP(hipMemcpyAsync(c->Ecell.p, e.data(), sizeof(double) * nc, hipMemcpyHostToDevice, s));
  int rc = prefix_canonical(c, c->Ecell.p, nc, c->cdf.p);
  if (rc) return rc;
  double Et = 0.0;
  IMC_HIP(hipMemcpyAsync(&Et, c->cdf.p + nc - 1, sizeof(double), hipMemcpyDeviceToHost, s));
  IMC_HIP(hipStreamSynchronize(s));
  c->n = 0;
  if (N <= 0 || !(Et > 0.0)) return kOk;
  if ((rc = c->A.grow(N, 0, s))) return rc;
  int g = launch_blocks(N, 256);
  Key key = make_key(c->seed);
  if (c->precision == IMC_FP64)
    k_mg_sample_ic<double><<<g, 256, 0, s>>>(c->A.view<double>(), N, c->geometry, c->nx, c->ny, c->G, c->e1.p,
                                             c->e2.p, c->edges.p, c->T.p, c->cdf.p, c->pcdf.p, key);
  else
    k_mg_sample_ic<float><<<g, 256, 0, s>>>(c->A.view<float>(), N, c->geometry, c->nx, c->ny, c->G, c->e1.p, c->e2.p,
                                            c->edges.p, c->T.p, c->cdf.p, c->pcdf.p, key);
  IMC_HIP(hipGetLastError());
  IMC_HIP(hipStreamSynchronize(s));
  c->n = N;
  c->have_material = false;  // T now holds Tr
  return kOk;
}

// Face source appended to the pool.  prof_cum / Tface (nullable, nface
// entries): running sum of per-face-cell energy and the face-cell
// temperatures.  E_by_group_out (nullable, G): energy of the new particles by group.
int mgimc_sample_boundary(mgimc_ctx* c, int side, int64_t n, double E_total, double dt, double Tb,
                          const double* prof_cum, const double* Tface, int nface, double* E_by_group_out) {
  IMC_REQUIRE(c && side >= 0 && side < 4 && n >= 0, kErrArg, "mgimc_sample_boundary: bad arguments");
  if (n == 0) {
    if (E_by_group_out)
      for (int g = 0; g < c->G; ++g) E_by_group_out[g] = 0.0;
    return kOk;
  }
  IMC_HIP(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  const double *prof = nullptr, *tf = nullptr;
  if (prof_cum) {
    IMC_REQUIRE(Tface && nface == (side < 2 ? c->ny : c->nx), kErrArg, "mgimc_sample_boundary: bad profile");
    IMC_HIP(hipMemcpyAsync(c->prof.p, prof_cum, sizeof(double) * nface, hipMemcpyHostToDevice, s));
    IMC_HIP(hipMemcpyAsync(c->Tface.p, Tface, sizeof(double) * nface, hipMemcpyHostToDevice, s));
    prof = c->prof.p;
    tf = c->Tface.p;
  }
  int rc = c->A.grow(c->n + n, c->n, s);
  if (rc) return rc;
  if (E_by_group_out) IMC_HIP(hipMemsetAsync(c->gcount.p, 0, sizeof(unsigned long long) * c->G, s));
  int g = launch_blocks(n, 256);
  Key key = make_key(c->seed);
  unsigned long long* gc = E_by_group_out ? c->gcount.p : nullptr;
  const size_t lds = gc ? sizeof(unsigned int) * (size_t)c->G : 0;
  if (c->precision == IMC_FP64)
    k_mg_sample_boundary<double><<<g, 256, lds, s>>>(c->A.view<double>(), c->n, n, c->geometry, side, E_total, dt, Tb,
                                                   c->nx, c->ny, c->e1.p, c->e2.p, c->G, c->edges.p, c->pcdf.p, prof,
                                                   tf, key, c->step, gc);
  else
    k_mg_sample_boundary<float><<<g, 256, lds, s>>>(c->A.view<float>(), c->n, n, c->geometry, side, E_total, dt, Tb,
                                                  c->nx, c->ny, c->e1.p, c->e2.p, c->G, c->edges.p, c->pcdf.p, prof,
                                                  tf, key, c->step, gc);
  IMC_HIP(hipGetLastError());
  if (E_by_group_out) {
    std::vector<unsigned long long> cnt(c->G);
    IMC_HIP(hipMemcpyAsync(cnt.data(), c->gcount.p, sizeof(unsigned long long) * c->G, hipMemcpyDeviceToHost, s));
    IMC_HIP(hipStreamSynchronize(s));
    double w = E_total / (double)n;
    for (int q = 0; q < c->G; ++q) E_by_group_out[q] = w * (double)cnt[q];
  } else if (prof) {
    IMC_HIP(hipStreamSynchronize(s));
  }
  c->n += n;
  return kOk;
}

// Fixed source appended to the pool: S (G, ncell) energy per (group, cell)
// this step (source x V x dt).  total_out: the source energy.
int mgimc_sample_source(mgimc_ctx* c, int64_t N, const double* S, double dt, double* total_out) {
  IMC_REQUIRE(c && S && N >= 0, kErrArg, "mgimc_sample_source: bad arguments");
  IMC_HIP(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  int64_t nc = c->ncell;
  int G = c->G;
  IMC_HIP(hipMemcpyAsync(c->tmp.p, S, sizeof(double) * c->nb, hipMemcpyHostToDevice, s));
  int rc;
  for (int g = 0; g < G; ++g)
    if ((rc = prefix_canonical(c, c->tmp.p + (int64_t)g * nc, nc, c->cdfg.p + (int64_t)g * nc))) return rc;
  std::vector<double> tot(G), gcdf(G);
  for (int g = 0; g < G; ++g)
    IMC_HIP(hipMemcpyAsync(&tot[g], c->cdfg.p + (int64_t)g * nc + nc - 1, sizeof(double), hipMemcpyDeviceToHost, s));
  IMC_HIP(hipStreamSynchronize(s));
  // _sample_group_piecewise_constant (MG_IMC2D.py:307-338)
  double sum = 0.0;
  for (int g = 0; g < G; ++g) sum += tot[g];
  double cum = 0.0;
  for (int g = 0; g < G; ++g) {
    cum += sum > 0.0 ? tot[g] / sum : 1.0 / (double)G;
    gcdf[g] = cum;
  }
  gcdf[G - 1] = 1.0;
  double total = 0.0;
  for (int g = 0; g < G; ++g) total += tot[g];
  if (total_out) *total_out = total;
  if (N <= 0 || !(total > 0.0)) return kOk;
  IMC_HIP(hipMemcpyAsync(c->gcdf.p, gcdf.data(), sizeof(double) * G, hipMemcpyHostToDevice, s));
  if ((rc = c->A.grow(c->n + N, c->n, s))) return rc;
  int gr = launch_blocks(N, 256);
  Key key = make_key(c->seed);
  if (c->precision == IMC_FP64)
    k_mg_sample_source<double><<<gr, 256, 0, s>>>(c->A.view<double>(), c->n, N, c->geometry, c->nx, c->ny, G,
                                                  c->e1.p, c->e2.p, c->cdfg.p, c->gcdf.p, total, dt, key, c->step);
  else
    k_mg_sample_source<float><<<gr, 256, 0, s>>>(c->A.view<float>(), c->n, N, c->geometry, c->nx, c->ny, G, c->e1.p,
                                                 c->e2.p, c->cdfg.p, c->gcdf.p, total, dt, key, c->step);
  IMC_HIP(hipGetLastError());
  IMC_HIP(hipStreamSynchronize(s));  // gcdf is a pageable source
  c->n += N;
  return kOk;
}

// Material emission appended to the pool: N particles of weight E/N from the
// energies of the last mgimc_set_material.  Returns the emitted energy.
int mgimc_sample_emission(mgimc_ctx* c, int64_t N, double dt, double* Etot_out) {
  IMC_REQUIRE(c && N >= 0, kErrArg, "mgimc_sample_emission: bad arguments");
  IMC_REQUIRE(c->have_material, kErrState, "mgimc_sample_emission: call mgimc_set_material first");
  IMC_HIP(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  double Et = 0.0;
  IMC_HIP(hipMemcpyAsync(&Et, c->cdf.p + c->ncell - 1, sizeof(double), hipMemcpyDeviceToHost, s));
  IMC_HIP(hipStreamSynchronize(s));
  if (Etot_out) *Etot_out = Et;
  if (N <= 0 || !(Et > 0.0)) return kOk;
  int rc = c->A.grow(c->n + N, c->n, s);
  if (rc) return rc;
  if ((rc = c->guide.reserve(kGuide + 1))) return rc;
  k_mg_guide<<<launch_blocks(kGuide + 1, 256), 256, 0, s>>>(c->cdf.p, c->ncell, c->guide.p);
  IMC_HIP(hipGetLastError());
  int g = launch_blocks(N, 256);
  Key key = make_key(c->seed);
  // sorted uniforms for the cells: N + 1 exponential spacings and their
  // running sum in three levels (the oracle's om_prefix3 order)
  const int64_t n1 = N + 1;
  if ((rc = c->Ecg.reserve(c->nb)) || (rc = c->spacing.reserve(n1)) || (rc = c->spacing_sum.reserve(n1)))
    return rc;
  k_mg_cell_major<<<launch_blocks(c->nb, 256), 256, 0, s>>>(c->E.p, c->ncell, c->G, c->Ecg.p);
  IMC_HIP(hipGetLastError());
  k_mg_emission_spacings<<<launch_blocks(n1, 256), 256, 0, s>>>(n1, key, c->step, c->spacing.p);
  IMC_HIP(hipGetLastError());
  if ((rc = prefix_three_level(c, c->spacing.p, n1, c->spacing_sum.p))) return rc;
  if (c->precision == IMC_FP64)
    k_mg_sample_emission<double><<<g, 256, 0, s>>>(c->A.view<double>(), c->n, N, c->geometry, c->nx, c->ny, c->G,
                                                   c->e1.p, c->e2.p, c->Ecg.p, c->cdf.p, c->guide.p, dt, key, c->step,
                                                   c->spacing_sum.p);
  else
    k_mg_sample_emission<float><<<g, 256, 0, s>>>(c->A.view<float>(), c->n, N, c->geometry, c->nx, c->ny, c->G,
                                                  c->e1.p, c->e2.p, c->Ecg.p, c->cdf.p, c->guide.p, dt, key, c->step,
                                                  c->spacing_sum.p);
  IMC_HIP(hipGetLastError());
  c->n += N;
  return kOk;
}

// Transport the pool (MG_IMC2D.py:774-989).  weight_floor < 0: 1e-10 sum(w)/N.
// Outputs (nullable): dep, si (G, ncell); bl[1 + G + 4] = total, by group,
// by side; stats[8] = events, boundary crossings, scatters, census, 0,
// floor kills, reflections, event-cap hits.
int mgimc_transport(mgimc_ctx* c, double dt, const int32_t* reflect, int64_t max_events, double weight_floor,
                    double* dep_out, double* si_out, double* bl_out, int64_t* stats_out) {
  IMC_REQUIRE(c && reflect && max_events > 0, kErrArg, "mgimc_transport: bad arguments");
  IMC_REQUIRE(c->have_material, kErrState, "mgimc_transport: call mgimc_set_material first");
  IMC_HIP(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  int64_t n = c->n, nb = c->nb;
  int G = c->G;
  IMC_HIP(hipMemsetAsync(c->dep_raw.p, 0, sizeof(double) * (nb + G + 4) * c->tally_copies, s));
  IMC_HIP(hipMemsetAsync(c->path_raw.p, 0, sizeof(double) * nb * c->tally_copies, s));
  IMC_HIP(hipMemsetAsync(c->stats, 0, sizeof(unsigned long long) * 8, s));
  IMC_HIP(hipMemsetAsync(c->next, 0, sizeof(unsigned long long), s));
  int rc;
  if (weight_floor < 0.0) {
    if ((rc = tree_sum((const double*)c->A.a[0], n, c->scal.p + 1, c->tree_scr.p, s))) return rc;
    k_mg_floor<<<1, 1, 0, s>>>(c->scal.p + 1, n, c->scal.p + 2);
    IMC_HIP(hipGetLastError());
  } else {
    IMC_HIP(hipMemcpyAsync(c->scal.p + 2, &weight_floor, sizeof(double), hipMemcpyHostToDevice, s));
  }
  int refl = 0;
  for (int a = 0; a < 4; ++a) refl |= (reflect[a] ? 1 : 0) << a;
  int block = 256, grid;
  const void* fn = c->precision == IMC_FP64
                       ? (c->geometry == IMC_GEOM_XY ? (const void*)k_mg_transport<double, 0>
                                                     : (const void*)k_mg_transport<double, 1>)
                       : (c->geometry == IMC_GEOM_XY ? (const void*)k_mg_transport<float, 0>
                                                     : (const void*)k_mg_transport<float, 1>);
  if (c->grid_override > 0) {
    grid = c->grid_override;
  } else {
    int per_cu = 0;
    IMC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, block, 0));
    grid = std::max(per_cu, 1) * c->n_cu;
  }
  grid = (int)std::max<int64_t>(std::min<int64_t>(grid, (n + block - 1) / block), 1);
  Key key = make_key(c->seed);
  // pool order: ordering the chunks by expected history length, as the gray
  // tracker does, measured no gain here, where the tally atomics bound the
  // time (crooked pipe, FP32 tracker: all chunks sorted 7.25 ms against 7.28
  // in pool order; only long chunks first 7.50 against 7.26)
  const QueueMap Q = pool_order_queue(n);
  IMC_HIP(hipEventRecord(c->ev0, s));
  if (n > 0) {
    if (c->precision == IMC_FP64) {
      MgArgs<double> A{c->A.view<double>(), n, c->e1p64.p, c->e2p64.p, c->nx, c->ny, G, c->sas64.p, dt, refl,
                       max_events, c->scal.p + 2, key, c->step, c->next, c->batch, c->dep_raw.p, c->path_raw.p, nb,
                       c->tally_copies - 1, c->stats, Q, c->refill_min};
      if (c->geometry == IMC_GEOM_XY)
        k_mg_transport<double, 0><<<grid, block, 0, s>>>(A);
      else
        k_mg_transport<double, 1><<<grid, block, 0, s>>>(A);
    } else {
      MgArgs<float> A{c->A.view<float>(), n, c->e1p32.p, c->e2p32.p, c->nx, c->ny, G, c->sas32.p, dt, refl,
                      max_events, c->scal.p + 2, key, c->step, c->next, c->batch, c->dep_raw.p, c->path_raw.p, nb,
                      c->tally_copies - 1, c->stats, Q, c->refill_min};
      if (c->geometry == IMC_GEOM_XY)
        k_mg_transport<float, 0><<<grid, block, 0, s>>>(A);
      else
        k_mg_transport<float, 1><<<grid, block, 0, s>>>(A);
    }
    IMC_HIP(hipGetLastError());
  }
  IMC_HIP(hipEventRecord(c->ev1, s));
  k_mg_finalize<<<launch_blocks(nb + G + 4, 256), 256, 0, s>>>(nb, c->ncell, G, c->tally_copies, c->dep_raw.p,
                                                                c->path_raw.p, c->sa.p, c->sas32.p,
                                                                c->precision == IMC_FP32_TRACK, c->vol.p, dt, c->dep.p,
                                                                c->si.p, c->bl.p);
  IMC_HIP(hipGetLastError());
  k_mg_bl_total<<<1, 1, 0, s>>>(G, c->bl.p);
  IMC_HIP(hipGetLastError());
  c->dep_valid = true;
  if (dep_out) IMC_HIP(hipMemcpyAsync(dep_out, c->dep.p, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
  // page-locked outputs (imc_host_alloc) take a direct DMA copy, others go
  // through the staging slot and a host copy after the synchronisation
  double* h_si = nullptr;
  if (si_out) {
    h_si = host_device_alias(si_out, sizeof(double) * nb) ? si_out : c->stage.slot(nb, 1);
    IMC_HIP(hipMemcpyAsync(h_si, c->si.p, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
  }
  if (bl_out) IMC_HIP(hipMemcpyAsync(bl_out, c->bl.p, sizeof(double) * (1 + G + 4), hipMemcpyDeviceToHost, s));
  unsigned long long st[8];
  IMC_HIP(hipMemcpyAsync(st, c->stats, sizeof(st), hipMemcpyDeviceToHost, s));
  IMC_HIP(hipStreamSynchronize(s));
  IMC_HIP(hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1));
  if (si_out && h_si != si_out) std::memcpy(si_out, h_si, sizeof(double) * nb);
  if (stats_out)
    for (int a = 0; a < 8; ++a) stats_out[a] = (int64_t)st[a];
  return kOk;
}

float mgimc_last_transport_ms(mgimc_ctx* c) { return c ? c->last_ms : -1.0f; }

// (group, cell) comb to about Nmax (MG_IMC2D.py:1314-1378).  Outputs
// (nullable): ecen and disc = ecen - ecen_after per bin, post = census energy
// per bin after the comb; ecen_total, total_rad scalars.
// Capacity of the comb's output buffer, which becomes the next pool: room
// for the copies and for a pool as large as the current one (census plus
// this step's sources), so the next step's sources need no reallocation.  A
// buffer short of that grows to 1.5 times it.  (Growing to the other pool
// buffer's capacity instead makes the two outgrow each other once their
// capacities differ: on the 240x240 x 10 group crooked pipe that reallocated
// 1.7 GB, 2.9 ms, every step.)
static int comb_output_capacity(mgimc_ctx* c, int64_t total, int64_t n) {
  const int64_t need = std::max<int64_t>(std::max<int64_t>(total, n), 1);
  if (c->B.cap >= need) return kOk;
  return c->B.grow(need + need / 2, 0, c->stream);
}

int mgimc_comb(mgimc_ctx* c, int64_t Nmax, double* ecen_out, double* disc_out, double* post_out,
               double* ecen_total_out, double* total_rad_out) {
  IMC_REQUIRE(c && Nmax >= 0, kErrArg, "mgimc_comb: bad arguments");
  IMC_HIP(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  int64_t n = c->n, nb = c->nb;
  int rc;
  if ((rc = mg_histogram(c, c->A, n, c->ecen.p))) return rc;
  if ((rc = c->sum_scr.reserve(canonical_sum_scratch(nb) + 8))) return rc;
  if ((rc = canonical_sum(c->ecen.p, nb, c->scal.p + 4, c->sum_scr.p, s))) return rc;
  double E = 0.0;
  IMC_HIP(hipMemcpyAsync(&E, c->scal.p + 4, sizeof(double), hipMemcpyDeviceToHost, s));
  IMC_HIP(hipStreamSynchronize(s));
  int passthrough = !(E > 0.0);
  k_mg_ew<<<launch_blocks(nb, 256), 256, 0, s>>>(nb, c->ecen.p, c->scal.p + 4, Nmax, c->ew.p);
  IMC_HIP(hipGetLastError());
  if ((rc = c->copies.reserve(n + 1)) || (rc = c->counts.reserve(n + 2)) ||
      (rc = c->scan_scr.reserve(scan_scratch(n) + 8)))
    return rc;
  Key key = make_key(c->seed);
  int g = launch_blocks(n, 256);
  if (n > 0) {
    if (c->precision == IMC_FP64)
      k_mg_copies<double><<<g, 256, 0, s>>>(c->A.view<double>(), n, c->nx, c->ny, c->ew.p, passthrough, key, c->step,
                                            c->copies.p);
    else
      k_mg_copies<float><<<g, 256, 0, s>>>(c->A.view<float>(), n, c->nx, c->ny, c->ew.p, passthrough, key, c->step,
                                           c->copies.p);
    IMC_HIP(hipGetLastError());
  }
  if ((rc = exclusive_scan_i32_to_i64(c->copies.p, n, c->counts.p, c->scan_scr.p, s))) return rc;
  int64_t total = 0;
  IMC_HIP(hipMemcpyAsync(&total, c->counts.p + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  IMC_HIP(hipStreamSynchronize(s));
  if ((rc = comb_output_capacity(c, total, n))) return rc;
  if (n > 0) {
    if (c->precision == IMC_FP64)
      k_mg_comb_write<double><<<g, 256, 0, s>>>(c->A.view<double>(), c->B.view<double>(), n, c->nx, c->ny,
                                                c->counts.p, c->ew.p, passthrough);
    else
      k_mg_comb_write<float><<<g, 256, 0, s>>>(c->A.view<float>(), c->B.view<float>(), n, c->nx, c->ny, c->counts.p,
                                               c->ew.p, passthrough);
    IMC_HIP(hipGetLastError());
  }
  std::swap(c->A, c->B);
  c->n = total;
  c->post_valid = disc_out || post_out;
  double* h_post = nullptr;
  if (disc_out || post_out) {
    if ((rc = mg_histogram(c, c->A, total, c->post.p))) return rc;
    if (disc_out) {
      k_sub<<<launch_blocks(nb, 256), 256, 0, s>>>(nb, c->ecen.p, c->post.p, c->tmp.p);
      IMC_HIP(hipGetLastError());
      IMC_HIP(hipMemcpyAsync(disc_out, c->tmp.p, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
    }
    if (post_out) {
      h_post = host_device_alias(post_out, sizeof(double) * nb) ? post_out : c->stage.slot(nb, 2);
      IMC_HIP(hipMemcpyAsync(h_post, c->post.p, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
    }
  }
  if (ecen_out) IMC_HIP(hipMemcpyAsync(ecen_out, c->ecen.p, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
  if ((rc = tree_sum((const double*)c->A.a[0], total, c->scal.p + 5, c->tree_scr.p, s))) return rc;
  double tr = 0.0;
  IMC_HIP(hipMemcpyAsync(&tr, c->scal.p + 5, sizeof(double), hipMemcpyDeviceToHost, s));
  IMC_HIP(hipStreamSynchronize(s));
  if (post_out && h_post != post_out) std::memcpy(post_out, h_post, sizeof(double) * nb);
  if (total_rad_out) *total_rad_out = tr;
  if (ecen_total_out) *ecen_total_out = E;
  return kOk;
}

// Per-cell sums over groups of a (G, ncell) device array, in the order of
// numpy's axis-0 sum: which = 0 emission energies of the last
// mgimc_set_material, 1 deposition of the last mgimc_transport, 2 comb
// discrepancy (pre-comb minus post-comb census energy) of the last
// mgimc_comb that produced disc or post.  out: ncell doubles.
int mgimc_group_sum(mgimc_ctx* c, int which, double* out) {
  IMC_REQUIRE(c && out && which >= 0 && which <= 2, kErrArg, "mgimc_group_sum: bad arguments");
  const double* a = nullptr;
  const double* b = nullptr;
  if (which == 0) {
    IMC_REQUIRE(c->e_valid, kErrState, "mgimc_group_sum: no emission energies yet (mgimc_set_material)");
    a = c->E.p;
  } else if (which == 1) {
    IMC_REQUIRE(c->dep_valid, kErrState, "mgimc_group_sum: no deposition yet (mgimc_transport)");
    a = c->dep.p;
  } else {
    IMC_REQUIRE(c->post_valid, kErrState, "mgimc_group_sum: the last mgimc_comb produced no post-comb census");
    a = c->ecen.p;
    b = c->post.p;
  }
  IMC_HIP(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  k_group_sum<<<launch_blocks(c->ncell, 256), 256, 0, s>>>(c->ncell, c->G, a, b, c->tmp.p);
  IMC_HIP(hipGetLastError());
  IMC_HIP(hipMemcpyAsync(out, c->tmp.p, sizeof(double) * c->ncell, hipMemcpyDeviceToHost, s));
  IMC_HIP(hipStreamSynchronize(s));
  return kOk;
}

// ---------------------------------------------------------------------------
// spherical 1-D multigroup entry points (MG_IMC/MG_IMC1D.py)
// ---------------------------------------------------------------------------
// Round-rule counts of E (nb entries, device) into c->counts, exclusive
// offsets into c->copies-free scratch c->scan: returns the total.
static int m1_counts_offsets(mgimc_ctx* c, const double* E, int64_t nb, int64_t N, int min_one, int64_t* total) {
  hipStream_t s = c->stream;
  int rc;
  if ((rc = c->sum_scr.reserve(canonical_sum_scratch(nb) + 8)) || (rc = c->counts.reserve(2 * (nb + 2))) ||
      (rc = c->scan_scr.reserve(scan_scratch(nb) + 8)))
    return rc;
  if ((rc = canonical_sum(E, nb, c->scal.p + 6, c->sum_scr.p, s))) return rc;
  int64_t* cnt = c->counts.p;
  int64_t* offs = c->counts.p + nb + 2;
  k_m1_counts<<<launch_blocks(nb, 256), 256, 0, s>>>(nb, E, N, min_one, c->scal.p + 6, cnt);
  IMC_HIP(hipGetLastError());
  if ((rc = exclusive_scan_i64(cnt, nb, offs, c->scan_scr.p, s))) return rc;
  IMC_HIP(hipMemcpyAsync(total, offs + nb, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  IMC_HIP(hipStreamSynchronize(s));
  return kOk;
}

static int require_sphere(mgimc_ctx* c, const char* fn) {
  IMC_REQUIRE(c && c->geometry == IMC_GEOM_SPHERE, kErrArg, "%s: needs a context from mgimc_create_sphere", fn);
  return kOk;
}

// Equilibrium census at t = 0 replacing the pool (MG_IMC1D.py:514-566).
int mgimc_sphere_equilibrium(mgimc_ctx* c, int64_t N, const double* Tr, double T_emit_floor) {
  int rc;
  if ((rc = require_sphere(c, "mgimc_sphere_equilibrium"))) return rc;
  IMC_REQUIRE(Tr && N >= 0, kErrArg, "mgimc_sphere_equilibrium: bad arguments");
  IMC_HIP(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  const int I = c->nx;
  IMC_HIP(hipMemcpyAsync(c->T.p, Tr, sizeof(double) * I, hipMemcpyHostToDevice, s));
  k_m1_eq_energy<<<launch_blocks(I, 256), 256, 0, s>>>(I, c->T.p, c->vol.p, T_emit_floor, c->Ecell.p);
  IMC_HIP(hipGetLastError());
  c->n = 0;
  c->have_material = false;  // T now holds Tr
  int64_t total = 0;
  if ((rc = m1_counts_offsets(c, c->Ecell.p, I, N, 1, &total))) return rc;
  if (N <= 0 || total <= 0) return kOk;
  if ((rc = c->A.grow(total, 0, s))) return rc;
  Key key = make_key(c->seed);
  const int64_t* offs = c->counts.p + I + 2;
  int g = launch_blocks(total, 256);
  if (c->precision == IMC_FP64)
    k_m1_sample_ic<double><<<g, 256, 0, s>>>(c->A.view<double>(), total, offs, I, c->e1.p, c->T.p, c->Ecell.p,
                                             c->G, c->edges.p, c->pcdf.p, key);
  else
    k_m1_sample_ic<float><<<g, 256, 0, s>>>(c->A.view<float>(), total, offs, I, c->e1.p, c->T.p, c->Ecell.p, c->G,
                                            c->edges.p, c->pcdf.p, key);
  IMC_HIP(hipGetLastError());
  IMC_HIP(hipStreamSynchronize(s));
  c->n = total;
  return kOk;
}

// Half-Lambertian source on the inner (side 0) or outer (side 1) sphere,
// appended to the pool.  E_out: a c T^4 / 4 dt 4 pi R^2 (MG_IMC1D.py:668).
int mgimc_sphere_face(mgimc_ctx* c, int side, int64_t n, double T, double dt, double* E_out) {
  int rc;
  if ((rc = require_sphere(c, "mgimc_sphere_face"))) return rc;
  IMC_REQUIRE((side == 0 || side == 1) && n >= 0, kErrArg, "mgimc_sphere_face: bad arguments");
  const int I = c->nx;
  double R = side == 0 ? c->h_e1[0] : c->h_e1[I];
  double t4 = (T * T) * (T * T);
  double E = kA * kC * t4 / 4.0 * dt * 4.0 * kPi * R * R;
  if (E_out) *E_out = E;
  if (n == 0) return kOk;
  IMC_HIP(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  if ((rc = c->A.grow(c->n + n, c->n, s))) return rc;
  Key key = make_key(c->seed);
  int g = launch_blocks(n, 256);
  if (c->precision == IMC_FP64)
    k_m1_sample_face<double><<<g, 256, 0, s>>>(c->A.view<double>(), c->n, n, side, E, T, dt, R, I, c->G, c->edges.p,
                                               c->pcdf.p, key, c->step);
  else
    k_m1_sample_face<float><<<g, 256, 0, s>>>(c->A.view<float>(), c->n, n, side, E, T, dt, R, I, c->G, c->edges.p,
                                              c->pcdf.p, key, c->step);
  IMC_HIP(hipGetLastError());
  c->n += n;
  return kOk;
}

// Volume source by (group, cell) bin appended to the pool.  kind
// IMC_SRC_EMISSION: the emission energies of the last mgimc_set_material,
// at least one particle per emitting bin; IMC_SRC_FIXED: QV_host (G, I)
// source energies this step, bins rounding to zero get none.  total_out:
// the energy of the particles added (0 when none are).
int mgimc_sphere_volume(mgimc_ctx* c, int kind, int64_t N, const double* QV_host, double dt, double* total_out) {
  int rc;
  if ((rc = require_sphere(c, "mgimc_sphere_volume"))) return rc;
  IMC_REQUIRE((kind == IMC_SRC_EMISSION || (kind == IMC_SRC_FIXED && QV_host)) && N >= 0, kErrArg,
              "mgimc_sphere_volume: bad arguments");
  IMC_REQUIRE(kind != IMC_SRC_EMISSION || c->have_material, kErrState,
              "mgimc_sphere_volume: call mgimc_set_material first");
  IMC_HIP(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  const int64_t nb = c->nb;
  const double* E = c->E.p;
  if (kind == IMC_SRC_FIXED) {
    IMC_HIP(hipMemcpyAsync(c->tmp.p, QV_host, sizeof(double) * nb, hipMemcpyHostToDevice, s));
    E = c->tmp.p;
  }
  if (total_out) *total_out = 0.0;
  if (N <= 0) return kOk;
  int64_t total = 0;
  if ((rc = m1_counts_offsets(c, E, nb, N, kind == IMC_SRC_EMISSION ? 1 : 0, &total))) return rc;
  if (total <= 0) return kOk;
  if ((rc = c->A.grow(c->n + total, c->n, s))) return rc;
  Key key = make_key(c->seed);
  const int64_t* offs = c->counts.p + nb + 2;
  uint32_t tag = kind == IMC_SRC_EMISSION ? kTagEmission : kTagSource;
  int g = launch_blocks(total, 256);
  if (c->precision == IMC_FP64)
    k_m1_sample_bins<double><<<g, 256, 0, s>>>(c->A.view<double>(), c->n, total, offs, nb, c->nx, c->e1.p, E, dt,
                                               key, c->step, tag);
  else
    k_m1_sample_bins<float><<<g, 256, 0, s>>>(c->A.view<float>(), c->n, total, offs, nb, c->nx, c->e1.p, E, dt, key,
                                              c->step, tag);
  IMC_HIP(hipGetLastError());
  // energy of the added particles: the bins that received particles
  if (total_out) {
    std::vector<double> e(nb);
    std::vector<int64_t> cnt(nb);
    IMC_HIP(hipMemcpyAsync(e.data(), E, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
    IMC_HIP(hipMemcpyAsync(cnt.data(), c->counts.p, sizeof(int64_t) * nb, hipMemcpyDeviceToHost, s));
    IMC_HIP(hipStreamSynchronize(s));
    double t = 0.0;
    for (int64_t b = 0; b < nb; ++b)
      if (cnt[b] > 0) t += e[b];
    *total_out = t;
  } else {
    IMC_HIP(hipStreamSynchronize(s));
  }
  c->n += total;
  return kOk;
}

// Transport the pool (MG_IMC1D.py:368-507).  weight_floor < 0: 1e-14 times
// the mean weight (canonical sum).  Outputs (nullable): dep, si (G, I); bl
// (G) weight escaping by group; stats[8] = events, boundary crossings,
// scatters, census, 0, floor kills, reflections, non-positive distances.
int mgimc_sphere_transport(mgimc_ctx* c, double dt, const int32_t* reflect, double weight_floor, double* dep_out,
                           double* si_out, double* bl_out, int64_t* stats_out) {
  int rc;
  if ((rc = require_sphere(c, "mgimc_sphere_transport"))) return rc;
  IMC_REQUIRE(reflect, kErrArg, "mgimc_sphere_transport: null reflect");
  IMC_REQUIRE(c->have_material, kErrState, "mgimc_sphere_transport: call mgimc_set_material first");
  IMC_HIP(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  int64_t n = c->n, nb = c->nb;
  int G = c->G;
  IMC_HIP(hipMemsetAsync(c->dep_raw.p, 0, sizeof(double) * (nb + G + 4) * c->tally_copies, s));
  IMC_HIP(hipMemsetAsync(c->path_raw.p, 0, sizeof(double) * nb * c->tally_copies, s));
  IMC_HIP(hipMemsetAsync(c->stats, 0, sizeof(unsigned long long) * 8, s));
  IMC_HIP(hipMemsetAsync(c->next, 0, sizeof(unsigned long long), s));
  if (weight_floor < 0.0) {
    if ((rc = c->sum_scr.reserve(canonical_sum_scratch(n) + 8))) return rc;
    if ((rc = canonical_sum((const double*)c->A.a[0], n, c->scal.p + 1, c->sum_scr.p, s))) return rc;
    k_m1_floor<<<1, 1, 0, s>>>(c->scal.p + 1, n, c->scal.p + 2);
    IMC_HIP(hipGetLastError());
  } else {
    IMC_HIP(hipMemcpyAsync(c->scal.p + 2, &weight_floor, sizeof(double), hipMemcpyHostToDevice, s));
  }
  int refl = (reflect[0] ? 1 : 0) | (reflect[1] ? 2 : 0);
  int block = 256, grid;
  const void* fn =
      c->precision == IMC_FP64 ? (const void*)k_m1_transport<double> : (const void*)k_m1_transport<float>;
  if (c->grid_override > 0) {
    grid = c->grid_override;
  } else {
    int per_cu = 0;
    IMC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, block, 0));
    grid = std::max(per_cu, 1) * c->n_cu;
  }
  grid = (int)std::max<int64_t>(std::min<int64_t>(grid, (n + block - 1) / block), 1);
  Key key = make_key(c->seed);
  IMC_HIP(hipEventRecord(c->ev0, s));
  if (n > 0) {
    if (c->precision == IMC_FP64) {
      MgArgs<double> A{c->A.view<double>(), n, c->e1p64.p, c->e2p64.p, c->nx, 1, G, c->sas64.p, dt, refl, 0,
                       c->scal.p + 2, key, c->step, c->next, c->batch, c->dep_raw.p, c->path_raw.p, nb,
                       c->tally_copies - 1, c->stats, pool_order_queue(n)};
      k_m1_transport<double><<<grid, block, 0, s>>>(A);
    } else {
      MgArgs<float> A{c->A.view<float>(), n, c->e1p32.p, c->e2p32.p, c->nx, 1, G, c->sas32.p, dt, refl, 0,
                      c->scal.p + 2, key, c->step, c->next, c->batch, c->dep_raw.p, c->path_raw.p, nb,
                      c->tally_copies - 1, c->stats, pool_order_queue(n)};
      k_m1_transport<float><<<grid, block, 0, s>>>(A);
    }
    IMC_HIP(hipGetLastError());
  }
  IMC_HIP(hipEventRecord(c->ev1, s));
  k_m1_finalize<<<launch_blocks(nb + G, 256), 256, 0, s>>>(nb, c->nx, G, c->tally_copies, c->dep_raw.p,
                                                            c->path_raw.p, c->sa.p, c->sas32.p,
                                                            c->precision == IMC_FP32_TRACK, c->vol.p, dt, c->dep.p,
                                                            c->si.p, c->bl.p);
  IMC_HIP(hipGetLastError());
  c->dep_valid = true;
  if (dep_out) IMC_HIP(hipMemcpyAsync(dep_out, c->dep.p, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
  if (si_out) IMC_HIP(hipMemcpyAsync(si_out, c->si.p, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
  if (bl_out) IMC_HIP(hipMemcpyAsync(bl_out, c->bl.p, sizeof(double) * G, hipMemcpyDeviceToHost, s));
  unsigned long long st[8];
  IMC_HIP(hipMemcpyAsync(st, c->stats, sizeof(st), hipMemcpyDeviceToHost, s));
  IMC_HIP(hipStreamSynchronize(s));
  IMC_HIP(hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1));
  if (stats_out)
    for (int a = 0; a < 8; ++a) stats_out[a] = (int64_t)st[a];
  return kOk;
}

// Post-transport census handling (MG_IMC1D.py:1147-1215): Nmax > 0 combs
// the live particles per (cell, group) bin to about Nmax; Nmax < 0 combs to
// |Nmax| only once at least |Nmax| particles live; otherwise dead particles
// are dropped and the rest kept.  radiation != 0: survivors of each bin are
// rescaled to the bin's pre-comb energy.  Times reset to 0.  Outputs
// (nullable): disc (G, I) = pre - post comb energy per bin (radiation mode:
// the energy of bins left without copies); n_alive, combed.
int mgimc_sphere_comb(mgimc_ctx* c, int64_t Nmax, int radiation, double* disc_out, int64_t* n_alive_out,
                      int* combed_out) {
  if (c) c->post_valid = false;  // the sphere comb keeps no post-comb histogram
  int rc;
  if ((rc = require_sphere(c, "mgimc_sphere_comb"))) return rc;
  IMC_HIP(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  int64_t n = c->n, nb = c->nb;
  const int I = c->nx;
  if ((rc = c->copies.reserve(n + 1)) || (rc = c->counts.reserve(n + 2)) ||
      (rc = c->scan_scr.reserve(scan_scratch(n) + 8)))
    return rc;
  Key key = make_key(c->seed);
  int g = launch_blocks(n, 256);
  // live particles (w > 0): pass-through copies
  if (n > 0) {
    if (c->precision == IMC_FP64)
      k_mg_copies<double><<<g, 256, 0, s>>>(c->A.view<double>(), n, I, 1, c->ew.p, 1, key, c->step, c->copies.p);
    else
      k_mg_copies<float><<<g, 256, 0, s>>>(c->A.view<float>(), n, I, 1, c->ew.p, 1, key, c->step, c->copies.p);
    IMC_HIP(hipGetLastError());
  }
  if ((rc = exclusive_scan_i32_to_i64(c->copies.p, n, c->counts.p, c->scan_scr.p, s))) return rc;
  int64_t alive = 0;
  IMC_HIP(hipMemcpyAsync(&alive, c->counts.p + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  IMC_HIP(hipStreamSynchronize(s));
  if (n_alive_out) *n_alive_out = alive;
  int64_t target = Nmax < 0 ? -Nmax : Nmax;
  bool comb = Nmax > 0 || (Nmax < 0 && alive >= target);
  if (combed_out) *combed_out = comb ? 1 : 0;
  if (comb) {
    if ((rc = mg_histogram(c, c->A, n, c->ecen.p))) return rc;
    if ((rc = c->sum_scr.reserve(canonical_sum_scratch(nb) + 8))) return rc;
    if ((rc = canonical_sum(c->ecen.p, nb, c->scal.p + 4, c->sum_scr.p, s))) return rc;
    double E = 0.0;
    IMC_HIP(hipMemcpyAsync(&E, c->scal.p + 4, sizeof(double), hipMemcpyDeviceToHost, s));
    IMC_HIP(hipStreamSynchronize(s));
    comb = E > 0.0;
    if (comb) {
      k_mg_ew<<<launch_blocks(nb, 256), 256, 0, s>>>(nb, c->ecen.p, c->scal.p + 4, target, c->ew.p);
      IMC_HIP(hipGetLastError());
      if (n > 0) {
        if (c->precision == IMC_FP64)
          k_mg_copies<double><<<g, 256, 0, s>>>(c->A.view<double>(), n, I, 1, c->ew.p, 0, key, c->step,
                                                c->copies.p);
        else
          k_mg_copies<float><<<g, 256, 0, s>>>(c->A.view<float>(), n, I, 1, c->ew.p, 0, key, c->step, c->copies.p);
        IMC_HIP(hipGetLastError());
      }
      if ((rc = exclusive_scan_i32_to_i64(c->copies.p, n, c->counts.p, c->scan_scr.p, s))) return rc;
    }
  }
  int64_t total = 0;
  IMC_HIP(hipMemcpyAsync(&total, c->counts.p + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  IMC_HIP(hipStreamSynchronize(s));
  if ((rc = comb_output_capacity(c, total, n))) return rc;
  if (n > 0) {
    if (c->precision == IMC_FP64)
      k_mg_comb_write<double><<<g, 256, 0, s>>>(c->A.view<double>(), c->B.view<double>(), n, I, 1, c->counts.p,
                                                c->ew.p, comb ? 0 : 1);
    else
      k_mg_comb_write<float><<<g, 256, 0, s>>>(c->A.view<float>(), c->B.view<float>(), n, I, 1, c->counts.p,
                                               c->ew.p, comb ? 0 : 1);
    IMC_HIP(hipGetLastError());
  }
  std::swap(c->A, c->B);
  c->n = total;
  if (comb && (radiation || disc_out)) {
    if ((rc = mg_histogram(c, c->A, total, c->post.p))) return rc;
    if (radiation) {
      if (total > 0) {
        if (c->precision == IMC_FP64)
          k_m1_rescale<double><<<launch_blocks(total, 256), 256, 0, s>>>(c->A.view<double>(), total, I, c->ecen.p,
                                                                         c->post.p);
        else
          k_m1_rescale<float><<<launch_blocks(total, 256), 256, 0, s>>>(c->A.view<float>(), total, I, c->ecen.p,
                                                                        c->post.p);
        IMC_HIP(hipGetLastError());
      }
      k_m1_lost<<<launch_blocks(nb, 256), 256, 0, s>>>(nb, c->ecen.p, c->post.p, c->tmp.p);
      IMC_HIP(hipGetLastError());
    } else {
      k_sub<<<launch_blocks(nb, 256), 256, 0, s>>>(nb, c->ecen.p, c->post.p, c->tmp.p);
      IMC_HIP(hipGetLastError());
    }
    if (disc_out) IMC_HIP(hipMemcpyAsync(disc_out, c->tmp.p, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
  } else if (disc_out) {
    for (int64_t b = 0; b < nb; ++b) disc_out[b] = 0.0;
  }
  IMC_HIP(hipStreamSynchronize(s));
  return kOk;
}

// Census diagnostics per (group, cell) bin: energy sum w and radial moment
// sum w mu (MG_IMC1D.py:1262-1276, before the c / V scaling); total_out: sum w.
int mgimc_sphere_census(mgimc_ctx* c, double* E_out, double* wmu_out, double* total_out) {
  int rc;
  if ((rc = require_sphere(c, "mgimc_sphere_census"))) return rc;
  IMC_HIP(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  int64_t n = c->n, nb = c->nb;
  if ((rc = mg_histogram(c, c->A, n, c->post.p))) return rc;
  if (E_out) IMC_HIP(hipMemcpyAsync(E_out, c->post.p, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
  if (wmu_out) {
    IMC_HIP(hipMemsetAsync(c->ecen_part.p, 0, sizeof(double) * nb * c->tally_copies, s));
    if (n > 0) {
      int g = std::min(launch_blocks(n, 256), c->n_cu * 4);
      if (c->precision == IMC_FP64)
        k_m1_flux<double><<<g, 256, 0, s>>>(c->A.view<double>(), n, c->nx, nb, c->tally_copies - 1, c->ecen_part.p);
      else
        k_m1_flux<float><<<g, 256, 0, s>>>(c->A.view<float>(), n, c->nx, nb, c->tally_copies - 1, c->ecen_part.p);
      IMC_HIP(hipGetLastError());
    }
    k_reduce_copies<double, double><<<launch_blocks(nb, 256), 256, 0, s>>>(nb, c->tally_copies, c->ecen_part.p,
                                                                           c->tmp.p);
    IMC_HIP(hipGetLastError());
    IMC_HIP(hipMemcpyAsync(wmu_out, c->tmp.p, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
  }
  if ((rc = tree_sum((const double*)c->A.a[0], n, c->scal.p + 5, c->tree_scr.p, s))) return rc;
  double tr = 0.0;
  IMC_HIP(hipMemcpyAsync(&tr, c->scal.p + 5, sizeof(double), hipMemcpyDeviceToHost, s));
  IMC_HIP(hipStreamSynchronize(s));
  if (total_out) *total_out = tr;
  return kOk;
}

}  // extern "C"

