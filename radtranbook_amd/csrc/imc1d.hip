// imc1d.hip -- batched gray 1-D (slab / sphere) Fleck-Cummings IMC on one
// MI355X.  One context holds P independent problems (P = 1 for the drop-in
// module); every step's transport pool is grouped by problem (census, then
// the step's source blocks), so the persistent tracker, the LDS-cached
// tallies and the comb each run once over all problems.  Philox index of a
// particle = (problem << 32) | index within its problem, so any problem of
// a batch is reproducible on its own.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/imc_hip.h"
#include "imc1d_cf_kernels.hpp"
#include "imc1d_kernels.hpp"
#include "imc_device.hpp"
#include "imc_util.hpp"

namespace imc {

template <typename R>
struct Pool1View {
  double* w;
  R* mu;
  R* t;
  R* x;
  int32_t* cell;  // cell index within the problem
  int32_t* prob;
};

struct Pool1Buf {
  double* w = nullptr;
  void* mu = nullptr;
  void* t = nullptr;
  void* x = nullptr;
  int32_t* cell = nullptr;
  int32_t* prob = nullptr;
  int64_t cap = 0;
  int rsz = 8;
  template <typename R>
  Pool1View<R> view() const {
    return Pool1View<R>{w, (R*)mu, (R*)t, (R*)x, cell, prob};
  }
  int reserve(int64_t n) {
    if (n <= cap) return kOk;
    int64_t nc = std::max<int64_t>(n, cap + cap / 4);
    release();
    IMC_HIP(hipMalloc(&w, 8 * (size_t)nc));
    IMC_HIP(hipMalloc(&mu, rsz * (size_t)nc));
    IMC_HIP(hipMalloc(&t, rsz * (size_t)nc));
    IMC_HIP(hipMalloc(&x, rsz * (size_t)nc));
    IMC_HIP(hipMalloc(&cell, 4 * (size_t)nc));
    IMC_HIP(hipMalloc(&prob, 4 * (size_t)nc));
    cap = nc;
    return kOk;
  }
  void release() {
    void* a[6] = {w, mu, t, x, cell, prob};
    for (void* p : a)
      if (p) (void)hipFree(p);
    w = nullptr; mu = t = x = nullptr; cell = prob = nullptr;
    cap = 0;
  }
};

// ---------------------------------------------------------------------------
// per-cell / per-problem kernels (no FMA contraction: counts must match the
// oracle bit for bit)
// ---------------------------------------------------------------------------
__global__ void k1_fleck(int64_t n, const int32_t* __restrict__ cprob, const double* __restrict__ T,
                         const double* __restrict__ sig, const double* __restrict__ cv,
                         const double* __restrict__ theta, const double* __restrict__ dt, double* __restrict__ sa,
                         double* __restrict__ ss, double2* __restrict__ sas64, float2* __restrict__ sas32,
                         int* __restrict__ bad) {
#pragma clang fp contract(off)
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  int p = cprob[c];
  double t = T[c];
  double beta = 4.0 * kA * (t * t * t) / cv[c];
  double f = 1.0 / (1.0 + theta[p] * beta * sig[c] * kC * dt[p]);
  if (!(f >= 0.0 && f <= 1.0)) atomicOr(bad, 1);  // IMC/IMC1D.py:936
  double s_s = sig[c] * (1.0 - f);
  double s_a = sig[c] * f;
  sa[c] = s_a;
  ss[c] = s_s;
  sas64[c] = make_double2(s_a, s_s);
  sas32[c] = make_float2((float)s_a, (float)s_s);
}

// Emission energy a c T^4 sigma_a dt V, zero below T_emit_floor (IMC/IMC1D.py:272-275).
__global__ void k1_emission_energy(int64_t n, const int32_t* __restrict__ cprob, const double* __restrict__ T,
                                   const double* __restrict__ sa, const double* __restrict__ vol,
                                   const double* __restrict__ dt, const double* __restrict__ tfloor,
                                   double* __restrict__ E) {
#pragma clang fp contract(off)
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  int p = cprob[c];
  double t = T[c] > 0.0 ? T[c] : 0.0;
  double t2 = t * t;
  double e = kA * kC * (t2 * t2) * sa[c] * dt[p] * vol[c];
  if (tfloor[p] > 0.0 && T[c] < tfloor[p]) e = 0.0;
  E[c] = e;
}

// Equilibrium energy a T^4 V (IMC/IMC1D.py:224-227).
__global__ void k1_equilibrium_energy(int64_t n, const int32_t* __restrict__ cprob, const double* __restrict__ Tr,
                                      const double* __restrict__ vol, const double* __restrict__ tfloor,
                                      double* __restrict__ E) {
#pragma clang fp contract(off)
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  int p = cprob[c];
  double t = Tr[c];
  double e = kA * ((t * t) * (t * t)) * vol[c];
  if (tfloor[p] > 0.0 && Tr[c] < tfloor[p]) e = 0.0;
  E[c] = e;
}

// Canonical sum of each problem's cells (chunks of 1024 left to right, then
// the chunk sums left to right); one workgroup per problem, each chunk staged
// through LDS and added in order by one lane.
__global__ __launch_bounds__(256) void k1_problem_sums(int P, const int64_t* __restrict__ coff,
                                                       const double* __restrict__ x, double* __restrict__ out) {
  __shared__ double v[1024];
  const int p = blockIdx.x;
  const int64_t lo = coff[p], hi = coff[p + 1];
  double tot = 0.0;
  for (int64_t c0 = lo; c0 < hi; c0 += 1024) {
    const int m = (int)((hi - c0) < 1024 ? (hi - c0) : 1024);
    __syncthreads();
    for (int i = threadIdx.x; i < m; i += 256) v[i] = x[c0 + i];
    __syncthreads();
    if (threadIdx.x == 0) {
      double s = 0.0;
#pragma unroll 16
      for (int i = 0; i < m; ++i) s += v[i];
      tot += s;
    }
  }
  if (threadIdx.x == 0) out[p] = tot;
}

// counts: mode 0 ceil(N E_c / E_p), zero when E_p <= 0 (IMC/IMC1D.py:277-281);
// mode 1 ceil(E_c / E_p * N), zero when N <= 0 or E_p <= 0 (IMC/IMC1D.py:229-232);
// Carter-Forest emission, none when E_p < 1e-30, else at least one per cell:
// mode 2 round(N (E_c / E_p)) (slab, IMC/IMC1D_CarterForest.py:504-505),
// mode 3 round(N E_c / E_p) (sphere, IMC/IMC1D_CarterForest.py:927); round
// half to even, as numpy.round
__global__ void k1_counts(int64_t n, const int32_t* __restrict__ cprob, const double* __restrict__ E,
                          const double* __restrict__ Ep, const int64_t* __restrict__ N, int mode,
                          int64_t* __restrict__ counts) {
#pragma clang fp contract(off)
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  int p = cprob[c];
  double et = Ep[p];
  int64_t k = 0;
  if (mode >= 2) {
    if (et >= 1e-30) {
      const double q = mode == 2 ? (double)N[p] * (E[c] / et) : (double)N[p] * E[c] / et;
      k = (int64_t)rint(q);
      if (k < 1) k = 1;
    }
  } else if (et > 0.0 && N[p] > 0) {
    if (mode == 0)
      k = (int64_t)ceil((double)N[p] * E[c] / et);
    else
      k = (int64_t)ceil(E[c] / et * (double)N[p]);
  }
  counts[c] = k;
}

// linear-in-T slab slopes (IMC/IMC1D.py:254-262) and linear-in-T^4 sphere
// slopes with the inner-edge coefficient (IMC/IMC1D.py:643-666)
__global__ void k1_slopes(int P, const int64_t* __restrict__ coff, const double* __restrict__ T,
                          const double2* __restrict__ ep, int sph, double* __restrict__ slope,
                          double* __restrict__ acoef) {
#pragma clang fp contract(off)
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= coff[P]) return;
  int lo = 0, hi = P;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (coff[mid + 1] > c)
      hi = mid;
    else
      lo = mid + 1;
  }
  const int64_t b = coff[lo];
  const int I = (int)(coff[lo + 1] - b);
  const int i = (int)(c - b);
  auto dxof = [&](int k) { return ep[b + k].y - ep[b + k].x; };
  if (!sph) {
    double dx = dxof(i);
    double s;
    if (I == 1)
      s = 0.0;
    else if (i == 0)
      s = (T[b + 1] - T[b]) / (dxof(1) + dx) * 2.0;
    else if (i == I - 1)
      s = (T[b + I - 1] - T[b + I - 2]) / (dx + dxof(I - 2)) * 2.0;
    else
      s = (T[b + i + 1] - T[b + i - 1]) / (dxof(i + 1) + dxof(i - 1)) * 2.0;
    double lv = T[c] - s * dx / 2.0, rv = T[c] + s * dx / 2.0;
    if (lv < 0.0 || rv < 0.0) s = 0.0;
    slope[c] = s;
    acoef[c] = 0.0;
  } else {
    auto t4 = [&](int k) { double t = T[b + k]; return t * t * t * t; };
    auto rm = [&](int k) { return 0.5 * (ep[b + k].x + ep[b + k].y); };
    double m;
    if (I == 1)
      m = 0.0;
    else if (i == 0)
      m = (t4(1) - t4(0)) / (rm(1) - rm(0));
    else if (i == I - 1)
      m = (t4(I - 1) - t4(I - 2)) / (rm(I - 1) - rm(I - 2));
    else
      m = (t4(i + 1) - t4(i - 1)) / (rm(i + 1) - rm(i - 1));
    double r0 = ep[c].x, r1 = ep[c].y, rmid = 0.5 * (r0 + r1);
    double a = t4(i) - m * (rmid - r0), bb = t4(i) + m * (r1 - rmid);
    if (a < 0.0 || bb < 0.0) m = 0.0;
    slope[c] = m;
    acoef[c] = t4(i) - m * (rmid - r0);
  }
}

__device__ double linear_density(double dx, double s, double T, double xi) {
#pragma clang fp contract(off)
  if (fabs(s) < 1e-6) return dx * xi;
  double smin = -2.0 * T / (dx * dx), smax = 2.0 * T / (dx * dx);
  if (s < smin) s = smin;
  if (s > smax) s = smax;
  double q = (s / (2.0 * T) - 1.0 / dx);
  return (-2.0 * T + s * dx + 2.0 * T * dx * sqrt(q * q + (2.0 * s * xi) / (T * dx))) / (2.0 * s);
}

__device__ double solve_sph_cdf(double r0, double r1, double a, double M, double xi) {
#pragma clang fp contract(off)
  double C = a / 3.0 * (r1 * r1 * r1 - r0 * r0 * r0) + M / 4.0 * (r1 * r1 * r1 * r1 - r0 * r0 * r0 * r0);
  if (C <= 0.0) return r0 + xi * (r1 - r0);
  double rhs = xi * C + (a / 3.0 * r0 * r0 * r0 + M / 4.0 * r0 * r0 * r0 * r0);
  double lo = r0, hi = r1, r = r0 + xi * (r1 - r0);
  for (int it = 0; it < 60; ++it) {
    double r3 = r * r * r, r4 = r3 * r;
    double g = a / 3.0 * r3 + M / 4.0 * r4 - rhs;
    if (fabs(g) < 1e-14 * (fabs(rhs) + 1e-30)) break;
    if (g < 0.0)
      lo = r;
    else
      hi = r;
    double dg = r * r * (a + M * r);
    double rn = (fabs(dg) > 1e-30) ? r - g / dg : 0.5 * (lo + hi);
    if (rn <= lo || rn >= hi) rn = 0.5 * (lo + hi);
    if (fabs(rn - r) < 1e-14 * (r1 - r0 + 1e-30)) {
      r = rn;
      break;
    }
    r = rn;
  }
  return r;
}

__device__ __forceinline__ int64_t bin_of(const int64_t* __restrict__ offs, int64_t n, int64_t k) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (offs[mid + 1] > k)
      hi = mid;
    else
      lo = mid + 1;
  }
  return lo;
}

// Volume sources (IMC/IMC1D.py:220-316, 535-741).  kind 0 emission (tilted,
// t ~ U(0,dt)), 1 equilibrium (uniform, t = 0), 2 fixed source (uniform; slab
// times U(0,1) as IMC/IMC1D.py:313), 3 Carter-Forest emission (tilted, the
// time from the truncated exponential of rate c sigma_a beta,
// IMC/IMC1D_CarterForest.py:443-549, 890-965).  Slots: 0 time, 1 position,
// 2 mu.
template <typename R>
__global__ void k1_sample_volume(Pool1View<R> D, int64_t base, int64_t total, const int64_t* __restrict__ offs,
                                 int64_t ncell, const int32_t* __restrict__ cprob, const int64_t* __restrict__ coff,
                                 const double2* __restrict__ ep, const double* __restrict__ QV,
                                 const double* __restrict__ T, const double* __restrict__ slope,
                                 const double* __restrict__ acoef, const double* __restrict__ dt, int kind, int sph,
                                 const double* __restrict__ cf_sa, const double* __restrict__ cf_beta, Key key,
                                 uint32_t step, uint32_t tag) {
#pragma clang fp contract(off)
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= total) return;
  int64_t c = bin_of(offs, ncell, k);
  int p = cprob[c];
  int64_t n = offs[c + 1] - offs[c];
  int64_t q = k - offs[coff[p]];  // index within the problem's block
  Ctr ctr = make_ctr(tag, ((uint64_t)p << 32) | (uint64_t)q, step);
  Pair b0 = block_d(ctr, key, 0), b1 = block_d(ctr, key, 1);
  double ut = b0.u0, ux = b0.u1, um = b1.u0;
  double r0 = ep[c].x, r1 = ep[c].y, dx = r1 - r0;
  double t = (kind == 1) ? 0.0 : ((kind == 2 && !sph) ? ut : dt[p] * ut);
  if (kind == 3) t = cf1_time(kC * cf_sa[c] * cf_beta[c], dt[p], ut);
  const bool tilted = kind == 0 || kind == 3;
  double x;
  if (!sph) {
    x = tilted ? linear_density(dx, slope[c], T[c], ux) + r0 : r0 + dx * ux;
  } else if (tilted && !(fabs(slope[c]) < 1e-30 || fabs(acoef[c]) < 1e-30)) {
    x = solve_sph_cdf(r0, r1, acoef[c], slope[c], ux);
  } else {
    x = pow(r0 * r0 * r0 + (r1 * r1 * r1 - r0 * r0 * r0) * ux, 1.0 / 3.0);
  }
  int64_t o = base + k;
  D.w[o] = QV[c] / (double)n;
  D.mu[o] = (R)(2.0 * um - 1.0);
  D.t[o] = (R)t;
  D.x[o] = (R)x;
  D.cell[o] = (int32_t)(c - coff[p]);
  D.prob[o] = p;
}

// Boundary sources (IMC/IMC1D.py:319-329, 744-763, 947-975).  side 0: left face,
// mu = sqrt(xi); side 1: right face, mu = -sqrt(xi).  Slots: 0 mu, 1 time.
template <typename R>
__global__ void k1_sample_boundary(Pool1View<R> D, int64_t base, int64_t total, const int64_t* __restrict__ boff,
                                   int P, int side, int sph, const double* __restrict__ E,
                                   const double* __restrict__ dt, const int64_t* __restrict__ coff,
                                   const double2* __restrict__ ep, Key key, uint32_t step) {
#pragma clang fp contract(off)
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= total) return;
  int p = (int)bin_of(boff, P, k);
  int64_t q = k - boff[p];
  int64_t n = boff[p + 1] - boff[p];
  Ctr ctr = make_ctr(kTagBoundary0 + side, ((uint64_t)p << 32) | (uint64_t)q, step);
  Pair b0 = block_d(ctr, key, 0);
  double m = sqrt(b0.u0);
  int64_t o = base + k;
  int I = (int)(coff[p + 1] - coff[p]);
  D.w[o] = E[p] / (double)n;
  D.t[o] = (R)(dt[p] * b0.u1);
  if (side == 0) {
    D.mu[o] = (R)m;
    D.x[o] = (R)(sph ? ep[coff[p]].x : 1e-8);
    D.cell[o] = 0;
  } else {
    D.mu[o] = (R)(-m);
    D.x[o] = (R)ep[coff[p + 1] - 1].y;
    D.cell[o] = I - 1;
  }
  D.prob[o] = p;
}

// Gather segments into the grouped transport pool: destination range
// [dst[s], dst[s+1]) comes from src_buf[s] starting at src[s].  One thread
// per destination; two threads of the workgroup find the segments of its
// first and last destinations by binary search, and every thread walks
// forward from the first (a workgroup spans few segments), instead of each
// running the whole search through global memory.
template <typename R>
__global__ __launch_bounds__(256) void k1_gather(Pool1View<R> A, Pool1View<R> N, Pool1View<R> Dst, int nseg,
                                                 const int64_t* __restrict__ dst, const int64_t* __restrict__ src,
                                                 const int32_t* __restrict__ from_new, int64_t total) {
  __shared__ int64_t sb;
  const int64_t k0 = (int64_t)blockIdx.x * blockDim.x;
  const int64_t k = k0 + threadIdx.x;
  if (threadIdx.x == 0) sb = bin_of(dst, nseg, k0);
  __syncthreads();
  if (k >= total) return;
  // the first segment whose end passes k (bin_of's answer): the walk ends by
  // the last destination's segment, since k < total = dst[nseg]
  int64_t s = sb;
  while (dst[s + 1] <= k) ++s;
  int64_t i = src[s] + (k - dst[s]);
  const Pool1View<R>& S = from_new[s] ? N : A;
  Dst.w[k] = S.w[i];
  Dst.mu[k] = S.mu[i];
  Dst.t[k] = S.t[i];
  Dst.x[k] = S.x[i];
  Dst.cell[k] = S.cell[i];
  Dst.prob[k] = S.prob[i];
}

// Deterministic per-problem sums over grouped pool segments: one workgroup
// of kSegThreads per problem, four independent accumulators per thread (a
// single dependent add chain per thread left the largest segments, ~2e6
// particles, waiting on one load at a time), combined in a fixed order.
constexpr int kSegThreads = 1024;
__global__ __launch_bounds__(kSegThreads) void k1_segment_sums(const double* __restrict__ w,
                                                               const int64_t* __restrict__ poff,
                                                               double* __restrict__ out) {
  __shared__ double ws[kSegThreads / 64];
  const int p = blockIdx.x;
  const int64_t lo = poff[p], hi = poff[p + 1];
  const int64_t bd = kSegThreads;
  double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
  int64_t i = lo + threadIdx.x;
  for (; i + 3 * bd < hi; i += 4 * bd) {
    v0 += w[i];
    v1 += w[i + bd];
    v2 += w[i + 2 * bd];
    v3 += w[i + 3 * bd];
  }
  for (; i < hi; i += bd) v0 += w[i];
  double v = (v0 + v1) + (v2 + v3);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int q = 0; q < kSegThreads / 64; ++q) t += ws[q];
    out[p] = t;
  }
}

__global__ void k1_floor(int P, const double* __restrict__ sums, const int64_t* __restrict__ poff,
                         double* __restrict__ floor_out) {
#pragma clang fp contract(off)
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  int64_t n = poff[p + 1] - poff[p];
  floor_out[p] = 1e-10 * sums[p] / (double)(n > 1 ? n : 1);
}

// out[p] = src[idx[p]] for p = 0..P (offsets at problem boundaries)
__global__ void k1_take(int P, const int64_t* __restrict__ idx, const int64_t* __restrict__ src,
                        int64_t* __restrict__ out) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p <= P) out[p] = src[idx[p]];
}

// ---------------------------------------------------------------------------
// transport (IMC/IMC1D.py:133-200 slab, 439-532 sphere)
// ---------------------------------------------------------------------------
template <typename R>
struct Transport1Args {
  using R2 = typename Vec2<R>::type;
  Pool1View<R> P;
  int64_t n;
  const int64_t* poff;   // pool offset of each problem (P+1)
  const int64_t* coff;   // cell offset of each problem (P+1)
  const R2* ep;          // per cell: FP64 (r0, r1); FP32 (r0, r1 - r0)
  const R2* sas;         // (sigma_a, sigma_s) per cell
  const double* dt;      // per problem
  const double* wfloor;  // per problem
  const int32_t* refl;   // 2 per problem
  Key key;
  uint32_t step;
  unsigned long long* next;
  int batch;
  double* dep_raw;  // copies x (ncell + 2P): cells, then (leak, floor leak) per problem
  double* path_raw; // copies x ncell: w*s where sigma_a <= 1e-10
  int64_t ncell;
  int nprob;
  int copy_mask;
  unsigned long long* stats;  // [8]
  QueueMap Q;
  int refill_min;  // a wave with active lanes refills once this many lanes are idle (0 or 1: any)
};

template <typename R>
__device__ __forceinline__ R vol_of(int sph, R r0, R r1) {
  return sph ? (R)(4.0 / 3.0) * (R)kPi * (r1 * r1 * r1 - r0 * r0 * r0) : (r1 - r0);
}

template <typename R, int SPH>
__global__ __launch_bounds__(256) void k1_transport(Transport1Args<R> A) {
  using R2 = typename Vec2<R>::type;
  const int64_t tally_n = A.ncell + 2 * (int64_t)A.nprob;
  const int copy = blockIdx.x & A.copy_mask;
  double* const dep_t = A.dep_raw + (int64_t)copy * tally_n;
  double* const path_t = A.path_raw + (int64_t)copy * A.ncell;
  __shared__ int lkey[kLdsSlots];
  __shared__ double lval[kLdsSlots];
  lds_tally_init(lkey, lval);
  __syncthreads();
  WaveQueue wq;
  bool active = false;
  int64_t k = 0;
  double w = 0.0;
  R mu = 0, dcen = 0;
  // FP64: the position.  FP32: the offset from the cell's inner edge
  // (kAtUpperFace: on the outer face), or, while `fresh`, the position read
  // from the pool; q holds the position to write back.
  R x = 0, q = 0;
  bool fresh = false;
  int cell = 0, p = 0, I = 0;
  int64_t cb = 0;
  double wfl = 0.0;
  uint32_t ev = 0;
  double acc_dep = 0.0, acc_path = 0.0;
  int acc_cell = -1;
  uint32_t st_ev = 0, st_bnd = 0, st_sc = 0, st_cen = 0, st_floor = 0, st_refl = 0, st_bad = 0;
  while (true) {
    uint64_t need = __ballot(!active);
    if (need && (__popcll(need) >= A.refill_min || need == ~0ull)) {
      const int64_t kk = wave_take(wq, !active, need, A.next, A.batch, A.Q, A.n);
      if (kk >= 0) {
        k = kk;
        w = A.P.w[k];
        mu = A.P.mu[k];
        x = A.P.x[k];
        q = x;
        fresh = true;
        cell = A.P.cell[k];
        p = A.P.prob[k];
        cb = A.coff[p];
        I = (int)(A.coff[p + 1] - cb);
        wfl = A.wfloor[p];
        dcen = (R)((A.dt[p] - (double)A.P.t[k]) * kC);
        ev = 0;
        active = dcen > (R)0;
      }
    }
    if (__ballot(active) == 0) {
      if (wq.exhausted) break;
      continue;
    }
    if (!active) continue;
    bool finish = false;
    ++st_ev;
    const int loc = cell;
    const int64_t c = cb + loc;
    const R2 edg = A.ep[c];
    R2 mat = A.sas[c];
    Move1D<R> o;
    Ctr ctr = make_ctr(kTagTransport, ((uint64_t)p << 32) | (uint64_t)(k - A.poff[p]), A.step);
    if constexpr (kF32<R>) {
      x = fresh ? fminf(fmaxf(x - edg.x, 0.0f), edg.y) : resolve_offset(x, edg.y);
      fresh = false;
    }
    move_1d<SPH>(ctr, A.key, ev, w, mu, x, edg.x, edg.y, mat.x, mat.y, dcen, o);
    ++ev;
    if (o.bad) {
      ++st_bad;
      finish = true;
    } else {
      if ((int)c != acc_cell) {
        if (acc_cell >= 0) {
          if (!lds_tally_add(lkey, lval, acc_cell, acc_dep)) atomic_add_f64(dep_t + acc_cell, acc_dep);
          if (acc_path != 0.0) atomic_add_f64(path_t + acc_cell, acc_path);
        }
        acc_cell = (int)c;
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
      if (cell == I) {
        if (A.refl[2 * p + 1]) {
          mu = -mu; cell -= 1; ++st_refl;
          enter_upper(x);
        } else {
          dcen = 0;
        }
      } else if (cell == -1) {
        if (A.refl[2 * p]) {
          mu = -mu; cell += 1; ++st_refl;
          enter_lower(x);
        } else {
          dcen = 0;
        }
      }
      if (w < wfl) {
        ++st_floor;
        if (0 < cell && cell < I - 1) {
          // deposit the remainder in the particle's cell (IMC/IMC1D.py:188-195)
          int64_t c2 = cb + cell;
          if (!lds_tally_add(lkey, lval, (int)c2, w)) atomic_add_f64(dep_t + c2, w);
        } else {
          int kf = (int)(A.ncell + 2 * p + 1);
          if (!lds_tally_add(lkey, lval, kf, w)) atomic_add_f64(dep_t + kf, w);
        }
        w = 0.0;
        dcen = 0;
      }
      if (!(dcen > (R)0)) finish = true;
    }
    if (finish) {
      if (cell < 0 || cell >= I) {  // escaped: boundary loss (IMC/IMC1D.py:1038-1057)
        int kl = (int)(A.ncell + 2 * p);
        if (w != 0.0 && !lds_tally_add(lkey, lval, kl, w)) atomic_add_f64(dep_t + kl, w);
        w = 0.0;
      }
      A.P.w[k] = w;
      A.P.mu[k] = mu;
      A.P.x[k] = kF32<R> ? q : x;
      A.P.cell[k] = cell;
      active = false;
    }
  }
  if (acc_cell >= 0) {
    if (!lds_tally_add(lkey, lval, acc_cell, acc_dep)) atomic_add_f64(dep_t + acc_cell, acc_dep);
    if (acc_path != 0.0) atomic_add_f64(path_t + acc_cell, acc_path);
  }
  __syncthreads();
  lds_tally_flush(lkey, lval, dep_t);
  unsigned long long s0 = wave_sum<unsigned long long>(st_ev);
  unsigned long long s1 = wave_sum<unsigned long long>(st_bnd);
  unsigned long long s2 = wave_sum<unsigned long long>(st_sc);
  unsigned long long s3 = wave_sum<unsigned long long>(st_cen);
  unsigned long long s4 = wave_sum<unsigned long long>(st_floor);
  unsigned long long s5 = wave_sum<unsigned long long>(st_refl);
  unsigned long long s6 = wave_sum<unsigned long long>(st_bad);
  if (__lane_id() == 0) {
    atomicAdd(A.stats + 0, s0);
    atomicAdd(A.stats + 1, s1);
    atomicAdd(A.stats + 2, s2);
    atomicAdd(A.stats + 3, s3);
    atomicAdd(A.stats + 4, s4);
    atomicAdd(A.stats + 5, s5);
    atomicAdd(A.stats + 6, s6);
  }
}

// dep = sum(dep)/V, si = sum(dep)/(sigma_a V dt) where sigma_a > 1e-10, else
// sum(w s)/(V dt) (IMC/IMC1D.py:117-121); leaks per problem.
__global__ void k1_finalize(int64_t ncell, int P, int copies, const int32_t* __restrict__ cprob,
                            const double* __restrict__ dep_raw, const double* __restrict__ path_raw,
                            const double* __restrict__ sa64, const float2* __restrict__ sas32, int f32,
                            const double* __restrict__ vol, const double* __restrict__ dt, double* __restrict__ dep,
                            double* __restrict__ si, double* __restrict__ leaks) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t tally_n = ncell + 2 * (int64_t)P;
  if (c < ncell) {
    double a = 0.0, b = 0.0;
    for (int k = 0; k < copies; ++k) {
      a += dep_raw[(int64_t)k * tally_n + c];
      b += path_raw[(int64_t)k * ncell + c];
    }
    double sa = f32 ? (double)sas32[c].x : sa64[c];
    bool thick = f32 ? (sas32[c].x > 1e-10f) : (sa64[c] > 1e-10);
    double d = dt[cprob[c]];
    dep[c] = a / vol[c];
    si[c] = thick ? a / sa / vol[c] / d : b / vol[c] / d;
  } else if (c < tally_n) {
    double a = 0.0;
    for (int k = 0; k < copies; ++k) a += dep_raw[(int64_t)k * tally_n + c];
    leaks[c - ncell] = a;
  }
}

// ---------------------------------------------------------------------------
// Carter-Forest transport (IMC/IMC1D_CarterForest.py:210-332 slab, 723-819
// sphere), FP64.  Per history: while the distance to census is positive, count
// the event (a positive event cap ends the history when reached), reflect or
// drop an out-of-range index, move, tally, stop on capture, then reflect or
// escape at a domain face.  A history whose weight is at or below the floor
// is not moved and stays in census (counted as a floor kill; the reference
// drops its weight unbooked).  Escaped weight is the problem's boundary loss.
// dep_raw: captured energy per cell then (leak, unused) per problem; path_raw:
// sum of w s per cell.  stats: events, boundary, absorb-continue, census,
// capture, floor, reflections, cap hits.
// ---------------------------------------------------------------------------
struct Cf1Args {
  Pool1View<double> P;
  int64_t n;
  const int64_t* poff;
  const int64_t* coff;
  const double2* ep;
  const double* sa;
  const double* beta;
  const double* dt;
  const double* wfloor;
  const int32_t* refl;
  int64_t max_events;  // <= 0: no cap
  double fastpath;
  Key key;
  uint32_t step;
  unsigned long long* next;
  int batch;
  double* dep_raw;
  double* path_raw;
  int64_t ncell;
  int nprob;
  int copy_mask;
  unsigned long long* stats;
  QueueMap Q;
};

template <int SPH>
__global__ __launch_bounds__(256) void k1_transport_cf(Cf1Args A) {
#pragma clang fp contract(off)
  const int64_t tally_n = A.ncell + 2 * (int64_t)A.nprob;
  const int copy = blockIdx.x & A.copy_mask;
  double* const dep_t = A.dep_raw + (int64_t)copy * tally_n;
  double* const path_t = A.path_raw + (int64_t)copy * A.ncell;
  __shared__ int lkey[kLdsSlots];
  __shared__ double lval[kLdsSlots];
  lds_tally_init(lkey, lval);
  __syncthreads();
  // keys [0, tally_n): captured energy and leaks; [tally_n, tally_n + ncell): path
  auto add = [&](int64_t key, double v) {
    if (!lds_tally_add(lkey, lval, (int)key, v))
      atomic_add_f64(key < tally_n ? dep_t + key : path_t + (key - tally_n), v);
  };
  WaveQueue wq;
  bool active = false;
  int64_t k = 0, cb = 0;
  double w = 0.0, mu = 0.0, x = 0.0, t = 0.0, dcen = 0.0;
  int cell = 0, p = 0, I = 0;
  uint32_t ev = 0;
  double acc_path = 0.0;
  int64_t acc_cell = -1;
  uint32_t st_ev = 0, st_bnd = 0, st_abs = 0, st_cen = 0, st_cap = 0, st_floor = 0, st_refl = 0, st_hit = 0;
  while (true) {
    uint64_t need = __ballot(!active);
    if (need) {
      const int64_t kk = wave_take(wq, !active, need, A.next, A.batch, A.Q, A.n);
      if (kk >= 0) {
        k = kk;
        w = A.P.w[k];
        mu = A.P.mu[k];
        x = A.P.x[k];
        t = A.P.t[k];
        cell = A.P.cell[k];
        p = A.P.prob[k];
        cb = A.coff[p];
        I = (int)(A.coff[p + 1] - cb);
        dcen = (A.dt[p] - t) * kC;
        ev = 0;
        if (w > A.wfloor[p]) {
          active = dcen > 0.0;
        } else {
          st_floor += w > 0.0;
          active = false;
        }
      }
    }
    if (__ballot(active) == 0) {
      if (wq.exhausted) break;
      continue;
    }
    if (!active) continue;
    bool finish = false;
    ++ev;
    ++st_ev;
    if (A.max_events > 0 && (int64_t)ev >= A.max_events) {
      ++st_hit;
      dcen = 0.0;
      finish = true;
    } else {
      if (cell < 0) {
        if (A.refl[2 * p]) { cell = 0; mu = -mu; } else finish = true;
      } else if (cell >= I) {
        if (A.refl[2 * p + 1]) { cell = I - 1; mu = -mu; } else finish = true;
      }
      if (!finish) {
        const int64_t c = cb + cell;
        const double2 e = A.ep[c];
        Ctr ctr = make_ctr(kTagTransport, ((uint64_t)p << 32) | (uint64_t)(k - A.poff[p]), A.step);
        Cf1Out o;
        if (SPH)
          cf1_move_sph(ctr, A.key, ev - 1u, w, mu, x, t, e.x, e.y, A.sa[c], A.beta[c], dcen, A.fastpath, o);
        else
          cf1_move_slab(ctr, A.key, ev - 1u, w, mu, x, t, e.x, e.y, A.sa[c], A.beta[c], dcen, A.fastpath, o);
        if (o.dep_w != 0.0) add(c, o.dep_w);
        if (c != acc_cell) {
          if (acc_cell >= 0 && acc_path != 0.0) add(tally_n + acc_cell, acc_path);
          acc_cell = c;
          acc_path = 0.0;
        }
        acc_path += o.path;
        w = o.w;
        mu = o.mu;
        x = o.x;
        t = o.t;
        dcen = o.dcen;
        cell += o.newloc;
        st_bnd += o.code == kCf1Boundary;
        st_abs += o.code == kCf1Absorb;
        st_cen += o.code == kCf1Census;
        st_cap += o.code == kCf1Capture;
        if (o.code == kCf1Capture) {
          finish = true;
        } else if (cell == I) {
          if (A.refl[2 * p + 1]) { mu = -mu; cell -= 1; ++st_refl; } else finish = true;
        } else if (cell == -1) {
          if (A.refl[2 * p]) { mu = -mu; cell += 1; ++st_refl; } else finish = true;
        }
        if (!(dcen > 0.0)) finish = true;
      }
    }
    if (finish) {
      if (cell < 0 || cell >= I) {  // escaped: boundary loss of the problem
        if (w != 0.0) add(A.ncell + 2 * p, w);
        w = 0.0;
      }
      A.P.w[k] = w;
      A.P.mu[k] = mu;
      A.P.x[k] = x;
      A.P.t[k] = t;
      A.P.cell[k] = cell;
      active = false;
    }
  }
  if (acc_cell >= 0 && acc_path != 0.0) add(tally_n + acc_cell, acc_path);
  __syncthreads();
  for (int s = threadIdx.x; s < kLdsSlots; s += blockDim.x) {
    const int kk = lkey[s];
    if (kk >= 0) {
      const double v = lval[s];
      if (v != 0.0) atomic_add_f64(kk < tally_n ? dep_t + kk : path_t + (kk - tally_n), v);
    }
  }
  unsigned long long s0 = wave_sum<unsigned long long>(st_ev);
  unsigned long long s1 = wave_sum<unsigned long long>(st_bnd);
  unsigned long long s2 = wave_sum<unsigned long long>(st_abs);
  unsigned long long s3 = wave_sum<unsigned long long>(st_cen);
  unsigned long long s4 = wave_sum<unsigned long long>(st_cap);
  unsigned long long s5 = wave_sum<unsigned long long>(st_floor);
  unsigned long long s6 = wave_sum<unsigned long long>(st_refl);
  unsigned long long s7 = wave_sum<unsigned long long>(st_hit);
  if (__lane_id() == 0) {
    atomicAdd(A.stats + 0, s0);
    atomicAdd(A.stats + 1, s1);
    atomicAdd(A.stats + 2, s2);
    atomicAdd(A.stats + 3, s3);
    atomicAdd(A.stats + 4, s4);
    atomicAdd(A.stats + 5, s5);
    atomicAdd(A.stats + 6, s6);
    atomicAdd(A.stats + 7, s7);
  }
}

// dep = captured / V, si = sum(w s) / (V dt) per cell (the reference's
// per-segment w s / V / dt, IMC/IMC1D_CarterForest.py:125, 282-283); leaks
// per problem.
__global__ void k1_cf_finalize(int64_t ncell, int P, int copies, const int32_t* __restrict__ cprob,
                               const double* __restrict__ dep_raw, const double* __restrict__ path_raw,
                               const double* __restrict__ vol, const double* __restrict__ dt,
                               double* __restrict__ dep, double* __restrict__ si, double* __restrict__ leaks) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t tally_n = ncell + 2 * (int64_t)P;
  if (c < ncell) {
    double a = 0.0, b = 0.0;
    for (int k = 0; k < copies; ++k) {
      a += dep_raw[(int64_t)k * tally_n + c];
      b += path_raw[(int64_t)k * ncell + c];
    }
    dep[c] = a / vol[c];
    si[c] = b / vol[c] / dt[cprob[c]];
  } else if (c < tally_n) {
    double a = 0.0;
    for (int k = 0; k < copies; ++k) a += dep_raw[(int64_t)k * tally_n + c];
    leaks[c - ncell] = a;
  }
}

// ---------------------------------------------------------------------------
// comb (IMC/IMC1D.py:770-828), per problem
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(256) void k1_comb_ecen(Pool1View<R> P, int64_t n, const int64_t* __restrict__ coff,
                                                    int64_t ncell, int copy_mask, double* __restrict__ part) {
  __shared__ int lkey[kLdsSlots];
  __shared__ double lval[kLdsSlots];
  lds_tally_init(lkey, lval);
  __syncthreads();
  double* const dst = part + (int64_t)(blockIdx.x & copy_mask) * ncell;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t trips = (n + stride - 1) / stride;
  for (int64_t t = 0; t < trips; ++t) {
    int64_t k = t * stride + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int key = -1;
    double v = 0.0;
    if (k < n) {
      int p = P.prob[k], cl = P.cell[k];
      int I = (int)(coff[p + 1] - coff[p]);
      if (P.w[k] > 0.0 && cl >= 0 && cl < I) {
        key = (int)(coff[p] + cl);
        v = P.w[k];
      }
    }
    bool head = wave_run_sum(key, v);
    if (head && key >= 0 && !lds_tally_add(lkey, lval, key, v)) atomic_add_f64(dst + key, v);
  }
  __syncthreads();
  lds_tally_flush(lkey, lval, dst);
}

__global__ void k1_comb_ew(int64_t ncell, const int32_t* __restrict__ cprob, const double* __restrict__ ecen,
                           const double* __restrict__ Ep, const int64_t* __restrict__ Nmax, double* __restrict__ ew) {
#pragma clang fp contract(off)
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  int p = cprob[c];
  double e = ecen[c], E = Ep[p];
  if (e > 0.0 && E > 0.0) {
    double d = rint((double)Nmax[p] * e / E);
    if (d < 1.0) d = 1.0;
    ew[c] = e / d;
  } else {
    ew[c] = 0.0;
  }
}

template <typename R>
__global__ void k1_comb_copies(Pool1View<R> P, int64_t n, const int64_t* __restrict__ poff,
                               const int64_t* __restrict__ coff, const double* __restrict__ ew,
                               const double* __restrict__ Ep, Key key, uint32_t step, int32_t* __restrict__ copies) {
#pragma clang fp contract(off)
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  int p = P.prob[k], cl = P.cell[k];
  int I = (int)(coff[p + 1] - coff[p]);
  double w = P.w[k];
  int32_t m = 0;
  if (w > 0.0 && cl >= 0 && cl < I) {
    if (!(Ep[p] > 0.0)) {
      m = 1;  // nothing to comb in this problem: survivors keep their weights
    } else {
      double et = ew[coff[p] + cl];
      if (et > 0.0) {
        double u = slot_d(make_ctr(kTagComb, ((uint64_t)p << 32) | (uint64_t)(k - poff[p]), step), key, 0);
        m = (int32_t)(w / et + u);
      }
    }
  }
  copies[k] = m;
}

template <typename R>
__global__ __launch_bounds__(256) void k1_comb_write(Pool1View<R> S, Pool1View<R> D, int64_t n,
                                                     const int64_t* __restrict__ coff,
                                                     const int64_t* __restrict__ offs, const double* __restrict__ ew,
                                                     const double* __restrict__ Ep, int copy_mask,
                                                     unsigned int* __restrict__ count_part, int64_t ncell) {
  __shared__ int64_t so[257];
  const int t = threadIdx.x;
  const int64_t k0 = (int64_t)blockIdx.x * 256;
  if (count_part) {
    const int64_t k = k0 + t;
    int key = -1;
    unsigned int cnt = 0;
    if (k < n) {
      int m = (int)(offs[k + 1] - offs[k]);
      if (m > 0) {
        key = (int)(coff[S.prob[k]] + S.cell[k]);
        cnt = (unsigned int)m;
      }
    }
    bool head = wave_run_sum(key, cnt);
    if (head && key >= 0) atomicAdd(count_part + (int64_t)(blockIdx.x & copy_mask) * ncell + key, cnt);
  }
  const int cnt = (n - k0 < 256) ? (int)(n - k0) : 256;
  if (t < cnt) so[t] = offs[k0 + t];
  if (t == 0) so[cnt] = offs[k0 + cnt];
  __syncthreads();
  const int64_t lo = so[0], hi = so[cnt];
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
    const int p = S.prob[src];
    const int cl = S.cell[src];
    D.w[o] = (Ep[p] > 0.0) ? ew[coff[p] + cl] : S.w[src];
    D.mu[o] = S.mu[src];
    D.x[o] = S.x[src];
    D.t[o] = (R)0;  // times reset after the comb (IMC/IMC1D.py:1068)
    D.cell[o] = cl;
    D.prob[o] = p;
  }
}

__global__ void k1_comb_disc(int64_t ncell, const int32_t* __restrict__ cprob, const double* __restrict__ ecen,
                             const double* __restrict__ ew, const unsigned long long* __restrict__ cnt,
                             const double* __restrict__ Ep, double* __restrict__ disc) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  disc[c] = (Ep[cprob[c]] > 0.0) ? ecen[c] - ew[c] * (double)cnt[c] : 0.0;
}

template <typename R>
__global__ void k1_put(Pool1View<R> P, int64_t n, const double* __restrict__ w, const double* __restrict__ mu,
                       const double* __restrict__ t, const double* __restrict__ x, const int64_t* __restrict__ cell,
                       const int32_t* __restrict__ prob) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  P.w[k] = w[k];
  P.mu[k] = (R)mu[k];
  P.t[k] = (R)t[k];
  P.x[k] = (R)x[k];
  P.cell[k] = (int32_t)cell[k];
  P.prob[k] = prob[k];
}

template <typename R>
__global__ void k1_get(Pool1View<R> P, int64_t n, double* __restrict__ w, double* __restrict__ mu,
                       double* __restrict__ t, double* __restrict__ x, int64_t* __restrict__ cell,
                       int32_t* __restrict__ prob) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  w[k] = P.w[k];
  mu[k] = (double)P.mu[k];
  t[k] = (double)P.t[k];
  x[k] = (double)P.x[k];
  cell[k] = P.cell[k];
  prob[k] = P.prob[k];
}

// Expected segment count of the histories of chunk c (particles c << shift
// onward), from its first particle: path min(distance to census, 23 /
// sigma_a) times the event rate min(sigma_s, 1e4 / dx) + 2 / dx (the
// tracker's opacity cap; dx the cell width).  Chunks expected to run at least
// kLongHistory1 segments get keys 0.. (longest first, quarter octaves); all
// others share the last key and keep close to pool order.  Also leaves the
// per-workgroup key histograms.
constexpr double kLongHistory1 = 100.0;
template <typename R>
__global__ __launch_bounds__(kChunkBlock) void k1_chunk_key(Pool1View<R> P, int64_t n, int shift,
                                                            const int64_t* __restrict__ coff,
                                                            const double2* __restrict__ ep,
                                                            const double* __restrict__ sa,
                                                            const double* __restrict__ ss,
                                                            const double* __restrict__ dt, uint8_t* __restrict__ key,
                                                            int32_t* __restrict__ hist) {
  __shared__ int h[kChunkKeys];
  chunk_hist_zero(h);
  const int64_t nchunk = (n + (1ll << shift) - 1) >> shift;
  for (int64_t c = blockIdx.x * (int64_t)kChunkBlock + threadIdx.x; c < nchunk;
       c += (int64_t)gridDim.x * kChunkBlock) {
    const int64_t k = c << shift;
    const int p = P.prob[k], cell = P.cell[k];
    int kk = kChunkKeys - 1;
    const int64_t cb = coff[p];
    if (cell >= 0 && cell < (int)(coff[p + 1] - cb)) {
      const double dc = (dt[p] - (double)P.t[k]) * kC;
      if (dc > 0.0) {
        const int64_t ci = cb + cell;
        const double dx = ep[ci].y - ep[ci].x;
        const double len = fmin(dc, 23.0 / fmax(sa[ci], 1e-30));
        const double segs = len * (fmin(ss[ci], 1e4 / dx) + 2.0 / dx);
        const double q = 4.0 * log2(1.0 + segs);  // NaN and inf handled before the conversion
        const int b = q >= (double)(kChunkKeys - 1) ? kChunkKeys - 1 : (q > 0.0 ? (int)q : 0);
        kk = segs >= kLongHistory1 ? kChunkKeys - 1 - b : kChunkKeys - 1;
      }
    }
    key[c] = (uint8_t)kk;
    atomicAdd(&h[kk], 1);
  }
  chunk_hist_publish(h, hist);
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
struct Imc1dCtx {
  int sph = 0, P = 0, precision = 0, device = 0, n_cu = 256;
  uint64_t seed = 1;
  uint32_t step = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  float last_ms = 0.0f;
  int64_t ncell = 0;
  std::vector<int64_t> h_coff;
  std::vector<int64_t> h_ncen;        // census particles per problem (pool A, grouped)
  std::vector<int64_t> h_blk_off;     // new-particle blocks: per block P+1 offsets (within block)
  std::vector<int64_t> h_blk_base;    // start of each block in buffer N
  int nblocks = 0;
  std::vector<int64_t> h_poff;        // transport pool offsets per problem (P+1)
  DevBuf<int64_t> coff, poff, boff, offs, counts, scan_scr, seg_dst, seg_src, Nvec, taken;
  DevBuf<int32_t> cprob, seg_new, refl, copies;
  DevBuf<double> vol, T, sig, cv, sa, ss, E, slope, acoef, dep_raw, path_raw, dep, si, leaks, ecen, ecen_part, ew,
      disc, Ep, theta, dt, tfloor, floor_v, psum, tree_scr, sum_scr, Ebnd;
  DevBuf<double2> ep64, sas64;
  DevBuf<float2> ep32, sas32;
  DevBuf<double> cfsa, cfbeta;  // Carter-Forest: true sigma_a and beta per cell
  bool cf_ready = false;
  DevBuf<uint8_t> chunk_key;
  DevBuf<int32_t> chunk_ord, chunk_hist;
  DevBuf<unsigned int> count_part;
  DevBuf<unsigned long long> cellcount;
  unsigned long long* stats = nullptr;
  unsigned long long* next = nullptr;
  int* bad = nullptr;
  Pool1Buf A, B, N;
  int64_t nA = 0, nN = 0, nB = 0;
  bool staged = false;  // B holds the grouped pool of the current census + blocks
  int grid_override = 0, batch = 256, tally_copies = kTallyCopies;
  int refill_min = 1;  // idle lanes before a wave of the tracker refills
};

}  // namespace imc

using namespace imc;

extern "C" {

struct imc1d_ctx : public Imc1dCtx {};

int imc1d_create(imc1d_ctx** out, int geometry, int nprob, const int32_t* ncells, const double* edges, int precision,
                 uint64_t seed, int device) {
  IMC_REQUIRE(out && ncells && edges && nprob > 0, kErrArg, "imc1d_create: bad arguments");
  IMC_REQUIRE(geometry == IMC_GEOM_SLAB || geometry == IMC_GEOM_SPHERE, kErrArg, "imc1d_create: bad geometry");
  IMC_REQUIRE(precision == IMC_FP64 || precision == IMC_FP32_TRACK, kErrArg, "imc1d_create: bad precision");
  IMC_HIP(hipSetDevice(device));
  imc1d_ctx* c = new imc1d_ctx();
  c->sph = geometry == IMC_GEOM_SPHERE;
  c->P = nprob;
  c->precision = precision;
  c->seed = seed;
  c->device = device;
  hipDeviceProp_t prop;
  IMC_HIP(hipGetDeviceProperties(&prop, device));
  c->n_cu = prop.multiProcessorCount;
  IMC_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  IMC_HIP(hipEventCreate(&c->ev0));
  IMC_HIP(hipEventCreate(&c->ev1));
  c->h_coff.assign(nprob + 1, 0);
  for (int p = 0; p < nprob; ++p) {
    IMC_REQUIRE(ncells[p] > 0 && ncells[p] <= (1 << 20), kErrArg, "imc1d_create: problem %d has %d cells", p,
                ncells[p]);
    c->h_coff[p + 1] = c->h_coff[p] + ncells[p];
  }
  int64_t nc = c->h_coff[nprob];
  IMC_REQUIRE(nc + 2 * (int64_t)nprob < (int64_t)0x7fffffff, kErrArg, "imc1d_create: too many cells");
  c->ncell = nc;
  std::vector<double2> e64(nc);
  std::vector<float2> e32(nc);
  std::vector<double> vol(nc);
  std::vector<int32_t> cp(nc);
  int64_t eo = 0;
  for (int p = 0; p < nprob; ++p) {
    for (int i = 0; i < ncells[p]; ++i) {
      double r0 = edges[eo + i], r1 = edges[eo + i + 1];
      IMC_REQUIRE(r1 > r0 && (!c->sph || r0 >= 0.0), kErrArg, "imc1d_create: bad edges in problem %d", p);
      int64_t ci = c->h_coff[p] + i;
      e64[ci] = make_double2(r0, r1);
      e32[ci] = make_float2((float)r0, (float)r1 - (float)r0);
      vol[ci] = c->sph ? (4.0 / 3.0) * M_PI * (r1 * r1 * r1 - r0 * r0 * r0) : (r1 - r0);
      cp[ci] = p;
    }
    eo += ncells[p] + 1;
  }
  int rc = 0;
  int P = nprob;
  if ((rc = c->coff.reserve(P + 1)) || (rc = c->poff.reserve(P + 1)) || (rc = c->boff.reserve(P + 1)) ||
      (rc = c->Nvec.reserve(P)) || (rc = c->taken.reserve(P + 1)) || (rc = c->cprob.reserve(nc)) || (rc = c->refl.reserve(2 * P)) ||
      (rc = c->ep64.reserve(nc)) || (rc = c->ep32.reserve(nc)) || (rc = c->sas64.reserve(nc)) ||
      (rc = c->sas32.reserve(nc)) || (rc = c->counts.reserve(nc + 1)) || (rc = c->offs.reserve(nc + 2)) ||
      (rc = c->scan_scr.reserve(scan_scratch(nc) + 8)) || (rc = c->tree_scr.reserve(1024)) ||
      (rc = c->sum_scr.reserve(canonical_sum_scratch(nc) + 8)))
    return rc;
  DevBuf<double>* cellbufs[] = {&c->vol, &c->T, &c->sig, &c->cv, &c->sa, &c->ss, &c->E, &c->slope, &c->acoef,
                                &c->dep, &c->si, &c->ecen, &c->ew, &c->disc};
  for (auto* b : cellbufs)
    if ((rc = b->reserve(nc))) return rc;
  DevBuf<double>* pbufs[] = {&c->leaks, &c->Ep, &c->theta, &c->dt, &c->tfloor, &c->floor_v, &c->psum, &c->Ebnd};
  for (auto* b : pbufs)
    if ((rc = b->reserve(2 * P))) return rc;
  if ((rc = c->dep_raw.reserve((nc + 2 * P) * kMaxTallyCopies))) return rc;
  if ((rc = c->path_raw.reserve(nc * kMaxTallyCopies))) return rc;
  if ((rc = c->ecen_part.reserve(nc * kMaxTallyCopies))) return rc;
  if ((rc = c->count_part.reserve(nc * kMaxTallyCopies))) return rc;
  if ((rc = c->cellcount.reserve(nc))) return rc;
  IMC_HIP(hipMemcpy(c->coff.p, c->h_coff.data(), sizeof(int64_t) * (P + 1), hipMemcpyHostToDevice));
  IMC_HIP(hipMemcpy(c->cprob.p, cp.data(), sizeof(int32_t) * nc, hipMemcpyHostToDevice));
  IMC_HIP(hipMemcpy(c->ep64.p, e64.data(), sizeof(double2) * nc, hipMemcpyHostToDevice));
  IMC_HIP(hipMemcpy(c->ep32.p, e32.data(), sizeof(float2) * nc, hipMemcpyHostToDevice));
  IMC_HIP(hipMemcpy(c->vol.p, vol.data(), sizeof(double) * nc, hipMemcpyHostToDevice));
  IMC_HIP(hipMalloc(&c->stats, sizeof(unsigned long long) * 8));
  IMC_HIP(hipMalloc(&c->next, sizeof(unsigned long long)));
  IMC_HIP(hipMalloc(&c->bad, sizeof(int)));
  c->A.rsz = c->B.rsz = c->N.rsz = (precision == IMC_FP64) ? 8 : 4;
  c->h_ncen.assign(P, 0);
  c->h_poff.assign(P + 1, 0);
  *out = c;
  return kOk;
}

void imc1d_destroy(imc1d_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  DevBuf<int64_t>* i64[] = {&c->coff, &c->poff, &c->boff, &c->offs, &c->counts, &c->scan_scr, &c->seg_dst,
                            &c->seg_src, &c->Nvec, &c->taken};
  for (auto* b : i64) b->release();
  DevBuf<int32_t>* i32[] = {&c->cprob, &c->seg_new, &c->refl, &c->copies};
  for (auto* b : i32) b->release();
  DevBuf<double>* d[] = {&c->vol, &c->T, &c->sig, &c->cv, &c->sa, &c->ss, &c->E, &c->slope, &c->acoef,
                         &c->dep_raw, &c->path_raw, &c->dep, &c->si, &c->leaks, &c->ecen, &c->ecen_part, &c->ew,
                         &c->disc, &c->Ep, &c->theta, &c->dt, &c->tfloor, &c->floor_v, &c->psum, &c->tree_scr,
                         &c->sum_scr, &c->Ebnd, &c->cfsa, &c->cfbeta};
  for (auto* b : d) b->release();
  c->ep64.release();
  c->sas64.release();
  c->chunk_key.release();
  c->chunk_ord.release();
  c->chunk_hist.release();
  c->ep32.release();
  c->sas32.release();
  c->count_part.release();
  c->cellcount.release();
  c->A.release();
  c->B.release();
  c->N.release();
  if (c->stats) (void)hipFree(c->stats);
  if (c->next) (void)hipFree(c->next);
  if (c->bad) (void)hipFree(c->bad);
  (void)hipEventDestroy(c->ev0);
  (void)hipEventDestroy(c->ev1);
  (void)hipStreamDestroy(c->stream);
  delete c;
}

int imc1d_set_step(imc1d_ctx* c, uint32_t step) {
  IMC_REQUIRE(c, kErrArg, "null ctx");
  c->step = step;
  return kOk;
}

int imc1d_set_seed(imc1d_ctx* c, uint64_t seed) {
  IMC_REQUIRE(c, kErrArg, "null ctx");
  c->seed = seed;
  return kOk;
}

int imc1d_set_launch(imc1d_ctx* c, int grid, int batch, int tally_copies) {
  IMC_REQUIRE(c && grid >= 0 && batch > 0, kErrArg, "imc1d_set_launch: bad arguments");
  IMC_REQUIRE(tally_copies >= 1 && tally_copies <= kMaxTallyCopies && (tally_copies & (tally_copies - 1)) == 0,
              kErrArg, "imc1d_set_launch: tally_copies must be a power of two in [1, %d]", kMaxTallyCopies);
  c->grid_override = grid;
  c->batch = batch;
  c->tally_copies = tally_copies;
  return kOk;
}

int imc1d_set_tracker(imc1d_ctx* c, int block, int refill_min) {
  IMC_REQUIRE(c, kErrArg, "null ctx");
  IMC_REQUIRE(block == 256, kErrArg, "imc1d_set_tracker: block must be 256");
  IMC_REQUIRE(refill_min >= 1 && refill_min <= 64, kErrArg, "imc1d_set_tracker: refill_min must be in [1, 64]");
  c->refill_min = refill_min;
  return kOk;
}

// census particles per problem (P values); returns the total
int64_t imc1d_census_counts(imc1d_ctx* c, int64_t* per_problem) {
  if (!c) return -1;
  if (per_problem)
    for (int p = 0; p < c->P; ++p) per_problem[p] = c->h_ncen[p];
  return c->nA;
}

// Census grouped by problem: counts[p] particles of problem p, in order.
int imc1d_put_census(imc1d_ctx* c, const int64_t* counts, const double* w, const double* mu, const double* t,
                     const double* x, const int64_t* cell) {
  IMC_REQUIRE(c && counts, kErrArg, "imc1d_put_census: bad arguments");
  IMC_HIP(hipSetDevice(c->device));
  int64_t n = 0;
  for (int p = 0; p < c->P; ++p) {
    IMC_REQUIRE(counts[p] >= 0, kErrArg, "imc1d_put_census: negative count");
    n += counts[p];
  }
  int rc = c->A.reserve(std::max<int64_t>(n, 1));
  if (rc) return rc;
  if (n > 0) {
    IMC_REQUIRE(w && mu && t && x && cell, kErrArg, "imc1d_put_census: null array");
    std::vector<int32_t> prob(n);
    int64_t q = 0;
    for (int p = 0; p < c->P; ++p)
      for (int64_t r = 0; r < counts[p]; ++r) prob[q++] = p;
    void* tmp[6];
    const void* src[6] = {w, mu, t, x, cell, prob.data()};
    size_t esz[6] = {8, 8, 8, 8, 8, 4};
    for (int a = 0; a < 6; ++a) {
      IMC_HIP(hipMalloc(&tmp[a], esz[a] * (size_t)n));
      IMC_HIP(hipMemcpyAsync(tmp[a], src[a], esz[a] * (size_t)n, hipMemcpyHostToDevice, c->stream));
    }
    int g = launch_blocks(n, 256);
    if (c->precision == IMC_FP64)
      k1_put<double><<<g, 256, 0, c->stream>>>(c->A.view<double>(), n, (double*)tmp[0], (double*)tmp[1],
                                               (double*)tmp[2], (double*)tmp[3], (int64_t*)tmp[4], (int32_t*)tmp[5]);
    else
      k1_put<float><<<g, 256, 0, c->stream>>>(c->A.view<float>(), n, (double*)tmp[0], (double*)tmp[1],
                                              (double*)tmp[2], (double*)tmp[3], (int64_t*)tmp[4], (int32_t*)tmp[5]);
    IMC_HIP(hipGetLastError());
    IMC_HIP(hipStreamSynchronize(c->stream));
    for (int a = 0; a < 6; ++a) IMC_HIP(hipFree(tmp[a]));
  }
  for (int p = 0; p < c->P; ++p) c->h_ncen[p] = counts[p];
  c->nA = n;
  c->nN = 0;
  c->staged = false;
  c->nblocks = 0;
  c->h_blk_off.clear();
  c->h_blk_base.clear();
  return kOk;
}

int imc1d_get_census(imc1d_ctx* c, int64_t n, double* w, double* mu, double* t, double* x, int64_t* cell,
                     int32_t* prob) {
  IMC_REQUIRE(c && n == c->nA, kErrArg, "imc1d_get_census: size mismatch");
  IMC_HIP(hipSetDevice(c->device));
  if (n == 0) return kOk;
  void* dst[6] = {w, mu, t, x, cell, prob};
  size_t esz[6] = {8, 8, 8, 8, 8, 4};
  void* tmp[6];
  for (int a = 0; a < 6; ++a) {
    IMC_REQUIRE(dst[a], kErrArg, "imc1d_get_census: null array");
    IMC_HIP(hipMalloc(&tmp[a], esz[a] * (size_t)n));
  }
  int g = launch_blocks(n, 256);
  if (c->precision == IMC_FP64)
    k1_get<double><<<g, 256, 0, c->stream>>>(c->A.view<double>(), n, (double*)tmp[0], (double*)tmp[1],
                                             (double*)tmp[2], (double*)tmp[3], (int64_t*)tmp[4], (int32_t*)tmp[5]);
  else
    k1_get<float><<<g, 256, 0, c->stream>>>(c->A.view<float>(), n, (double*)tmp[0], (double*)tmp[1],
                                            (double*)tmp[2], (double*)tmp[3], (int64_t*)tmp[4], (int32_t*)tmp[5]);
  IMC_HIP(hipGetLastError());
  for (int a = 0; a < 6; ++a)
    IMC_HIP(hipMemcpyAsync(dst[a], tmp[a], esz[a] * (size_t)n, hipMemcpyDeviceToHost, c->stream));
  IMC_HIP(hipStreamSynchronize(c->stream));
  for (int a = 0; a < 6; ++a) IMC_HIP(hipFree(tmp[a]));
  return kOk;
}

// Census energy per cell (np.bincount of the weights by cell,
// IMC/IMC1D.py:884-886) and per problem (np.sum(weights), IMC/IMC1D.py:889);
// both outputs nullable.
int imc1d_census_energy(imc1d_ctx* c, double* per_cell, double* per_problem) {
  IMC_REQUIRE(c, kErrArg, "null ctx");
  IMC_HIP(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  int P = c->P;
  int64_t n = c->nA, nc = c->ncell;
  if (per_cell) {
    IMC_HIP(hipMemsetAsync(c->ecen_part.p, 0, sizeof(double) * nc * c->tally_copies, s));
    if (n > 0) {
      // the census carries each particle's problem; pool offsets are not needed
      int g = std::min(launch_blocks(n, 256), c->n_cu * 4);
      if (c->precision == IMC_FP64)
        k1_comb_ecen<double><<<g, 256, 0, s>>>(c->A.view<double>(), n, c->coff.p, nc, c->tally_copies - 1,
                                               c->ecen_part.p);
      else
        k1_comb_ecen<float><<<g, 256, 0, s>>>(c->A.view<float>(), n, c->coff.p, nc, c->tally_copies - 1,
                                              c->ecen_part.p);
      IMC_HIP(hipGetLastError());
    }
    k_reduce_copies<double, double><<<launch_blocks(nc, 256), 256, 0, s>>>(nc, c->tally_copies, c->ecen_part.p,
                                                                          c->ecen.p);
    IMC_HIP(hipGetLastError());
    IMC_HIP(hipMemcpyAsync(per_cell, c->ecen.p, sizeof(double) * nc, hipMemcpyDeviceToHost, s));
  }
  if (per_problem) {
    std::vector<int64_t> off(P + 1, 0);
    for (int p = 0; p < P; ++p) off[p + 1] = off[p] + c->h_ncen[p];
    IMC_HIP(hipMemcpyAsync(c->taken.p, off.data(), sizeof(int64_t) * (P + 1), hipMemcpyHostToDevice, s));
    if (n > 0) {
      k1_segment_sums<<<P, kSegThreads, 0, s>>>(c->A.w, c->taken.p, c->psum.p);
      IMC_HIP(hipGetLastError());
    } else {
      IMC_HIP(hipMemsetAsync(c->psum.p, 0, sizeof(double) * P, s));
    }
    IMC_HIP(hipMemcpyAsync(per_problem, c->psum.p, sizeof(double) * P, hipMemcpyDeviceToHost, s));
    IMC_HIP(hipStreamSynchronize(s));  // `off` is a pageable source
  }
  IMC_HIP(hipStreamSynchronize(s));
  return kOk;
}

// Fleck factors (IMC/IMC1D.py:933-938); fails if f leaves [0,1].  theta, dt per problem.
int imc1d_set_material(imc1d_ctx* c, const double* T, const double* sigma_true, const double* cv, const double* theta,
                       const double* dt, double* sigma_a_out, double* sigma_s_out) {
  IMC_REQUIRE(c && T && sigma_true && cv && theta && dt, kErrArg, "imc1d_set_material: null argument");
  IMC_HIP(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  size_t bytes = sizeof(double) * c->ncell;
  IMC_HIP(hipMemcpyAsync(c->T.p, T, bytes, hipMemcpyHostToDevice, s));
  IMC_HIP(hipMemcpyAsync(c->sig.p, sigma_true, bytes, hipMemcpyHostToDevice, s));
  IMC_HIP(hipMemcpyAsync(c->cv.p, cv, bytes, hipMemcpyHostToDevice, s));
  IMC_HIP(hipMemcpyAsync(c->theta.p, theta, sizeof(double) * c->P, hipMemcpyHostToDevice, s));
  IMC_HIP(hipMemcpyAsync(c->dt.p, dt, sizeof(double) * c->P, hipMemcpyHostToDevice, s));
  IMC_HIP(hipMemsetAsync(c->bad, 0, sizeof(int), s));
  k1_fleck<<<launch_blocks(c->ncell, 256), 256, 0, s>>>(c->ncell, c->cprob.p, c->T.p, c->sig.p, c->cv.p, c->theta.p,
                                                        c->dt.p, c->sa.p, c->ss.p, c->sas64.p, c->sas32.p, c->bad);
  IMC_HIP(hipGetLastError());
  int bad = 0;
  IMC_HIP(hipMemcpyAsync(&bad, c->bad, sizeof(int), hipMemcpyDeviceToHost, s));
  if (sigma_a_out) IMC_HIP(hipMemcpyAsync(sigma_a_out, c->sa.p, bytes, hipMemcpyDeviceToHost, s));
  if (sigma_s_out) IMC_HIP(hipMemcpyAsync(sigma_s_out, c->ss.p, bytes, hipMemcpyDeviceToHost, s));
  IMC_HIP(hipStreamSynchronize(s));
  IMC_REQUIRE(!bad, kErrArg, "Fleck factor out of bounds");
  // a new step starts: no source blocks yet
  c->staged = false;
  c->nN = 0;
  c->nblocks = 0;
  c->h_blk_off.clear();
  c->h_blk_base.clear();
  return kOk;
}

static int add_block(imc1d_ctx* c, const std::vector<int64_t>& off) {
  int64_t total = off[c->P];
  int rc = c->N.reserve(0);
  (void)rc;
  if (c->nN + total > c->N.cap) {
    // grow keeping existing blocks
    Pool1Buf nb;
    nb.rsz = c->N.rsz;
    if ((rc = nb.reserve(std::max<int64_t>(c->nN + total, 1)))) return rc;
    if (c->nN > 0) {
      size_t r = (size_t)c->N.rsz, n = (size_t)c->nN;
      IMC_HIP(hipMemcpyAsync(nb.w, c->N.w, 8 * n, hipMemcpyDeviceToDevice, c->stream));
      IMC_HIP(hipMemcpyAsync(nb.mu, c->N.mu, r * n, hipMemcpyDeviceToDevice, c->stream));
      IMC_HIP(hipMemcpyAsync(nb.t, c->N.t, r * n, hipMemcpyDeviceToDevice, c->stream));
      IMC_HIP(hipMemcpyAsync(nb.x, c->N.x, r * n, hipMemcpyDeviceToDevice, c->stream));
      IMC_HIP(hipMemcpyAsync(nb.cell, c->N.cell, 4 * n, hipMemcpyDeviceToDevice, c->stream));
      IMC_HIP(hipMemcpyAsync(nb.prob, c->N.prob, 4 * n, hipMemcpyDeviceToDevice, c->stream));
      IMC_HIP(hipStreamSynchronize(c->stream));
    }
    c->N.release();
    c->N = nb;
  }
  c->staged = false;
  c->h_blk_base.push_back(c->nN);
  c->h_blk_off.insert(c->h_blk_off.end(), off.begin(), off.end());
  c->nblocks += 1;
  c->nN += total;
  return kOk;
}

// Face source on side 0 (left, mu > 0) or 1 (right, mu < 0): n[p] particles
// carrying E[p] in total; problems with n[p] = 0 get none.
int imc1d_sample_boundary(imc1d_ctx* c, int side, const int64_t* n, const double* E) {
  IMC_REQUIRE(c && n && E && (side == 0 || side == 1), kErrArg, "imc1d_sample_boundary: bad arguments");
  IMC_HIP(hipSetDevice(c->device));
  std::vector<int64_t> off(c->P + 1, 0);
  for (int p = 0; p < c->P; ++p) off[p + 1] = off[p] + std::max<int64_t>(n[p], 0);
  int64_t total = off[c->P];
  if (total == 0) return kOk;
  int64_t base = c->nN;
  int rc = add_block(c, off);
  if (rc) return rc;
  IMC_HIP(hipMemcpyAsync(c->boff.p, off.data(), sizeof(int64_t) * (c->P + 1), hipMemcpyHostToDevice, c->stream));
  IMC_HIP(hipMemcpyAsync(c->Ebnd.p, E, sizeof(double) * c->P, hipMemcpyHostToDevice, c->stream));
  Key key = make_key(c->seed);
  int g = launch_blocks(total, 256);
  if (c->precision == IMC_FP64)
    k1_sample_boundary<double><<<g, 256, 0, c->stream>>>(c->N.view<double>(), base, total, c->boff.p, c->P, side,
                                                         c->sph, c->Ebnd.p, c->dt.p, c->coff.p, c->ep64.p, key,
                                                         c->step);
  else
    k1_sample_boundary<float><<<g, 256, 0, c->stream>>>(c->N.view<float>(), base, total, c->boff.p, c->P, side,
                                                        c->sph, c->Ebnd.p, c->dt.p, c->coff.p, c->ep64.p, key,
                                                        c->step);
  IMC_HIP(hipGetLastError());
  IMC_HIP(hipStreamSynchronize(c->stream));  // boff / Ebnd are reused by the next call
  return kOk;
}

// counts -> per-cell offsets -> per-problem block offsets (host)
static int volume_counts(imc1d_ctx* c, const double* E, const int64_t* N, int mode, std::vector<int64_t>& off,
                         double* Ep_host) {
  hipStream_t s = c->stream;
  int P = c->P;
  IMC_HIP(hipMemcpyAsync(c->Nvec.p, N, sizeof(int64_t) * P, hipMemcpyHostToDevice, s));
  k1_problem_sums<<<P, 256, 0, s>>>(P, c->coff.p, E, c->Ep.p);
  IMC_HIP(hipGetLastError());
  k1_counts<<<launch_blocks(c->ncell, 256), 256, 0, s>>>(c->ncell, c->cprob.p, E, c->Ep.p, c->Nvec.p, mode,
                                                         c->counts.p);
  IMC_HIP(hipGetLastError());
  int rc = exclusive_scan_i64(c->counts.p, c->ncell, c->offs.p, c->scan_scr.p, s);
  if (rc) return rc;
  k1_take<<<launch_blocks(P + 1, 64), 64, 0, s>>>(P, c->coff.p, c->offs.p, c->taken.p);
  IMC_HIP(hipGetLastError());
  off.assign(P + 1, 0);
  IMC_HIP(hipMemcpyAsync(off.data(), c->taken.p, sizeof(int64_t) * (P + 1), hipMemcpyDeviceToHost, s));
  if (Ep_host) IMC_HIP(hipMemcpyAsync(Ep_host, c->Ep.p, sizeof(double) * P, hipMemcpyDeviceToHost, s));
  IMC_HIP(hipStreamSynchronize(s));
  return kOk;
}

static int launch_volume1(imc1d_ctx* c, Pool1Buf& dst, int64_t base, int64_t total, const double* QV, int kind,
                          uint32_t tag, uint32_t step) {
  if (total <= 0) return kOk;
  Key key = make_key(c->seed);
  int g = launch_blocks(total, 256);
  if (c->precision == IMC_FP64)
    k1_sample_volume<double><<<g, 256, 0, c->stream>>>(dst.view<double>(), base, total, c->offs.p, c->ncell,
                                                       c->cprob.p, c->coff.p, c->ep64.p, QV, c->T.p, c->slope.p,
                                                       c->acoef.p, c->dt.p, kind, c->sph, c->cfsa.p, c->cfbeta.p,
                                                       key, step, tag);
  else
    k1_sample_volume<float><<<g, 256, 0, c->stream>>>(dst.view<float>(), base, total, c->offs.p, c->ncell,
                                                      c->cprob.p, c->coff.p, c->ep64.p, QV, c->T.p, c->slope.p,
                                                      c->acoef.p, c->dt.p, kind, c->sph, c->cfsa.p, c->cfbeta.p,
                                                      key, step, tag);
  IMC_HIP(hipGetLastError());
  return kOk;
}

// Equilibrium census (IMC/IMC1D.py:220-248, 535-575): replaces the census.
int imc1d_equilibrium_sample(imc1d_ctx* c, const int64_t* N, const double* Tr, const double* T_emit_floor) {
  IMC_REQUIRE(c && N && Tr && T_emit_floor, kErrArg, "imc1d_equilibrium_sample: bad arguments");
  IMC_HIP(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  IMC_HIP(hipMemcpyAsync(c->T.p, Tr, sizeof(double) * c->ncell, hipMemcpyHostToDevice, s));
  IMC_HIP(hipMemcpyAsync(c->tfloor.p, T_emit_floor, sizeof(double) * c->P, hipMemcpyHostToDevice, s));
  k1_equilibrium_energy<<<launch_blocks(c->ncell, 256), 256, 0, s>>>(c->ncell, c->cprob.p, c->T.p, c->vol.p,
                                                                     c->tfloor.p, c->E.p);
  IMC_HIP(hipGetLastError());
  std::vector<int64_t> off;
  int rc = volume_counts(c, c->E.p, N, 1, off, nullptr);
  if (rc) return rc;
  int64_t total = off[c->P];
  if ((rc = c->A.reserve(std::max<int64_t>(total, 1)))) return rc;
  if ((rc = launch_volume1(c, c->A, 0, total, c->E.p, 1, kTagIC, 0))) return rc;
  IMC_HIP(hipStreamSynchronize(s));
  for (int p = 0; p < c->P; ++p) c->h_ncen[p] = off[p + 1] - off[p];
  c->nA = total;
  c->nN = 0;
  c->staged = false;
  c->nblocks = 0;
  c->h_blk_off.clear();
  c->h_blk_base.clear();
  return kOk;
}

// Volume source block: IMC_SRC_EMISSION (a c T^4 sigma_a dt V from the last
// imc1d_set_material, linear tilt), IMC_SRC_FIXED (QV per cell, uniform) or
// IMC_SRC_CF (QV = Carter-Forest emission energies, tilt from the T of the
// last imc1d_set_material_cf, truncated-exponential times).
int imc1d_sample_volume(imc1d_ctx* c, int kind, const int64_t* N, const double* QV_host, const double* T_emit_floor,
                        double* E_out, double* Etot_out) {
  IMC_REQUIRE(c && N && T_emit_floor, kErrArg, "imc1d_sample_volume: bad arguments");
  IMC_REQUIRE(kind == IMC_SRC_EMISSION || ((kind == IMC_SRC_FIXED || kind == IMC_SRC_CF) && QV_host), kErrArg,
              "imc1d_sample_volume: bad kind");
  IMC_REQUIRE(kind != IMC_SRC_CF || c->cf_ready, kErrArg,
              "imc1d_sample_volume: Carter-Forest emission needs imc1d_set_material_cf");
  IMC_HIP(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  IMC_HIP(hipMemcpyAsync(c->tfloor.p, T_emit_floor, sizeof(double) * c->P, hipMemcpyHostToDevice, s));
  const double* QV;
  if (kind == IMC_SRC_EMISSION) {
    k1_emission_energy<<<launch_blocks(c->ncell, 256), 256, 0, s>>>(c->ncell, c->cprob.p, c->T.p, c->sa.p, c->vol.p,
                                                                    c->dt.p, c->tfloor.p, c->E.p);
    IMC_HIP(hipGetLastError());
    k1_slopes<<<launch_blocks(c->ncell, 256), 256, 0, s>>>(c->P, c->coff.p, c->T.p, c->ep64.p, c->sph, c->slope.p,
                                                           c->acoef.p);
    IMC_HIP(hipGetLastError());
    QV = c->E.p;
  } else {
    IMC_HIP(hipMemcpyAsync(c->ecen.p, QV_host, sizeof(double) * c->ncell, hipMemcpyHostToDevice, s));
    QV = c->ecen.p;
    if (kind == IMC_SRC_CF) {
      k1_slopes<<<launch_blocks(c->ncell, 256), 256, 0, s>>>(c->P, c->coff.p, c->T.p, c->ep64.p, c->sph,
                                                             c->slope.p, c->acoef.p);
      IMC_HIP(hipGetLastError());
    }
  }
  const int count_mode = kind == IMC_SRC_CF ? (c->sph ? 3 : 2) : 0;
  std::vector<int64_t> off;
  int rc = volume_counts(c, QV, N, count_mode, off, Etot_out);
  if (rc) return rc;
  if (E_out) {
    IMC_HIP(hipMemcpyAsync(E_out, QV, sizeof(double) * c->ncell, hipMemcpyDeviceToHost, s));
    IMC_HIP(hipStreamSynchronize(s));
  }
  int64_t total = off[c->P];
  if (total == 0) return kOk;
  int64_t base = c->nN;
  if ((rc = add_block(c, off))) return rc;
  const int skind = kind == IMC_SRC_EMISSION ? 0 : (kind == IMC_SRC_CF ? 3 : 2);
  if ((rc = launch_volume1(c, c->N, base - off[0], total, QV, skind,
                           kind == IMC_SRC_FIXED ? kTagSource : kTagEmission, c->step)))
    return rc;
  IMC_HIP(hipStreamSynchronize(s));
  return kOk;
}

// Build the grouped transport pool: per problem the census, then each source
// block in call order (the reference's concatenation order, IMC/IMC1D.py:947-1002).
static int build_transport_pool(imc1d_ctx* c) {
  int P = c->P;
  std::vector<int64_t> dst, src;
  std::vector<int32_t> fromn;
  c->h_poff.assign(P + 1, 0);
  int64_t pos = 0;
  int64_t aoff = 0;
  for (int p = 0; p < P; ++p) {
    c->h_poff[p] = pos;
    if (c->h_ncen[p] > 0) {
      dst.push_back(pos);
      src.push_back(aoff);
      fromn.push_back(0);
      pos += c->h_ncen[p];
    }
    aoff += c->h_ncen[p];
    for (int b = 0; b < c->nblocks; ++b) {
      const int64_t* bo = &c->h_blk_off[(size_t)b * (P + 1)];
      int64_t cnt = bo[p + 1] - bo[p];
      if (cnt > 0) {
        dst.push_back(pos);
        src.push_back(c->h_blk_base[b] + bo[p]);
        fromn.push_back(1);
        pos += cnt;
      }
    }
  }
  c->h_poff[P] = pos;
  c->nB = pos;
  c->staged = true;
  int nseg = (int)dst.size();
  dst.push_back(pos);
  int rc;
  if ((rc = c->B.reserve(std::max<int64_t>(pos, 1)))) return rc;
  if ((rc = c->seg_dst.reserve(nseg + 1)) || (rc = c->seg_src.reserve(nseg + 1)) ||
      (rc = c->seg_new.reserve(nseg + 1)))
    return rc;
  hipStream_t s = c->stream;
  IMC_HIP(hipMemcpyAsync(c->poff.p, c->h_poff.data(), sizeof(int64_t) * (P + 1), hipMemcpyHostToDevice, s));
  if (pos > 0) {
    IMC_HIP(hipMemcpyAsync(c->seg_dst.p, dst.data(), sizeof(int64_t) * (nseg + 1), hipMemcpyHostToDevice, s));
    IMC_HIP(hipMemcpyAsync(c->seg_src.p, src.data(), sizeof(int64_t) * nseg, hipMemcpyHostToDevice, s));
    IMC_HIP(hipMemcpyAsync(c->seg_new.p, fromn.data(), sizeof(int32_t) * nseg, hipMemcpyHostToDevice, s));
    int g = launch_blocks(pos, 256);
    if (c->precision == IMC_FP64)
      k1_gather<double><<<g, 256, 0, s>>>(c->A.view<double>(), c->N.view<double>(), c->B.view<double>(), nseg,
                                          c->seg_dst.p, c->seg_src.p, c->seg_new.p, pos);
    else
      k1_gather<float><<<g, 256, 0, s>>>(c->A.view<float>(), c->N.view<float>(), c->B.view<float>(), nseg,
                                         c->seg_dst.p, c->seg_src.p, c->seg_new.p, pos);
    IMC_HIP(hipGetLastError());
  }
  IMC_HIP(hipStreamSynchronize(s));
  return kOk;
}

// Stage the grouped transport pool (census + this step's source blocks, per
// problem) without tracking it; per_problem (nullable) receives its counts.
int64_t imc1d_stage_pool(imc1d_ctx* c, int64_t* per_problem) {
  if (!c) return kErrArg;
  if (hipSetDevice(c->device) != hipSuccess) return kErrHip;
  int rc = build_transport_pool(c);
  if (rc) return rc;
  if (per_problem)
    for (int p = 0; p < c->P; ++p) per_problem[p] = c->h_poff[p + 1] - c->h_poff[p];
  return c->nB;
}

// Download the staged (or, after imc1d_transport, the transported) pool.
int imc1d_get_pool(imc1d_ctx* c, int64_t n, double* w, double* mu, double* t, double* x, int64_t* cell,
                   int32_t* prob) {
  IMC_REQUIRE(c && n == c->nB, kErrArg, "imc1d_get_pool: size mismatch");
  IMC_HIP(hipSetDevice(c->device));
  if (n == 0) return kOk;
  void* dst[6] = {w, mu, t, x, cell, prob};
  size_t esz[6] = {8, 8, 8, 8, 8, 4};
  void* tmp[6];
  for (int a = 0; a < 6; ++a) {
    IMC_REQUIRE(dst[a], kErrArg, "imc1d_get_pool: null array");
    IMC_HIP(hipMalloc(&tmp[a], esz[a] * (size_t)n));
  }
  int g = launch_blocks(n, 256);
  if (c->precision == IMC_FP64)
    k1_get<double><<<g, 256, 0, c->stream>>>(c->B.view<double>(), n, (double*)tmp[0], (double*)tmp[1],
                                             (double*)tmp[2], (double*)tmp[3], (int64_t*)tmp[4], (int32_t*)tmp[5]);
  else
    k1_get<float><<<g, 256, 0, c->stream>>>(c->B.view<float>(), n, (double*)tmp[0], (double*)tmp[1],
                                            (double*)tmp[2], (double*)tmp[3], (int64_t*)tmp[4], (int32_t*)tmp[5]);
  IMC_HIP(hipGetLastError());
  for (int a = 0; a < 6; ++a)
    IMC_HIP(hipMemcpyAsync(dst[a], tmp[a], esz[a] * (size_t)n, hipMemcpyDeviceToHost, c->stream));
  IMC_HIP(hipStreamSynchronize(c->stream));
  for (int a = 0; a < 6; ++a) IMC_HIP(hipFree(tmp[a]));
  return kOk;
}

// Transport census + this step's sources (IMC/IMC1D.py:1004-1019).
// reflect: 2 per problem; weight_floor per problem (< 0: 1e-10 sum(w)/N).
// Outputs: dep, si per cell; leak, floor_leak per problem; stats[8] = events,
// boundary crossings, scatters, census events, floor kills, reflections,
// non-positive distances (the reference asserts), particles transported.
int imc1d_transport(imc1d_ctx* c, const int32_t* reflect, const double* weight_floor, double* dep_out,
                    double* si_out, double* leak_out, double* floor_leak_out, int64_t* stats_out) {
  IMC_REQUIRE(c && reflect && weight_floor, kErrArg, "imc1d_transport: bad arguments");
  IMC_HIP(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  int P = c->P;
  int rc = build_transport_pool(c);
  if (rc) return rc;
  int64_t n = c->nB;
  int64_t tally_n = c->ncell + 2 * (int64_t)P;
  IMC_HIP(hipMemsetAsync(c->dep_raw.p, 0, sizeof(double) * tally_n * c->tally_copies, s));
  IMC_HIP(hipMemsetAsync(c->path_raw.p, 0, sizeof(double) * c->ncell * c->tally_copies, s));
  IMC_HIP(hipMemsetAsync(c->stats, 0, sizeof(unsigned long long) * 8, s));
  IMC_HIP(hipMemsetAsync(c->next, 0, sizeof(unsigned long long), s));
  IMC_HIP(hipMemcpyAsync(c->refl.p, reflect, sizeof(int32_t) * 2 * P, hipMemcpyHostToDevice, s));
  // weight floors: given, or 1e-10 * sum(w) / N per problem
  std::vector<double> fl(weight_floor, weight_floor + P);
  bool any_auto = false;
  for (int p = 0; p < P; ++p) any_auto |= fl[p] < 0.0;
  if (any_auto) {
    if (n > 0) {
      k1_segment_sums<<<P, kSegThreads, 0, s>>>(c->B.w, c->poff.p, c->psum.p);
      IMC_HIP(hipGetLastError());
    } else {
      IMC_HIP(hipMemsetAsync(c->psum.p, 0, sizeof(double) * P, s));
    }
    k1_floor<<<launch_blocks(P, 64), 64, 0, s>>>(P, c->psum.p, c->poff.p, c->floor_v.p);
    IMC_HIP(hipGetLastError());
    std::vector<double> autof(P);
    IMC_HIP(hipMemcpyAsync(autof.data(), c->floor_v.p, sizeof(double) * P, hipMemcpyDeviceToHost, s));
    IMC_HIP(hipStreamSynchronize(s));
    for (int p = 0; p < P; ++p)
      if (fl[p] < 0.0) fl[p] = autof[p];
  }
  IMC_HIP(hipMemcpyAsync(c->floor_v.p, fl.data(), sizeof(double) * P, hipMemcpyHostToDevice, s));
  int block = 256, grid;
  if (c->grid_override > 0) {
    grid = c->grid_override;
  } else {
    int per_cu = 0;
    const void* fn = c->precision == IMC_FP64 ? (c->sph ? (const void*)k1_transport<double, 1>
                                                        : (const void*)k1_transport<double, 0>)
                                              : (c->sph ? (const void*)k1_transport<float, 1>
                                                        : (const void*)k1_transport<float, 0>);
    IMC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, block, 0));
    grid = std::max(per_cu, 1) * c->n_cu;
  }
  grid = (int)std::max<int64_t>(std::min<int64_t>(grid, (n + block - 1) / block), 1);
  Key key = make_key(c->seed);
  QueueMap Q = pool_order_queue(n);
  if (n > 0 && (c->batch & (c->batch - 1)) == 0) {
    // pickup order: chunks of `batch` particles with long expected histories
    // first (256 slab problems of 50-2000 cells, default launch: FP32 tracker
    // 60.9 -> 48.5 ms, FP64 125 -> 105 ms; one slab of 1e5 particles unchanged)
    int shift = 0;
    while ((1 << shift) < c->batch) ++shift;
    const int64_t nchunk = (n + c->batch - 1) >> shift;
    if ((rc = c->chunk_key.reserve(nchunk)) || (rc = c->chunk_ord.reserve(nchunk)) ||
        (rc = c->chunk_hist.reserve((int64_t)kChunkGrid * kChunkKeys)))
      return rc;
    const int gk = chunk_grid(nchunk);
    if (c->precision == IMC_FP64)
      k1_chunk_key<double><<<gk, kChunkBlock, 0, s>>>(c->B.view<double>(), n, shift, c->coff.p, c->ep64.p, c->sa.p,
                                                      c->ss.p, c->dt.p, c->chunk_key.p, c->chunk_hist.p);
    else
      k1_chunk_key<float><<<gk, kChunkBlock, 0, s>>>(c->B.view<float>(), n, shift, c->coff.p, c->ep64.p, c->sa.p,
                                                     c->ss.p, c->dt.p, c->chunk_key.p, c->chunk_hist.p);
    IMC_HIP(hipGetLastError());
    if ((rc = chunk_order(c->chunk_key.p, nchunk, gk, c->chunk_hist.p, c->chunk_ord.p, s))) return rc;
    Q = QueueMap{c->chunk_ord.p, nchunk << shift, shift};
  }
  IMC_HIP(hipEventRecord(c->ev0, s));
  if (n > 0) {
    if (c->precision == IMC_FP64) {
      Transport1Args<double> A{c->B.view<double>(), n, c->poff.p, c->coff.p, c->ep64.p, c->sas64.p, c->dt.p,
                               c->floor_v.p, c->refl.p, key, c->step, c->next, c->batch, c->dep_raw.p,
                               c->path_raw.p, c->ncell, P, c->tally_copies - 1, c->stats, Q, c->refill_min};
      if (c->sph)
        k1_transport<double, 1><<<grid, block, 0, s>>>(A);
      else
        k1_transport<double, 0><<<grid, block, 0, s>>>(A);
    } else {
      Transport1Args<float> A{c->B.view<float>(), n, c->poff.p, c->coff.p, c->ep32.p, c->sas32.p, c->dt.p,
                              c->floor_v.p, c->refl.p, key, c->step, c->next, c->batch, c->dep_raw.p,
                              c->path_raw.p, c->ncell, P, c->tally_copies - 1, c->stats, Q, c->refill_min};
      if (c->sph)
        k1_transport<float, 1><<<grid, block, 0, s>>>(A);
      else
        k1_transport<float, 0><<<grid, block, 0, s>>>(A);
    }
    IMC_HIP(hipGetLastError());
  }
  IMC_HIP(hipEventRecord(c->ev1, s));
  k1_finalize<<<launch_blocks(tally_n, 256), 256, 0, s>>>(c->ncell, P, c->tally_copies, c->cprob.p, c->dep_raw.p,
                                                          c->path_raw.p, c->sa.p, c->sas32.p,
                                                          c->precision == IMC_FP32_TRACK, c->vol.p, c->dt.p, c->dep.p,
                                                          c->si.p, c->leaks.p);
  IMC_HIP(hipGetLastError());
  size_t bytes = sizeof(double) * c->ncell;
  if (dep_out) IMC_HIP(hipMemcpyAsync(dep_out, c->dep.p, bytes, hipMemcpyDeviceToHost, s));
  if (si_out) IMC_HIP(hipMemcpyAsync(si_out, c->si.p, bytes, hipMemcpyDeviceToHost, s));
  std::vector<double> lk(2 * P);
  IMC_HIP(hipMemcpyAsync(lk.data(), c->leaks.p, sizeof(double) * 2 * P, hipMemcpyDeviceToHost, s));
  unsigned long long st[8];
  IMC_HIP(hipMemcpyAsync(st, c->stats, sizeof(st), hipMemcpyDeviceToHost, s));
  IMC_HIP(hipStreamSynchronize(s));
  IMC_HIP(hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1));
  for (int p = 0; p < P; ++p) {
    if (leak_out) leak_out[p] = lk[2 * p];
    if (floor_leak_out) floor_leak_out[p] = lk[2 * p + 1];
  }
  if (stats_out) {
    for (int a = 0; a < 7; ++a) stats_out[a] = (int64_t)st[a];
    stats_out[7] = n;
  }
  return kOk;
}

float imc1d_last_transport_ms(imc1d_ctx* c) { return c ? c->last_ms : -1.0f; }

// ---- Carter-Forest variant (IMC/IMC1D_CarterForest.py), FP64 contexts ----
int imc1d_set_material_cf(imc1d_ctx* c, const double* T, const double* sigma_a, const double* beta,
                          const double* dt) {
  IMC_REQUIRE(c && T && sigma_a && beta && dt, kErrArg, "imc1d_set_material_cf: null argument");
  IMC_REQUIRE(c->precision == IMC_FP64, kErrArg, "imc1d_set_material_cf: the Carter-Forest tracker runs in FP64");
  IMC_HIP(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  int rc;
  if ((rc = c->cfsa.reserve(c->ncell)) || (rc = c->cfbeta.reserve(c->ncell))) return rc;
  size_t bytes = sizeof(double) * c->ncell;
  IMC_HIP(hipMemcpyAsync(c->T.p, T, bytes, hipMemcpyHostToDevice, s));
  IMC_HIP(hipMemcpyAsync(c->cfsa.p, sigma_a, bytes, hipMemcpyHostToDevice, s));
  IMC_HIP(hipMemcpyAsync(c->cfbeta.p, beta, bytes, hipMemcpyHostToDevice, s));
  IMC_HIP(hipMemcpyAsync(c->dt.p, dt, sizeof(double) * c->P, hipMemcpyHostToDevice, s));
  IMC_HIP(hipStreamSynchronize(s));
  c->cf_ready = true;
  // a new step starts: no source blocks yet
  c->staged = false;
  c->nN = 0;
  c->nblocks = 0;
  c->h_blk_off.clear();
  c->h_blk_base.clear();
  return kOk;
}

int imc1d_transport_cf(imc1d_ctx* c, const int32_t* reflect, const double* weight_floor, int64_t max_events,
                       double fastpath_threshold, double* dep_out, double* si_out, double* leak_out,
                       int64_t* stats_out, int64_t* transported_out) {
  IMC_REQUIRE(c && reflect && weight_floor, kErrArg, "imc1d_transport_cf: bad arguments");
  IMC_REQUIRE(c->cf_ready, kErrArg, "imc1d_transport_cf: call imc1d_set_material_cf first");
  IMC_REQUIRE(c->ncell + 2 * (int64_t)c->P + c->ncell < INT32_MAX, kErrArg, "imc1d_transport_cf: too many cells");
  IMC_HIP(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  int P = c->P;
  int rc = build_transport_pool(c);
  if (rc) return rc;
  int64_t n = c->nB;
  int64_t tally_n = c->ncell + 2 * (int64_t)P;
  IMC_HIP(hipMemsetAsync(c->dep_raw.p, 0, sizeof(double) * tally_n * c->tally_copies, s));
  IMC_HIP(hipMemsetAsync(c->path_raw.p, 0, sizeof(double) * c->ncell * c->tally_copies, s));
  IMC_HIP(hipMemsetAsync(c->stats, 0, sizeof(unsigned long long) * 8, s));
  IMC_HIP(hipMemsetAsync(c->next, 0, sizeof(unsigned long long), s));
  IMC_HIP(hipMemcpyAsync(c->refl.p, reflect, sizeof(int32_t) * 2 * P, hipMemcpyHostToDevice, s));
  // weight floors: given, or 1e-10 * sum(w) / N per problem (IMC/IMC1D_CarterForest.py:1243)
  std::vector<double> fl(weight_floor, weight_floor + P);
  bool any_auto = false;
  for (int p = 0; p < P; ++p) any_auto |= fl[p] < 0.0;
  if (any_auto) {
    if (n > 0) {
      k1_segment_sums<<<P, kSegThreads, 0, s>>>(c->B.w, c->poff.p, c->psum.p);
      IMC_HIP(hipGetLastError());
    } else {
      IMC_HIP(hipMemsetAsync(c->psum.p, 0, sizeof(double) * P, s));
    }
    k1_floor<<<launch_blocks(P, 64), 64, 0, s>>>(P, c->psum.p, c->poff.p, c->floor_v.p);
    IMC_HIP(hipGetLastError());
    std::vector<double> autof(P);
    IMC_HIP(hipMemcpyAsync(autof.data(), c->floor_v.p, sizeof(double) * P, hipMemcpyDeviceToHost, s));
    IMC_HIP(hipStreamSynchronize(s));
    for (int p = 0; p < P; ++p)
      if (fl[p] < 0.0) fl[p] = autof[p];
  }
  IMC_HIP(hipMemcpyAsync(c->floor_v.p, fl.data(), sizeof(double) * P, hipMemcpyHostToDevice, s));
  int block = 256, grid;
  if (c->grid_override > 0) {
    grid = c->grid_override;
  } else {
    int per_cu = 0;
    const void* fn = c->sph ? (const void*)k1_transport_cf<1> : (const void*)k1_transport_cf<0>;
    IMC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, block, 0));
    grid = std::max(per_cu, 1) * c->n_cu;
  }
  grid = (int)std::max<int64_t>(std::min<int64_t>(grid, (n + block - 1) / block), 1);
  Key key = make_key(c->seed);
  IMC_HIP(hipEventRecord(c->ev0, s));
  if (n > 0) {
    Cf1Args A{c->B.view<double>(), n, c->poff.p, c->coff.p, c->ep64.p, c->cfsa.p, c->cfbeta.p, c->dt.p,
              c->floor_v.p, c->refl.p, max_events, fastpath_threshold, key, c->step, c->next, c->batch,
              c->dep_raw.p, c->path_raw.p, c->ncell, P, c->tally_copies - 1, c->stats, pool_order_queue(n)};
    if (c->sph)
      k1_transport_cf<1><<<grid, block, 0, s>>>(A);
    else
      k1_transport_cf<0><<<grid, block, 0, s>>>(A);
    IMC_HIP(hipGetLastError());
  }
  IMC_HIP(hipEventRecord(c->ev1, s));
  k1_cf_finalize<<<launch_blocks(tally_n, 256), 256, 0, s>>>(c->ncell, P, c->tally_copies, c->cprob.p,
                                                             c->dep_raw.p, c->path_raw.p, c->vol.p, c->dt.p,
                                                             c->dep.p, c->si.p, c->leaks.p);
  IMC_HIP(hipGetLastError());
  size_t bytes = sizeof(double) * c->ncell;
  if (dep_out) IMC_HIP(hipMemcpyAsync(dep_out, c->dep.p, bytes, hipMemcpyDeviceToHost, s));
  if (si_out) IMC_HIP(hipMemcpyAsync(si_out, c->si.p, bytes, hipMemcpyDeviceToHost, s));
  std::vector<double> lk(2 * P);
  IMC_HIP(hipMemcpyAsync(lk.data(), c->leaks.p, sizeof(double) * 2 * P, hipMemcpyDeviceToHost, s));
  unsigned long long st[8];
  IMC_HIP(hipMemcpyAsync(st, c->stats, sizeof(st), hipMemcpyDeviceToHost, s));
  IMC_HIP(hipStreamSynchronize(s));
  IMC_HIP(hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1));
  if (leak_out)
    for (int p = 0; p < P; ++p) leak_out[p] = lk[2 * p];
  if (stats_out)
    for (int a = 0; a < 8; ++a) stats_out[a] = (int64_t)st[a];
  if (transported_out) *transported_out = n;
  return kOk;
}

// Comb each problem to Nmax[p] (IMC/IMC1D.py:770-828); the transported pool
// becomes the new census, grouped by problem.  Optional outputs: ecen, disc
// per cell; ecen_total, total_rad per problem.
int imc1d_comb(imc1d_ctx* c, const int64_t* Nmax, double* ecen_out, double* disc_out, double* ecen_total_out,
               double* total_rad_out) {
  IMC_REQUIRE(c && Nmax, kErrArg, "imc1d_comb: bad arguments");
  IMC_REQUIRE(c->staged, kErrState, "imc1d_comb: no transported or staged pool");
  IMC_HIP(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  int P = c->P;
  int64_t n = c->nB, nc = c->ncell;
  int rc;
  IMC_HIP(hipMemsetAsync(c->ecen_part.p, 0, sizeof(double) * nc * c->tally_copies, s));
  if (n > 0) {
    int g = std::min(launch_blocks(n, 256), c->n_cu * 4);
    if (c->precision == IMC_FP64)
      k1_comb_ecen<double><<<g, 256, 0, s>>>(c->B.view<double>(), n, c->coff.p, nc, c->tally_copies - 1,
                                             c->ecen_part.p);
    else
      k1_comb_ecen<float><<<g, 256, 0, s>>>(c->B.view<float>(), n, c->coff.p, nc, c->tally_copies - 1,
                                            c->ecen_part.p);
    IMC_HIP(hipGetLastError());
  }
  k_reduce_copies<double, double><<<launch_blocks(nc, 256), 256, 0, s>>>(nc, c->tally_copies, c->ecen_part.p, c->ecen.p);
  IMC_HIP(hipGetLastError());
  k1_problem_sums<<<P, 256, 0, s>>>(P, c->coff.p, c->ecen.p, c->Ep.p);
  IMC_HIP(hipGetLastError());
  IMC_HIP(hipMemcpyAsync(c->Nvec.p, Nmax, sizeof(int64_t) * P, hipMemcpyHostToDevice, s));
  k1_comb_ew<<<launch_blocks(nc, 256), 256, 0, s>>>(nc, c->cprob.p, c->ecen.p, c->Ep.p, c->Nvec.p, c->ew.p);
  IMC_HIP(hipGetLastError());
  if ((rc = c->copies.reserve(n + 1)) || (rc = c->counts.reserve(n + 2)) ||
      (rc = c->scan_scr.reserve(scan_scratch(n) + 8)))
    return rc;
  Key key = make_key(c->seed);
  int g = launch_blocks(n, 256);
  if (n > 0) {
    if (c->precision == IMC_FP64)
      k1_comb_copies<double><<<g, 256, 0, s>>>(c->B.view<double>(), n, c->poff.p, c->coff.p, c->ew.p, c->Ep.p, key,
                                               c->step, c->copies.p);
    else
      k1_comb_copies<float><<<g, 256, 0, s>>>(c->B.view<float>(), n, c->poff.p, c->coff.p, c->ew.p, c->Ep.p, key,
                                              c->step, c->copies.p);
    IMC_HIP(hipGetLastError());
  }
  if ((rc = exclusive_scan_i32_to_i64(c->copies.p, n, c->counts.p, c->scan_scr.p, s))) return rc;
  // per-problem census counts from the scan at the problem boundaries
  std::vector<int64_t> newoff(P + 1);
  k1_take<<<launch_blocks(P + 1, 64), 64, 0, s>>>(P, c->poff.p, c->counts.p, c->taken.p);
  IMC_HIP(hipGetLastError());
  IMC_HIP(hipMemcpyAsync(newoff.data(), c->taken.p, sizeof(int64_t) * (P + 1), hipMemcpyDeviceToHost, s));
  IMC_HIP(hipStreamSynchronize(s));
  int64_t total = newoff[P];
  if ((rc = c->A.reserve(std::max<int64_t>(total, 1)))) return rc;
  bool want_counts = disc_out != nullptr;
  if (want_counts) IMC_HIP(hipMemsetAsync(c->count_part.p, 0, sizeof(unsigned int) * nc * c->tally_copies, s));
  if (n > 0) {
    if (c->precision == IMC_FP64)
      k1_comb_write<double><<<g, 256, 0, s>>>(c->B.view<double>(), c->A.view<double>(), n, c->coff.p, c->counts.p,
                                              c->ew.p, c->Ep.p, c->tally_copies - 1,
                                              want_counts ? c->count_part.p : nullptr, nc);
    else
      k1_comb_write<float><<<g, 256, 0, s>>>(c->B.view<float>(), c->A.view<float>(), n, c->coff.p, c->counts.p,
                                             c->ew.p, c->Ep.p, c->tally_copies - 1,
                                             want_counts ? c->count_part.p : nullptr, nc);
    IMC_HIP(hipGetLastError());
  }
  size_t bytes = sizeof(double) * nc;
  if (want_counts) {
    k_reduce_copies<unsigned int, unsigned long long><<<launch_blocks(nc, 256), 256, 0, s>>>(nc, c->tally_copies, c->count_part.p, c->cellcount.p);
    IMC_HIP(hipGetLastError());
    k1_comb_disc<<<launch_blocks(nc, 256), 256, 0, s>>>(nc, c->cprob.p, c->ecen.p, c->ew.p, c->cellcount.p, c->Ep.p,
                                                        c->disc.p);
    IMC_HIP(hipGetLastError());
    IMC_HIP(hipMemcpyAsync(disc_out, c->disc.p, bytes, hipMemcpyDeviceToHost, s));
  }
  if (ecen_out) IMC_HIP(hipMemcpyAsync(ecen_out, c->ecen.p, bytes, hipMemcpyDeviceToHost, s));
  if (ecen_total_out) IMC_HIP(hipMemcpyAsync(ecen_total_out, c->Ep.p, sizeof(double) * P, hipMemcpyDeviceToHost, s));
  for (int p = 0; p < P; ++p) c->h_ncen[p] = newoff[p + 1] - newoff[p];
  c->nA = total;
  c->staged = false;
  if (total_rad_out) {
    IMC_HIP(hipMemcpyAsync(c->poff.p, newoff.data(), sizeof(int64_t) * (P + 1), hipMemcpyHostToDevice, s));
    if (total > 0) {
      k1_segment_sums<<<P, kSegThreads, 0, s>>>(c->A.w, c->poff.p, c->psum.p);
      IMC_HIP(hipGetLastError());
    } else {
      IMC_HIP(hipMemsetAsync(c->psum.p, 0, sizeof(double) * P, s));
    }
    IMC_HIP(hipMemcpyAsync(total_rad_out, c->psum.p, sizeof(double) * P, hipMemcpyDeviceToHost, s));
  }
  IMC_HIP(hipStreamSynchronize(s));
  c->nN = 0;
  c->nblocks = 0;
  c->h_blk_off.clear();
  c->h_blk_base.clear();
  return kOk;
}

}  // extern "C"
