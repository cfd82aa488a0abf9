/*
 * imc_oracle_1d.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C FP64 restatement of the reference's gray 1-D IMC particle step,
 * IMC/IMC1D.py: slab (move_particle 75-130, move_particles 133-200,
 * sample_linear_density 203-217, equilibrium_sample 220-248,
 * emitted_particles 251-291, sample_source 294-316, create_boundary 319-329)
 * and spherical (move_particle_spherical 336-436, _move_sph_slice 439-497,
 * equilibrium_sample_spherical 535-575, _solve_spherical_cdf 578-626,
 * emitted_particles_spherical 629-707, sample_source_spherical 710-741,
 * create_boundary_spherical 744-763).
 *
 * Random numbers: the Philox streams of philox_ref.h with index
 * (problem << 32) | local, so problem p of a batched device run is reproduced
 * by passing stream_base = p << 32.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "philox_ref.h"

#define O1_C 29.98
#define O1_A 0.01372
#define O1_PI 3.14159265358979323846

/* cell "volume": dx (slab) or 4/3 pi (r1^3 - r0^3) (IMC/IMC1D.py:853-857) */
static double o1_vol(int sph, double r0, double r1) {
  return sph ? (4.0 / 3.0) * O1_PI * (r1 * r1 * r1 - r0 * r0 * r0) : (r1 - r0);
}

/* ------------------------------------------------------------------------- */
/* sourcing                                                                   */
/* ------------------------------------------------------------------------- */
/* T-slopes for the slab linear tilt (IMC/IMC1D.py:254-262). */
void o1_slab_slopes(int I, const double* T, const double* edges, double* slopes) {
  for (int i = 0; i < I; ++i) {
    double dx = edges[i + 1] - edges[i];
    double s;
    if (I == 1)
      s = 0.0;
    else if (i == 0)
      s = (T[1] - T[0]) / ((edges[2] - edges[1]) + dx) * 2.0;
    else if (i == I - 1)
      s = (T[I - 1] - T[I - 2]) / (dx + (edges[I - 1] - edges[I - 2])) * 2.0;
    else
      s = (T[i + 1] - T[i - 1]) / ((edges[i + 2] - edges[i + 1]) + (edges[i] - edges[i - 1])) * 2.0;
    double lv = T[i] - s * dx / 2.0, rv = T[i] + s * dx / 2.0;
    if (lv < 0.0 || rv < 0.0) s = 0.0;
    slopes[i] = s;
  }
}

/* sample_linear_density for one uniform xi (IMC/IMC1D.py:203-217) */
static double o1_linear_density(double dx, double s, double T, double xi) {
  if (fabs(s) < 1e-6) return dx * xi;
  double smin = -2.0 * T / (dx * dx), smax = 2.0 * T / (dx * dx);
  if (s < smin) s = smin;
  if (s > smax) s = smax;
  double q = (s / (2.0 * T) - 1.0 / dx);
  return (-2.0 * T + s * dx + 2.0 * T * dx * sqrt(q * q + (2.0 * s * xi) / (T * dx))) / (2.0 * s);
}

/* T^4 slopes in r for the spherical tilt (IMC/IMC1D.py:643-659) */
void o1_sph_slopes(int I, const double* T, const double* edges, double* M) {
  for (int i = 0; i < I; ++i) {
    double t4 = T[i] * T[i] * T[i] * T[i];
    double m;
    if (I == 1) {
      m = 0.0;
    } else {
#define RM(k) (0.5 * (edges[(k)] + edges[(k) + 1]))
#define T4(k) (T[(k)] * T[(k)] * T[(k)] * T[(k)])
      if (i == 0)
        m = (T4(1) - T4(0)) / (RM(1) - RM(0));
      else if (i == I - 1)
        m = (T4(I - 1) - T4(I - 2)) / (RM(I - 1) - RM(I - 2));
      else
        m = (T4(i + 1) - T4(i - 1)) / (RM(i + 1) - RM(i - 1));
#undef RM
#undef T4
    }
    double r0 = edges[i], r1 = edges[i + 1], rm = 0.5 * (r0 + r1);
    double a = t4 - m * (rm - r0), b = t4 + m * (r1 - rm);
    if (a < 0.0 || b < 0.0) m = 0.0;
    M[i] = m;
  }
}

static double o1_solve_sph_cdf(double r0, double r1, double a, double M, double xi) {
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

/* Volume source in one problem.  kind: 0 emission (linear tilt, t ~ U(0,dt)),
 * 1 equilibrium (uniform, t = 0), 2 fixed source (uniform; slab times
 * U(0,1) as IMC/IMC1D.py:313, spherical U(0,dt) as IMC/IMC1D.py:737).
 * Slot map: 0 time, 1 position, 2 mu.
 */
void o1_sample_volume(int sph, int I, const double* edges, const double* QV, const int64_t* counts, int32_t kind,
                      const double* T, double dt, uint64_t seed, uint32_t step, int32_t tag, uint64_t stream_base,
                      double* w, double* mu, double* t, double* x, int32_t* cell) {
  double* sl = (double*)malloc(sizeof(double) * (size_t)I);
  if (kind == 0) {
    if (sph)
      o1_sph_slopes(I, T, edges, sl);
    else
      o1_slab_slopes(I, T, edges, sl);
  }
  int64_t off = 0;
  for (int i = 0; i < I; ++i) {
    int64_t n = counts[i];
    double r0 = edges[i], r1 = edges[i + 1], dx = r1 - r0;
    double wc = n > 0 ? QV[i] / (double)n : 0.0;
    double a_c = 0.0;
    int uniform_sph = 1;
    if (kind == 0 && sph) {
      double t4 = T[i] * T[i] * T[i] * T[i];
      a_c = t4 - sl[i] * (0.5 * (r0 + r1) - r0);
      uniform_sph = (fabs(sl[i]) < 1e-30 || fabs(a_c) < 1e-30);
    }
    for (int64_t q = 0; q < n; ++q) {
      int64_t k = off + q;
      orc_stream s;
      orc_stream_init(&s, seed, tag, stream_base + (uint64_t)k, step);
      w[k] = wc;
      double ut = orc_u01(&s, 0), ux = orc_u01(&s, 1), um = orc_u01(&s, 2);
      if (kind == 1)
        t[k] = 0.0;
      else if (kind == 2 && !sph)
        t[k] = ut; /* U(0,1): IMC/IMC1D.py:313 */
      else
        t[k] = dt * ut;
      if (!sph) {
        x[k] = (kind == 0) ? o1_linear_density(dx, sl[i], T[i], ux) + r0 : r0 + dx * ux;
      } else if (kind == 0 && !uniform_sph) {
        x[k] = o1_solve_sph_cdf(r0, r1, a_c, sl[i], ux);
      } else {
        x[k] = pow(r0 * r0 * r0 + (r1 * r1 * r1 - r0 * r0 * r0) * ux, 1.0 / 3.0);
      }
      mu[k] = 2.0 * um - 1.0;
      cell[k] = i;
    }
    off += n;
  }
  free(sl);
}

/* Boundary source (IMC/IMC1D.py:319-329, 744-763 and their placement in
 * step() 947-975).  side 0: left face, mu > 0, cell 0 at x = 1e-8 (slab) or
 * r = R_in (sphere); side 1: right face, mu < 0, cell I-1 at x = R_out.
 * Slot map: 0 mu, 1 time. */
double o1_boundary_energy(int sph, double Tb, double dt, double R) {
  if (!sph) return O1_A * O1_C * (Tb * Tb * Tb * Tb) / 4.0 * dt;
  return O1_A * O1_C * (Tb * Tb * Tb * Tb) / 4.0 * dt * 4.0 * O1_PI * R * R;
}

void o1_sample_boundary(int sph, int side, int64_t n, double E, double dt, int I, const double* edges, uint64_t seed,
                        uint32_t step, uint64_t stream_base, double* w, double* mu, double* t, double* x,
                        int32_t* cell) {
  for (int64_t k = 0; k < n; ++k) {
    orc_stream s;
    orc_stream_init(&s, seed, ORC_TAG_BOUNDARY0 + side, stream_base + (uint64_t)k, step);
    w[k] = E / (double)n;
    double m = sqrt(orc_u01(&s, 0));
    t[k] = dt * orc_u01(&s, 1);
    if (side == 0) {
      mu[k] = m;
      x[k] = sph ? edges[0] : 1e-8;
      cell[k] = 0;
    } else {
      mu[k] = -m;
      x[k] = edges[I];
      cell[k] = I - 1;
    }
  }
}

/* ------------------------------------------------------------------------- */
/* transport (IMC/IMC1D.py:75-200 slab, 336-497 spherical)                    */
/* Slots for event e: 4e scatter distance, 4e+1 capped redraw, 4e+2 new mu.   */
/* stats[7]: events, boundary, scatter, census, floor kills, reflections,     */
/*           assertion failures.                                              */
/* Outputs per cell (I): dep, si; leak: boundary loss of escaped particles;   */
/* floor_leak: weight dropped by floor kills outside the deposit cells.       */
/* ------------------------------------------------------------------------- */
void o1_transport(int sph, int64_t n, double* w, double* mu, const double* times, double* x, int32_t* cell, int I,
                  const double* edges, const double* sigma_a, const double* sigma_s, double dt, const int32_t* refl,
                  double weight_floor, uint64_t seed, uint32_t step, uint64_t stream_base, double* dep, double* si,
                  double* leak_out, double* floor_leak_out, int64_t* stats_out, int32_t nthreads) {
  int nt = 1;
#ifdef _OPENMP
  nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#endif
  double* depT = (double*)calloc((size_t)nt * (size_t)I, sizeof(double));
  double* siT = (double*)calloc((size_t)nt * (size_t)I, sizeof(double));
  double* lkT = (double*)calloc((size_t)nt * 2, sizeof(double));
  int64_t* stT = (int64_t*)calloc((size_t)nt * 8, sizeof(int64_t));
#pragma omp parallel for schedule(dynamic, 256) num_threads(nt)
  for (int64_t k = 0; k < n; ++k) {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    double* dp = depT + (size_t)tid * I;
    double* sp = siT + (size_t)tid * I;
    int64_t* st = stT + (size_t)tid * 8;
    orc_stream rs;
    orc_stream_init(&rs, seed, ORC_TAG_TRANSPORT, stream_base + (uint64_t)k, step);
    double dcen = (dt - times[k]) * O1_C;
    uint32_t ev = 0;
    while (dcen > 0.0) {
      int loc = cell[k];
      st[0] += 1;
      double r0 = edges[loc], r1 = edges[loc + 1];
      double sa = sigma_a[loc], ss = sigma_s[loc];
      double wt = w[k], m = mu[k], r = x[k];
      double dxc = r1 - r0;
      double db, dsc = 1e30;
      int hit_outer = 1;
      if (!sph) {
        db = (m > 0) ? (r1 - r) / m : (r0 - r) / m;
      } else {
        double b2 = r * r * (1.0 - m * m);
        double dout = -r * m + sqrt(r1 * r1 - b2);
        double din = 1e30;
        double discin = r0 * r0 - b2;
        if (m < 0.0 && discin > 0.0) din = -r * m - sqrt(discin);
        db = dout < din ? dout : din;
        hit_outer = (dout <= din);
      }
      if (ss > 1e-10) dsc = -log(1.0 - orc_u01(&rs, 4 * ev + 0)) / ss;
      if (ss * dxc > 10000) dsc = -log(1.0 - orc_u01(&rs, 4 * ev + 1)) / (10000.0 / dxc);
      double s = db;
      if (dsc < s) s = dsc;
      if (dcen < s) s = dcen;
      if (!(s > 0.0)) {
        st[6] += 1; /* AssertionError in the reference */
        dcen = 0.0;
        ev++;
        break;
      }
      int newloc = 0, code;
      double rn, mn = m;
      if (!sph) {
        rn = r + m * s;
        if (fabs(db - s) < 1e-10) {
          if (m > 0) { rn = r1; newloc = 1; } else { rn = r0; newloc = -1; }
          code = 1;
        } else if (s == dsc) {
          mn = 2.0 * orc_u01(&rs, 4 * ev + 2) - 1.0;
          code = 2;
        } else {
          code = 3;
        }
      } else {
        double rn2 = r * r + 2.0 * r * s * m + s * s;
        rn = rn2 > 0.0 ? sqrt(rn2) : 0.0;
        mn = rn > 1e-15 ? (r * m + s) / rn : 1.0;
        mn = mn < -1.0 ? -1.0 : (mn > 1.0 ? 1.0 : mn);
        if (fabs(db - s) < 1e-10) {
          if (hit_outer) {
            rn = r1;
            mn = r1 > 1e-15 ? (r * m + s) / r1 : 1.0;
            newloc = 1;
          } else {
            rn = r0;
            mn = r0 > 1e-15 ? (r * m + s) / r0 : -1.0;
            newloc = -1;
          }
          mn = mn < -1.0 ? -1.0 : (mn > 1.0 ? 1.0 : mn);
          code = 1;
        } else if (s == dsc) {
          mn = 2.0 * orc_u01(&rs, 4 * ev + 2) - 1.0;
          code = 2;
        } else {
          code = 3;
        }
      }
      ev++;
      double norm = sph ? o1_vol(1, r0, r1) : dxc;
      dcen -= s;
      double wf = exp(-sa * s);
      double di = (sa > 1e-10) ? wt * (1.0 - wf) / sa / norm : wt * s / norm;
      double dw = di * sa;
      w[k] = wt * wf;
      mu[k] = mn;
      x[k] = rn;
      cell[k] += newloc;
      dp[loc] += dw;
      sp[loc] += di / dt;
      if (code == 1) st[1] += 1;
      else if (code == 2) st[2] += 1;
      else st[3] += 1;
      if (cell[k] == I) {
        if (refl[1]) { mu[k] = -mu[k]; cell[k] -= 1; st[5] += 1; }
        else dcen = 0.0;
      } else if (cell[k] == -1) {
        if (refl[0]) { mu[k] = -mu[k]; cell[k] += 1; st[5] += 1; }
        else dcen = 0.0;
      }
      if (w[k] < weight_floor) {
        st[4] += 1;
        int l2 = cell[k];
        if (0 < l2 && l2 < I - 1) {
          double nrm = sph ? o1_vol(1, edges[l2], edges[l2 + 1]) : (edges[l2 + 1] - edges[l2]);
          dp[l2] += w[k] / nrm;
          sp[l2] += w[k] / dt / sigma_a[l2] / nrm;
        } else {
          lkT[tid * 2 + 1] += w[k];
        }
        w[k] = 0.0;
        dcen = 0.0;
      }
    }
    if (cell[k] < 0 || cell[k] >= I) {
      lkT[tid * 2 + 0] += w[k];
      w[k] = 0.0;
    }
  }
  double lk = 0.0, fl = 0.0;
  for (int a = 0; a < 8; ++a) stats_out[a] = 0;
  for (int th = 0; th < nt; ++th) {
    for (int i = 0; i < I; ++i) {
      dep[i] += depT[(size_t)th * I + i];
      si[i] += siT[(size_t)th * I + i];
    }
    lk += lkT[th * 2];
    fl += lkT[th * 2 + 1];
    for (int a = 0; a < 8; ++a) stats_out[a] += stT[th * 8 + a];
  }
  *leak_out = lk;
  *floor_leak_out = fl;
  free(depT);
  free(siT);
  free(lkT);
  free(stT);
}

/* Equilibrium counts of IMC/IMC1D.py:231-232: ceil(E_i / E * N), no minimum;
 * empty when N <= 0 or E <= 0.  Returns the total. */
extern double orc_canonical_sum(const double* x, int64_t n);
int64_t o1_counts_ic(int I, const double* E, int64_t N, int64_t* counts) {
  double Et = orc_canonical_sum(E, I);
  int64_t tot = 0;
  for (int i = 0; i < I; ++i) {
    int64_t k = 0;
    if (N > 0 && Et > 0.0) k = (int64_t)ceil(E[i] / Et * (double)N);
    counts[i] = k;
    tot += k;
  }
  return tot;
}

/* ------------------------------------------------------------------------- */
/* Carter-Forest variant, IMC/IMC1D_CarterForest.py                           */
/* move_particle_cf 64-207, move_particles_cf 210-332 (slab);                 */
/* move_particle_spherical_cf 615-720, move_particles_spherical_cf 723-819.   */
/* Draw slots for event e: 6e absorption distance, 6e+1 capped redraw, 6e+2   */
/* delay, 6e+3 fast-path test, 6e+4 new direction.  A history whose weight is */
/* at or below the floor is not moved and stays in census (counted as a floor */
/* kill).  Outputs per cell: dep_raw = captured weight, path_raw = sum w s.   */
/* stats[8]: events, boundary, absorb-continue, census, capture, floor,       */
/* reflections, event-cap hits.                                               */
/* ------------------------------------------------------------------------- */

/* Truncated-exponential emission time (IMC/IMC1D_CarterForest.py:531-539). */
double o1_cf_time(double rate, double dt, double xi) {
  if (rate > 1e-10) return -log(1.0 - xi * (1.0 - exp(-rate * dt))) / rate;
  return xi * dt;
}

/* Times of an emission batch written by o1_sample_volume (same seed, step,
 * tag, stream base): slot 0 of each particle's stream through o1_cf_time. */
void o1_cf_emission_times(int64_t n, const int32_t* cell, const double* rate, double dt, uint64_t seed,
                          uint32_t step, int32_t tag, uint64_t stream_base, double* t) {
  for (int64_t k = 0; k < n; ++k) {
    orc_stream s;
    orc_stream_init(&s, seed, tag, stream_base + (uint64_t)k, step);
    t[k] = o1_cf_time(rate[cell[k]], dt, orc_u01(&s, 0));
  }
}

/* One Carter-Forest segment; returns the event code (1 boundary, 2 absorb and
 * continue, 3 census, 4 capture) and updates the particle in place. */
static int o1_cf_segment(int sph, orc_stream* rs, uint32_t e, double* w, double* mu, double* x, double* t,
                         double r0, double r1, double sa, double beta, double* dcen, double fastpath, int* newloc,
                         double* dep_w, double* path) {
  *newloc = 0;
  *dep_w = 0.0;
  *path = 0.0;
  if (*dcen < 1e-14) {
    *dcen = 0.0;
    return 3;
  }
  const double m = *mu, r = *x;
  const double width = r1 - r0;
  double db;
  int outer = 1;
  if (!sph) {
    if (m > 0.0)
      db = fmax(0.0, (r1 - r) / m);
    else if (m < 0.0)
      db = fmax(0.0, (r0 - r) / m);
    else
      db = 1e30;
  } else {
    double b2 = r * r * (1.0 - m * m);
    double dout = -r * m + sqrt(fmax(r1 * r1 - b2, 0.0));
    double din = 1e30;
    double dd = r0 * r0 - b2;
    if (m < 0.0 && dd > 0.0) din = -r * m - sqrt(dd);
    db = din < dout ? din : dout;
    outer = dout <= din;
  }
  double da = 1e30;
  if (sa > 1e-10) da = -log(1.0 - orc_u01(rs, 6 * e + 0)) / sa;
  if (sa * width > 10000) da = -log(1.0 - orc_u01(rs, 6 * e + 1)) / (10000.0 / width);
  double s = db;
  if (da < s) s = da;
  if (*dcen < s) s = *dcen;
  if (s <= 0.0) {
    *dcen = 0.0;
    return 3;
  }
  double xn, mn = m;
  if (!sph) {
    xn = r + m * s;
  } else {
    double q = r * r + 2.0 * r * s * m + s * s;
    xn = q > 0.0 ? sqrt(q) : 0.0;
    mn = xn > 1e-15 ? (r * m + s) / xn : 1.0;
    mn = fmax(-1.0, fmin(1.0, mn));
  }
  *path = *w * s;
  int code = 3;
  if (fabs(db - s) < 1e-10) {
    code = 1;
    if (!sph) {
      if (m > 0.0) { xn = r1; *newloc = 1; } else { xn = r0; *newloc = -1; }
    } else if (outer) {
      xn = r1;
      mn = fmax(-1.0, fmin(1.0, r1 > 1e-15 ? (r * m + s) / r1 : 1.0));
      *newloc = 1;
    } else {
      xn = r0;
      mn = fmax(-1.0, fmin(1.0, r0 > 1e-15 ? (r * m + s) / r0 : -1.0));
      *newloc = -1;
    }
    *dcen -= s;
  } else if (fabs(da - s) < 1e-10) {
    code = 2;
    if (beta > 1e-15 && sa > 1e-15) {
      double rate = O1_C * sa * beta;
      double tr = (*dcen - s) / O1_C;
      int cap;
      if (fastpath > 0.0 && tr > 1e-14 && rate * tr > fastpath) {
        cap = orc_u01(rs, 6 * e + 3) < exp(-rate * tr);
        if (!cap) {
          mn = -1.0 + 2.0 * orc_u01(rs, 6 * e + 4);
          *dcen -= s;
        }
      } else {
        double xi = orc_u01(rs, 6 * e + 2);
        if (xi < 1e-15) xi = 1e-15;
        double delay = -log(xi) / rate;
        cap = !(tr > 1e-14 && delay < tr);
        if (!cap) {
          mn = -1.0 + 2.0 * orc_u01(rs, 6 * e + 4);
          *t = (*t + s / O1_C) + delay;
          *dcen = (tr - delay) * O1_C;
        }
      }
      if (cap) {
        *dep_w = *w;
        *w = 0.0;
        code = 4;
      }
    } else {
      mn = -1.0 + 2.0 * orc_u01(rs, 6 * e + 4);
      *dcen -= s;
    }
  } else {
    *dcen = 0.0;
  }
  *x = xn;
  *mu = mn;
  return code;
}

void o1_transport_cf(int sph, int64_t n, double* w, double* mu, double* times, double* x, int32_t* cell, int I,
                     const double* edges, const double* sigma_a, const double* beta, double dt, const int32_t* refl,
                     double weight_floor, int64_t max_events, double fastpath, uint64_t seed, uint32_t step,
                     uint64_t stream_base, double* dep_raw, double* path_raw, double* leak_out, int64_t* stats_out,
                     int32_t nthreads) {
  int nt = 1;
#ifdef _OPENMP
  nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#endif
  double* depT = (double*)calloc((size_t)nt * (size_t)I, sizeof(double));
  double* pathT = (double*)calloc((size_t)nt * (size_t)I, sizeof(double));
  double* lkT = (double*)calloc((size_t)nt, sizeof(double));
  int64_t* stT = (int64_t*)calloc((size_t)nt * 8, sizeof(int64_t));
#pragma omp parallel for schedule(dynamic, 256) num_threads(nt)
  for (int64_t k = 0; k < n; ++k) {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    int64_t* st = stT + (size_t)tid * 8;
    orc_stream rs;
    orc_stream_init(&rs, seed, ORC_TAG_TRANSPORT, stream_base + (uint64_t)k, step);
    double dcen = (dt - times[k]) * O1_C;
    if (!(w[k] > weight_floor)) {
      if (w[k] > 0.0) st[5] += 1;
      continue;
    }
    int64_t events = 0;
    while (dcen > 0.0) {
      st[0] += 1;
      events += 1;
      if (max_events > 0 && events >= max_events) {
        st[7] += 1;
        dcen = 0.0;
        break;
      }
      int loc = cell[k];
      if (loc < 0) {
        if (refl[0]) { cell[k] = loc = 0; mu[k] = -mu[k]; } else break;
      } else if (loc >= I) {
        if (refl[1]) { cell[k] = loc = I - 1; mu[k] = -mu[k]; } else break;
      }
      int newloc;
      double dw, pth;
      int code = o1_cf_segment(sph, &rs, (uint32_t)(events - 1), &w[k], &mu[k], &x[k], &times[k], edges[loc],
                               edges[loc + 1], sigma_a[loc], beta[loc], &dcen, fastpath, &newloc, &dw, &pth);
      cell[k] += newloc;
      depT[(size_t)tid * I + loc] += dw;
      pathT[(size_t)tid * I + loc] += pth;
      if (code == 1) st[1] += 1;
      else if (code == 2) st[2] += 1;
      else if (code == 4) st[4] += 1;
      else st[3] += 1;
      if (code == 4) break;
      if (cell[k] == I) {
        if (refl[1]) { mu[k] = -mu[k]; cell[k] -= 1; st[6] += 1; } else break;
      } else if (cell[k] == -1) {
        if (refl[0]) { mu[k] = -mu[k]; cell[k] += 1; st[6] += 1; } else break;
      }
    }
    if (cell[k] < 0 || cell[k] >= I) {
      lkT[tid] += w[k];
      w[k] = 0.0;
    }
  }
  for (int a = 0; a < 8; ++a) stats_out[a] = 0;
  double lk = 0.0;
  for (int th = 0; th < nt; ++th) {
    for (int i = 0; i < I; ++i) {
      dep_raw[i] += depT[(size_t)th * I + i];
      path_raw[i] += pathT[(size_t)th * I + i];
    }
    lk += lkT[th];
    for (int a = 0; a < 8; ++a) stats_out[a] += stT[th * 8 + a];
  }
  *leak_out = lk;
  free(depT);
  free(pathT);
  free(lkT);
  free(stT);
}
