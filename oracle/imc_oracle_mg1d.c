/*
 * imc_oracle_mg1d.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C FP64 restatement of the reference's multigroup 1-D spherical IMC
 * particle step, MG_IMC/MG_IMC1D.py: the equilibrium census (514-566),
 * material emission by (group, cell) bin (569-650), the half-Lambertian
 * sphere sources (653-676), fixed sources (679-743), and the group-aware
 * spherical tracker (_move_particle_spherical 277-365, _move_sph_slice_mg
 * 368-461).  The Planck-mean Fleck factor and emission energies are the
 * multigroup 2-D ones (om_fleck_mg, om_emission_mg in imc_oracle_mg.c;
 * MG_IMC1D.py:987-1002 and 600-612 compute the same quantities), and the
 * (cell, group) comb counts are orc_comb_counts of imc_oracle.c on bins
 * cell + group * I (MG_IMC1D.py:746-843).
 *
 * Differences from the reference, all deliberate:
 *  - Philox streams (philox_ref.h) replace np.random / random.random.
 *  - B_g is the Planck group integral from its series (om_bg); the
 *    reference calls the external planck_integrals module.
 *  - Particles of a (group, cell) bin are consecutive, bins in the
 *    reference's loop order (group outer, cell inner).
 *
 * Slot maps (stream index = position within the sampled block):
 *   equilibrium (tag 5, step 0): 0 radius, 1 mu, 2..6 Planck group at Tr
 *   emission (tag 2) / fixed source (tag 3): 0 radius, 1 mu, 2 time
 *   sphere face (tag 8 + side): 0 |mu| = sqrt(u), 1 time, 2..6 Planck group
 *   transport (tag 1, index = pool position): 4e scatter distance,
 *     4e+1 capped redraw, 4e+2 new mu (the gray 1-D tracker's layout)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "philox_ref.h"

#define M1_C 29.98
#define M1_A 0.01372
#define M1_PI 3.14159265358979323846

int32_t om_planck_group(orc_stream* st, uint32_t slot0, double T, int G, const double* edges, const double* cdf);
double orc_canonical_sum(const double* x, int64_t n);

/* 4/3 pi (r1^3 - r0^3) (MG_IMC1D.py:524, 920) */
static double m1_vol(double r0, double r1) { return (4.0 / 3.0) * M1_PI * (r1 * r1 * r1 - r0 * r0 * r0); }

void m1_volumes(int I, const double* edges, double* vol) {
  for (int i = 0; i < I; ++i) vol[i] = m1_vol(edges[i], edges[i + 1]);
}

/* Particle counts by rounding (np.round: half to even).  E[b] > 0 gets
 * max(min_one, round(N E[b] / E_total)); others 0 (MG_IMC1D.py:536-539
 * min_one = 1, 626-628 min_one = 1, 712-714 min_one = 0). */
int64_t m1_counts(int64_t n, const double* E, int64_t N, int min_one, int64_t* cnt) {
  double Et = orc_canonical_sum(E, n);
  int64_t tot = 0;
  for (int64_t b = 0; b < n; ++b) {
    int64_t m = 0;
    if (E[b] > 0.0 && Et > 0.0) {
      double d = nearbyint((double)N * E[b] / Et);
      if (d < (double)min_one) d = (double)min_one;
      m = (int64_t)d;
    }
    cnt[b] = m;
    tot += m;
  }
  return tot;
}

/* Equilibrium energy a Tr^4 V per cell, zero below T_emit_floor. */
void m1_equilibrium_energy(int I, const double* Tr, const double* vol, double T_emit_floor, double* E) {
  for (int i = 0; i < I; ++i) {
    double t = Tr[i];
    double e = M1_A * ((t * t) * (t * t)) * vol[i];
    if (T_emit_floor > 0.0 && t < T_emit_floor) e = 0.0;
    E[i] = e;
  }
}

static double m1_radius(double r0, double r1, double u) {
  return pow(r0 * r0 * r0 + (r1 * r1 * r1 - r0 * r0 * r0) * u, 1.0 / 3.0);
}

/* Equilibrium census (MG_IMC1D.py:514-566): counts per cell, uniform in
 * volume, isotropic, t = 0, group from the Planck spectrum at Tr. */
void m1_sample_ic(int I, const double* edges, const double* Tr, const double* E, const int64_t* cnt, int G,
                  const double* energy_edges, const double* pcdf, uint64_t seed, double* w, double* mu, double* t,
                  double* x, int32_t* cell, int32_t* grp) {
  int64_t k = 0;
  for (int i = 0; i < I; ++i) {
    int64_t n = cnt[i];
    for (int64_t q = 0; q < n; ++q, ++k) {
      orc_stream s;
      orc_stream_init(&s, seed, ORC_TAG_IC, (uint64_t)k, 0);
      w[k] = E[i] / (double)n;
      x[k] = m1_radius(edges[i], edges[i + 1], orc_u01(&s, 0));
      mu[k] = 2.0 * orc_u01(&s, 1) - 1.0;
      t[k] = 0.0;
      cell[k] = i;
      grp[k] = om_planck_group(&s, 2, Tr[i], G, energy_edges, pcdf);
    }
  }
}

/* Volume source by (group, cell) bin b = g I + i (emission 569-650, fixed
 * source 679-743): w = E_b / n_b, uniform in volume, isotropic, t ~ U(0,dt). */
void m1_sample_bins(int I, int G, const double* edges, const double* E, const int64_t* cnt, double dt,
                    uint64_t seed, uint32_t step, int32_t tag, double* w, double* mu, double* t, double* x,
                    int32_t* cell, int32_t* grp) {
  int64_t k = 0;
  for (int g = 0; g < G; ++g) {
    for (int i = 0; i < I; ++i) {
      int64_t b = (int64_t)g * I + i;
      int64_t n = cnt[b];
      for (int64_t q = 0; q < n; ++q, ++k) {
        orc_stream s;
        orc_stream_init(&s, seed, tag, (uint64_t)k, step);
        w[k] = E[b] / (double)n;
        x[k] = m1_radius(edges[i], edges[i + 1], orc_u01(&s, 0));
        mu[k] = 2.0 * orc_u01(&s, 1) - 1.0;
        t[k] = dt * orc_u01(&s, 2);
        cell[k] = i;
        grp[k] = g;
      }
    }
  }
}

/* Face energy a c T^4 / 4 dt 4 pi R^2 (MG_IMC1D.py:668). */
double m1_face_energy(double T, double dt, double R) {
  double t4 = (T * T) * (T * T);
  return M1_A * M1_C * t4 / 4.0 * dt * 4.0 * M1_PI * R * R;
}

/* Half-Lambertian sphere source (MG_IMC1D.py:653-676, placed by step()
 * 1061-1088): side 0 inner face, mu = +sqrt(u), cell 0; side 1 outer face,
 * mu = -sqrt(u), cell I-1; r = R; group from the Planck spectrum at T. */
void m1_sample_face(int side, int64_t n, double E, double T, double dt, double R, int I, int G,
                    const double* energy_edges, const double* pcdf, uint64_t seed, uint32_t step, double* w,
                    double* mu, double* t, double* x, int32_t* cell, int32_t* grp) {
  for (int64_t k = 0; k < n; ++k) {
    orc_stream s;
    orc_stream_init(&s, seed, ORC_TAG_BOUNDARY0 + side, (uint64_t)k, step);
    w[k] = E / (double)n;
    double m = sqrt(orc_u01(&s, 0));
    mu[k] = side == 0 ? m : -m;
    t[k] = dt * orc_u01(&s, 1);
    x[k] = R;
    cell[k] = side == 0 ? 0 : I - 1;
    grp[k] = om_planck_group(&s, 2, T, G, energy_edges, pcdf);
  }
}

/* ------------------------------------------------------------------------- */
/* transport (MG_IMC1D.py:277-461)                                            */
/* Raw outputs per bin b = g I + i: dep_raw += w (1 - e^{-sa s}) (and floor   */
/* kills' remaining weight), path_raw += w s.  bl[g]: weight escaping either  */
/* face.  stats: events, boundary, scatter, census, floor kills, reflections, */
/* non-positive distances (the reference's AssertionError).                   */
/* ------------------------------------------------------------------------- */
void m1_transport(int64_t n, double* w, double* mu, const double* times, double* x, int32_t* cell,
                  const int32_t* grp, int I, const double* edges, int G, const double* sigma_a,
                  const double* sigma_s, double dt, const int32_t* refl, double weight_floor, uint64_t seed,
                  uint32_t step, double* dep_raw, double* path_raw, double* bl, int64_t* stats_out,
                  int32_t nthreads) {
  int nt = 1;
#ifdef _OPENMP
  nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#endif
  int64_t nb = (int64_t)G * I;
  double* depT = (double*)calloc((size_t)nt * (size_t)nb, sizeof(double));
  double* pathT = (double*)calloc((size_t)nt * (size_t)nb, sizeof(double));
  double* blT = (double*)calloc((size_t)nt * (size_t)G, sizeof(double));
  int64_t* stT = (int64_t*)calloc((size_t)nt * 8, sizeof(int64_t));
#pragma omp parallel for schedule(dynamic, 256) num_threads(nt)
  for (int64_t k = 0; k < n; ++k) {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    double* dp = depT + (size_t)tid * nb;
    double* pp = pathT + (size_t)tid * nb;
    double* bp = blT + (size_t)tid * G;
    int64_t* st = stT + (size_t)tid * 8;
    orc_stream rs;
    orc_stream_init(&rs, seed, ORC_TAG_TRANSPORT, (uint64_t)k, step);
    const int g = grp[k];
    double dcen = (dt - times[k]) * M1_C;
    uint32_t ev = 0;
    while (dcen > 1e-15) {
      int loc = cell[k];
      if (loc < 0 || loc >= I) break;
      st[0] += 1;
      int64_t b = (int64_t)g * I + loc;
      double r0 = edges[loc], r1 = edges[loc + 1];
      double sa = sigma_a[b], ss = sigma_s[b];
      double wt = w[k], m = mu[k], r = x[k];
      double b2 = r * r * (1.0 - m * m);
      double dout = -r * m + sqrt(r1 * r1 - b2);
      double din = 1e30;
      double discin = r0 * r0 - b2;
      if (m < 0.0 && discin > 0.0) din = -r * m - sqrt(discin);
      double db = dout < din ? dout : din;
      int hit_outer = (dout <= din);
      double dr = r1 - r0;
      double dsc = 1e30;
      if (ss > 1e-10) dsc = -log(1.0 - orc_u01(&rs, 4 * ev + 0)) / ss;
      if (ss * dr > 10000.0) dsc = -log(1.0 - orc_u01(&rs, 4 * ev + 1)) / (10000.0 / dr);
      double s = db;
      if (dsc < s) s = dsc;
      if (dcen < s) s = dcen;
      if (!(s > 0.0)) {
        st[6] += 1;
        break;
      }
      double rn2 = r * r + 2.0 * r * s * m + s * s;
      double rn = rn2 > 0.0 ? sqrt(rn2) : 0.0;
      double mn = rn > 1e-15 ? (r * m + s) / rn : 1.0;
      mn = mn < -1.0 ? -1.0 : (mn > 1.0 ? 1.0 : mn);
      int newloc = 0, code;
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
      ev++;
      dcen -= s;
      double wf = exp(-sa * s);
      dp[b] += wt * (1.0 - wf);
      pp[b] += wt * s;
      w[k] = wt * wf;
      mu[k] = mn;
      x[k] = rn;
      cell[k] += newloc;
      if (code == 1) st[1] += 1;
      else if (code == 2) st[2] += 1;
      else st[3] += 1;
      if (cell[k] >= I) {
        if (refl[1]) {
          mu[k] = -mu[k];
          cell[k] = I - 1;
          st[5] += 1;
        } else {
          bp[g] += w[k];
          w[k] = 0.0;
          break;
        }
      } else if (cell[k] < 0) {
        if (refl[0]) {
          mu[k] = -mu[k];
          cell[k] = 0;
          st[5] += 1;
        } else {
          bp[g] += w[k];
          w[k] = 0.0;
          break;
        }
      }
      if (w[k] < weight_floor && w[k] > 0.0) {
        /* deposit the remainder locally, in any cell (MG_IMC1D.py:449-461) */
        st[4] += 1;
        dp[(int64_t)g * I + cell[k]] += w[k];
        w[k] = 0.0;
        dcen = 0.0;
      }
    }
  }
  for (int a = 0; a < 8; ++a) stats_out[a] = 0;
  for (int th = 0; th < nt; ++th) {
    for (int64_t b = 0; b < nb; ++b) {
      dep_raw[b] += depT[(size_t)th * nb + b];
      path_raw[b] += pathT[(size_t)th * nb + b];
    }
    for (int g = 0; g < G; ++g) bl[g] += blT[(size_t)th * G + g];
    for (int a = 0; a < 8; ++a) stats_out[a] += stT[th * 8 + a];
  }
  free(depT);
  free(pathT);
  free(blT);
  free(stT);
}
