/*
 * imc_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, FP64, history-based restatement of the reference's Fleck-Cummings
 * IMC particle step (DrRyanMc/RadTranBook, IMC/IMC2D.py and IMC/IMC1D.py).  It
 * is the checker the HIP library is compared against and the timed CPU
 * baseline; nothing in the shipped library links, loads or calls it.
 *
 * Random numbers come from the counter-based Philox stream of philox_ref.h
 * instead of the reference's hidden MT19937, so that a device history and an
 * oracle history with the same (seed, step, index) see the same numbers.
 *
 * Build: make -C oracle   (gcc -O2 -fopenmp -ffp-contract=off, no -ffast-math)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "philox_ref.h"

#define ORC_C 29.98
#define ORC_A 0.01372
#define ORC_PI 3.14159265358979323846

/* ------------------------------------------------------------------------- */
/* Philox export (pinned by known-answer vectors)                             */
/* ------------------------------------------------------------------------- */
void orc_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
  orc_philox4x32_10(ctr, key, out);
}

double orc_stream_u01(uint64_t seed, int32_t tag, uint64_t idx, uint32_t step, uint32_t slot) {
  orc_stream s;
  orc_stream_init(&s, seed, tag, idx, step);
  return orc_u01(&s, slot);
}

/* ------------------------------------------------------------------------- */
/* Canonical blocked sum: chunks of 1024 summed left to right, then the chunk */
/* sums reduced the same way.  Deterministic; the device uses the same order. */
/* ------------------------------------------------------------------------- */
double orc_canonical_sum(const double* x, int64_t n) {
  if (n <= 0) return 0.0;
  if (n <= 1024) {
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += x[i];
    return s;
  }
  int64_t nc = (n + 1023) / 1024;
  double* p = (double*)malloc(sizeof(double) * (size_t)nc);
  for (int64_t c = 0; c < nc; ++c) {
    int64_t lo = c * 1024, hi = lo + 1024 < n ? lo + 1024 : n;
    double s = 0.0;
    for (int64_t i = lo; i < hi; ++i) s += x[i];
    p[c] = s;
  }
  double r = orc_canonical_sum(p, nc);
  free(p);
  return r;
}

/* ------------------------------------------------------------------------- */
/* a1: Fleck factor and effective opacities                                   */
/* IMC/IMC2D.py:1646-1651 (clip), IMC/IMC1D.py:933-938 (assert in [0,1]).     */
/* ------------------------------------------------------------------------- */
int orc_fleck(int64_t n, const double* T, const double* sigma_true, const double* cv, double theta,
              double dt, int32_t clip, double* sigma_a, double* sigma_s, double* f_out) {
  int bad = 0;
  for (int64_t c = 0; c < n; ++c) {
    double t = T[c];
    double beta = 4.0 * ORC_A * (t * t * t) / cv[c];
    double f = 1.0 / (1.0 + theta * beta * sigma_true[c] * ORC_C * dt);
    if (clip) {
      if (f < 0.0) f = 0.0;
      if (f > 1.0) f = 1.0;
    } else if (!(f >= 0.0 && f <= 1.0)) {
      bad = 1;
    }
    if (f_out) f_out[c] = f;
    sigma_s[c] = sigma_true[c] * (1.0 - f);
    sigma_a[c] = sigma_true[c] * f;
  }
  return bad ? -1 : 0;
}

/* a4: E_c = a c T^4 sigma_a dt V with the Fleck-reduced sigma_a, T clamped at 0
 * (IMC/IMC2D.py:1732). */
void orc_emission_energy(int64_t n, const double* T, const double* sigma_a, const double* vol, double dt,
                         double* E) {
  for (int64_t c = 0; c < n; ++c) {
    double t = T[c] > 0.0 ? T[c] : 0.0;
    double t2 = t * t;
    double t4 = t2 * t2;
    E[c] = ORC_A * ORC_C * t4 * sigma_a[c] * dt * vol[c];
  }
}

/* Particle counts per cell: ceil(N E_c / E) (IMC/IMC2D.py:1205), optionally at
 * least one per cell (IMC/IMC2D.py:1115).  Returns the total. */
int64_t orc_counts(int64_t n, const double* E, int64_t N, int32_t min_one, int64_t* counts) {
  double Et = orc_canonical_sum(E, n);
  double denom = min_one ? (Et > 1e-300 ? Et : 1e-300) : Et;
  int64_t tot = 0;
  for (int64_t c = 0; c < n; ++c) {
    int64_t k;
    if (!min_one && !(Et > 0.0)) {
      k = 0;
    } else {
      double v = ceil((double)N * E[c] / denom);
      k = (int64_t)v;
      if (min_one && k < 1) k = 1;
    }
    counts[c] = k;
    tot += k;
  }
  return tot;
}

/* ------------------------------------------------------------------------- */
/* Volume sources (emission, fixed source, equilibrium census)                */
/* xy: IMC/IMC2D.py:1191-1238 and 1110-1146; rz uniform 1241-1290,            */
/* rz linear 1293-1445, rz equilibrium 1149-1188.                             */
/* kind: 0 timed (t ~ U(0,dt)), 1 equilibrium (t = 0)                         */
/* geometry: 0 xy, 1 rz-uniform, 2 rz-linear                                  */
/* ------------------------------------------------------------------------- */
static double orc_solve_r_linear_cdf(double r0, double r1, double a, double M, double xi) {
  double C = (a * 0.5 * (r1 * r1 - r0 * r0) + M / 3.0 * (r1 * r1 * r1 - r0 * r0 * r0));
  if (C <= 0.0) return sqrt(r0 * r0 + (r1 * r1 - r0 * r0) * xi);
  double rhs = xi * C + a * 0.5 * (r0 * r0) + M / 3.0 * (r0 * r0 * r0);
  double lo = r0, hi = r1;
  double r = sqrt(r0 * r0 + (r1 * r1 - r0 * r0) * xi);
  for (int it = 0; it < 50; ++it) {
    double g = a * 0.5 * r * r + M / 3.0 * (r * r * r) - rhs;
    if (fabs(g) < 1e-14 * (fabs(rhs) + 1e-30)) break;
    if (g < 0.0)
      lo = r;
    else
      hi = r;
    double dg = a * r + M * r * r;
    double rn;
    if (fabs(dg) > 1e-30)
      rn = r - g / dg;
    else
      rn = 0.5 * (lo + hi);
    if (rn <= lo || rn >= hi) rn = 0.5 * (lo + hi);
    if (fabs(rn - r) < 1e-14 * (r1 - r0 + 1e-30)) {
      r = rn;
      break;
    }
    r = rn;
  }
  return r;
}

static double orc_t4(double t) {
  double u = t > 0.0 ? t : 0.0;
  double u2 = u * u;
  return u2 * u2;
}

/* Per-cell T^4 slopes for the rz linear source (IMC/IMC2D.py:1364-1384). */
void orc_rz_slopes(int nx, int ny, const double* T, const double* e1, const double* e2, double* Mr, double* Mz) {
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) {
#define T4AT(ii, jj) orc_t4(T[(ii) * ny + (jj)])
#define RC(ii) (0.5 * (e1[(ii)] + e1[(ii) + 1]))
#define ZC(jj) (0.5 * (e2[(jj)] + e2[(jj) + 1]))
      double mr, mz;
      if (nx == 1)
        mr = 0.0;
      else if (i == 0)
        mr = (T4AT(i + 1, j) - T4AT(i, j)) / (RC(i + 1) - RC(i));
      else if (i == nx - 1)
        mr = (T4AT(i, j) - T4AT(i - 1, j)) / (RC(i) - RC(i - 1));
      else
        mr = (T4AT(i + 1, j) - T4AT(i - 1, j)) / (RC(i + 1) - RC(i - 1));
      if (ny == 1)
        mz = 0.0;
      else if (j == 0)
        mz = (T4AT(i, j + 1) - T4AT(i, j)) / (ZC(j + 1) - ZC(j));
      else if (j == ny - 1)
        mz = (T4AT(i, j) - T4AT(i, j - 1)) / (ZC(j) - ZC(j - 1));
      else
        mz = (T4AT(i, j + 1) - T4AT(i, j - 1)) / (ZC(j + 1) - ZC(j - 1));
      Mr[i * ny + j] = mr;
      Mz[i * ny + j] = mz;
#undef T4AT
#undef RC
#undef ZC
    }
}

void orc_sample_volume_2d(int nx, int ny, const double* e1, const double* e2, const double* QV,
                          const int64_t* counts, int32_t kind, int32_t geometry, const double* T, double dt,
                          uint64_t seed, uint32_t step, int32_t tag, double* w, double* d1, double* d2, double* t,
                          double* p1, double* p2, int32_t* ci, int32_t* cj) {
  int64_t ncell = (int64_t)nx * ny;
  int64_t* off = (int64_t*)malloc(sizeof(int64_t) * (size_t)(ncell + 1));
  off[0] = 0;
  for (int64_t c = 0; c < ncell; ++c) off[c + 1] = off[c] + counts[c];
  double* Mr = NULL;
  double* Mz = NULL;
  if (geometry == 2) {
    Mr = (double*)malloc(sizeof(double) * (size_t)ncell);
    Mz = (double*)malloc(sizeof(double) * (size_t)ncell);
    orc_rz_slopes(nx, ny, T, e1, e2, Mr, Mz);
  }
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t c = 0; c < ncell; ++c) {
    int64_t n = counts[c];
    if (n <= 0) continue;
    int i = (int)(c / ny), j = (int)(c % ny);
    double x0 = e1[i], x1 = e1[i + 1], y0 = e2[j], y1 = e2[j + 1];
    double wc = QV[c] / (double)n;
    /* rz linear per-cell reconstruction (IMC/IMC2D.py:1397-1418) */
    double a_r = 0.0, Mloc_r = 0.0, Pn = 0.5;
    if (geometry == 2) {
      double T4c = orc_t4(T[c]);
      double rcen = 0.5 * (x0 + x1), zcen = 0.5 * (y0 + y1);
      double dr = x1 - x0, dz = y1 - y0;
      a_r = T4c - Mr[c] * (rcen - x0);
      double t4r0 = a_r, t4r1 = a_r + Mr[c] * dr;
      if (t4r0 < 0.0 || t4r1 < 0.0) {
        a_r = T4c > 0.0 ? T4c : 0.0;
        Mloc_r = 0.0;
      } else {
        Mloc_r = Mr[c];
      }
      double a_z = T4c - Mz[c] * (zcen - y0);
      double t4z0 = a_z > 0.0 ? a_z : 0.0;
      double t4z1 = a_z + Mz[c] * dz;
      t4z1 = t4z1 > 0.0 ? t4z1 : 0.0;
      double den = t4z0 + t4z1;
      Pn = den > 0.0 ? t4z0 / den : 0.5;
    }
    for (int64_t q = 0; q < n; ++q) {
      int64_t k = off[c] + q;
      orc_stream s;
      orc_stream_init(&s, seed, tag, (uint64_t)k, step);
      w[k] = wc;
      t[k] = (kind == 1) ? 0.0 : dt * orc_u01(&s, 0);
      if (geometry == 0) {
        p1[k] = x0 + (x1 - x0) * orc_u01(&s, 1);
        p2[k] = y0 + (y1 - y0) * orc_u01(&s, 2);
        double uz = 2.0 * orc_u01(&s, 3) - 1.0;
        double phi = 2.0 * ORC_PI * orc_u01(&s, 4);
        double rxy = sqrt(fmax(0.0, 1.0 - uz * uz));
        d1[k] = rxy * cos(phi);
        d2[k] = rxy * sin(phi);
      } else if (geometry == 1) {
        p1[k] = sqrt(x0 * x0 + (x1 * x1 - x0 * x0) * orc_u01(&s, 1));
        p2[k] = y0 + (y1 - y0) * orc_u01(&s, 2);
        d2[k] = 2.0 * orc_u01(&s, 3) - 1.0;
        d1[k] = cos(2.0 * ORC_PI * orc_u01(&s, 4));
      } else {
        double dz = y1 - y0;
        p1[k] = orc_solve_r_linear_cdf(x0, x1, a_r, Mloc_r, orc_u01(&s, 1));
        double xb = orc_u01(&s, 2), xs = orc_u01(&s, 3);
        p2[k] = (xb <= Pn) ? (y1 - dz * sqrt(xs)) : (y0 + dz * sqrt(xs));
        d2[k] = 2.0 * orc_u01(&s, 4) - 1.0;
        d1[k] = cos(2.0 * ORC_PI * orc_u01(&s, 5));
      }
      ci[k] = i;
      cj[k] = j;
    }
  }
  free(off);
  free(Mr);
  free(Mz);
}

/* ------------------------------------------------------------------------- */
/* a3: boundary sources, IMC/IMC2D.py:157-319 (xy) and 322-489 (rz).          */
/* side: 0 left/rmin, 1 right/rmax, 2 bottom/zmin, 3 top/zmax.                */
/* prof_cum: optional running sum of per-face-cell emission (length nface);   */
/* each particle picks a face cell from it (a multinomial split, like         */
/* np.random.multinomial at IMC/IMC2D.py:206).  Returns the total energy.     */
/* ------------------------------------------------------------------------- */
static int orc_locate(const double* e, int n, double x) {
  /* searchsorted(edges, x, 'right') - 1, clamped into [0, n-1] */
  int lo = 0, hi = n + 1; /* first index with e[idx] > x */
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (e[mid] > x)
      hi = mid;
    else
      lo = mid + 1;
  }
  int idx = lo - 1;
  if (idx < 0) idx = 0;
  if (idx > n - 1) idx = n - 1;
  return idx;
}

double orc_boundary_energy(int32_t geometry, int32_t side, double Tb, double dt, const double* e1, int nx,
                           const double* e2, int ny) {
  double x0 = e1[0], x1 = e1[nx], y0 = e2[0], y1 = e2[ny];
  double area;
  if (geometry == 0)
    area = (side < 2) ? (y1 - y0) : (x1 - x0);
  else if (side < 2) {
    double R = side == 0 ? x0 : x1;
    area = 2.0 * ORC_PI * R * (y1 - y0);
  } else
    area = ORC_PI * (x1 * x1 - x0 * x0);
  return ORC_A * ORC_C * (Tb * Tb * Tb * Tb) / 4.0 * area * dt;
}

void orc_sample_boundary_2d(int32_t geometry, int32_t side, int64_t n, double E_total, double dt, const double* e1,
                            int nx, const double* e2, int ny, const double* prof_cum, uint64_t seed, uint32_t step,
                            double* w, double* d1, double* d2, double* t, double* p1, double* p2, int32_t* ci,
                            int32_t* cj) {
  double x0 = e1[0], x1 = e1[nx], y0 = e2[0], y1 = e2[ny];
  int nface = (side < 2) ? ny : nx;
  const double* fe = (side < 2) ? e2 : e1; /* edges along the face */
  for (int64_t k = 0; k < n; ++k) {
    orc_stream s;
    orc_stream_init(&s, seed, ORC_TAG_BOUNDARY0 + side, (uint64_t)k, step);
    w[k] = E_total / (double)n;
    t[k] = dt * orc_u01(&s, 0);
    /* face cell (profile) or whole face */
    double a = (side < 2) ? y0 : x0, b = (side < 2) ? y1 : x1;
    if (prof_cum) {
      double tot = prof_cum[nface - 1];
      double u = orc_u01(&s, 6) * tot;
      int m = 0;
      while (m < nface - 1 && !(u < prof_cum[m])) ++m;
      while (m > 0 && prof_cum[m] == prof_cum[m - 1]) --m; /* never land on an empty cell */
      a = fe[m];
      b = fe[m + 1];
    }
    if (geometry == 0) {
      double mun = sqrt(orc_u01(&s, 1));
      double mut = sqrt(fmax(0.0, 1.0 - mun * mun));
      double phi = 2.0 * ORC_PI * orc_u01(&s, 2);
      double along = a + (b - a) * orc_u01(&s, 3);
      if (side == 0) {
        d1[k] = mun; d2[k] = mut * cos(phi); p1[k] = x0 + 1e-12; p2[k] = along;
      } else if (side == 1) {
        d1[k] = -mun; d2[k] = mut * cos(phi); p1[k] = x1 - 1e-12; p2[k] = along;
      } else if (side == 2) {
        d2[k] = mun; d1[k] = mut * cos(phi); p2[k] = y0 + 1e-12; p1[k] = along;
      } else {
        d2[k] = -mun; d1[k] = mut * cos(phi); p2[k] = y1 - 1e-12; p1[k] = along;
      }
    } else {
      double eta = 2.0 * orc_u01(&s, 1) - 1.0;
      double muabs = sqrt(orc_u01(&s, 2));
      if (side == 0) {
        d1[k] = muabs; d2[k] = eta; p1[k] = x0 + 1e-12; p2[k] = a + (b - a) * orc_u01(&s, 3);
      } else if (side == 1) {
        d1[k] = -muabs; d2[k] = eta; p1[k] = x1 - 1e-12; p2[k] = a + (b - a) * orc_u01(&s, 3);
      } else {
        double e = sqrt(orc_u01(&s, 3));
        d2[k] = side == 2 ? e : -e;
        d1[k] = cos(2.0 * ORC_PI * orc_u01(&s, 4));
        p2[k] = side == 2 ? y0 + 1e-12 : y1 - 1e-12;
        p1[k] = sqrt(a * a + (b * b - a * a) * orc_u01(&s, 5));
      }
    }
    ci[k] = orc_locate(e1, nx, p1[k]);
    cj[k] = orc_locate(e2, ny, p2[k]);
  }
}

/* ------------------------------------------------------------------------- */
/* a8-a10: transport, IMC/IMC2D.py:711-1107                                   */
/* Draw slots for event e (0-based): 4e scatter distance, 4e+1 capped redraw, */
/* 4e+2 and 4e+3 the new direction.                                           */
/* ------------------------------------------------------------------------- */
enum { X_NONE = 0, X_IP = 1, X_IM = 2, X_JP = 3, X_JM = 4 };
enum { E_CENSUS = 0, E_BOUNDARY = 1, E_SCATTER = 2 };

static double orc_dist_radial(double r, double mu, double eta, double R) {
  double ome2 = 1.0 - eta * eta;
  if (ome2 < 1e-15) return 1e30;
  double b2 = r * r * (1.0 - mu * mu);
  double disc = R * R - b2;
  if (disc <= 0.0) return 1e30;
  double root = sqrt(disc);
  double lp = -r * mu + root, lm = -r * mu - root;
  double den = sqrt(ome2);
  double smin = 1e30;
  if (lp > 1e-15) {
    double s = lp / den;
    if (s < smin) smin = s;
  }
  if (lm > 1e-15) {
    double s = lm / den;
    if (s < smin) smin = s;
  }
  return smin;
}

typedef struct {
  double w, d1, d2, p1, p2, dep_w, dep_i, dcen, s;
  int crossed, evt;
} orc_move_out;

static void orc_scatter_dist(orc_stream* s, uint32_t ev, double ss, double cell_len, double* s_scat) {
  *s_scat = 1e30;
  if (ss > 1e-12) *s_scat = -log(1.0 - orc_u01(s, 4 * ev + 0)) / ss;
  if (ss * cell_len > 10000.0) *s_scat = -log(1.0 - orc_u01(s, 4 * ev + 1)) / (10000.0 / cell_len);
}

static void orc_absorb(orc_move_out* o, double w, double sa, double s) {
  double wf = exp(-sa * s);
  double di;
  if (sa > 1e-12)
    di = w * (1.0 - wf) / sa;
  else
    di = w * s;
  o->dep_i = di;
  o->dep_w = di * sa;
  o->w = w * wf;
}

static void orc_move_xy(orc_stream* st, uint32_t ev, double w, double ux, double uy, double x, double y, int i, int j,
                        const double* e1, const double* e2, double sa, double ss, double dcen, orc_move_out* o) {
  double xl = e1[i], xr = e1[i + 1], yl = e2[j], yr = e2[j + 1];
  double sx = 1e30, sy = 1e30;
  if (ux > 0.0)
    sx = (xr - x) / ux;
  else if (ux < 0.0)
    sx = (xl - x) / ux;
  if (uy > 0.0)
    sy = (yr - y) / uy;
  else if (uy < 0.0)
    sy = (yl - y) / uy;
  double cl = (xr - xl) < (yr - yl) ? (xr - xl) : (yr - yl);
  double s_scat;
  orc_scatter_dist(st, ev, ss, cl, &s_scat);
  double sb = sx < sy ? sx : sy;
  double se = sb;
  if (s_scat < se) se = s_scat;
  if (dcen < se) se = dcen;
  o->d1 = ux; o->d2 = uy; o->p1 = x; o->p2 = y; o->crossed = X_NONE; o->s = 0.0;
  if (se <= 0.0) {
    o->w = w; o->dep_w = 0.0; o->dep_i = 0.0; o->dcen = 0.0; o->evt = E_CENSUS;
    return;
  }
  o->s = se;
  double xn = x + ux * se, yn = y + uy * se;
  int evt = E_CENSUS;
  if (fabs(se - sb) < 1e-11) {
    evt = E_BOUNDARY;
    if (sx <= sy) {
      if (ux > 0.0) { xn = xr; o->crossed = X_IP; } else { xn = xl; o->crossed = X_IM; }
    } else {
      if (uy > 0.0) { yn = yr; o->crossed = X_JP; } else { yn = yl; o->crossed = X_JM; }
    }
  } else if (se == s_scat) {
    evt = E_SCATTER;
    double uz = 2.0 * orc_u01(st, 4 * ev + 2) - 1.0;
    double phi = 2.0 * ORC_PI * orc_u01(st, 4 * ev + 3);
    double rxy = sqrt(fmax(0.0, 1.0 - uz * uz));
    o->d1 = rxy * cos(phi);
    o->d2 = rxy * sin(phi);
  }
  o->p1 = xn; o->p2 = yn; o->evt = evt;
  o->dcen = dcen - se;
  orc_absorb(o, w, sa, se);
}

static void orc_move_rz(orc_stream* st, uint32_t ev, double w, double mu, double eta, double r, double z, int i, int j,
                        const double* e1, const double* e2, double sa, double ss, double dcen, orc_move_out* o) {
  double rin = e1[i], rout = e1[i + 1], zlo = e2[j], zhi = e2[j + 1];
  double s_rin = rin > 0.0 ? orc_dist_radial(r, mu, eta, rin) : 1e30;
  double s_rout = orc_dist_radial(r, mu, eta, rout);
  double s_zlo, s_zhi;
  if (eta > 1e-15) {
    s_zlo = 1e30; s_zhi = (zhi - z) / eta;
  } else if (eta < -1e-15) {
    s_zlo = (zlo - z) / eta; s_zhi = 1e30;
  } else {
    s_zlo = 1e30; s_zhi = 1e30;
  }
  double cl = (rout - rin) < (zhi - zlo) ? (rout - rin) : (zhi - zlo);
  double s_scat;
  orc_scatter_dist(st, ev, ss, cl, &s_scat);
  double sb = s_rin;
  if (s_rout < sb) sb = s_rout;
  if (s_zlo < sb) sb = s_zlo;
  if (s_zhi < sb) sb = s_zhi;
  double se = sb;
  if (s_scat < se) se = s_scat;
  if (dcen < se) se = dcen;
  o->d1 = mu; o->d2 = eta; o->p1 = r; o->p2 = z; o->crossed = X_NONE; o->s = 0.0;
  if (se <= 0.0) {
    o->w = w; o->dep_w = 0.0; o->dep_i = 0.0; o->dcen = 0.0; o->evt = E_CENSUS;
    return;
  }
  o->s = se;
  double ome2 = 1.0 - eta * eta;
  if (ome2 < 0.0) ome2 = 0.0;
  double l = se * sqrt(ome2);
  double rn2 = r * r + 2.0 * r * mu * l + l * l;
  if (rn2 < 0.0) rn2 = 0.0;
  double rn = sqrt(rn2);
  double zn = z + se * eta;
  double mun;
  if (rn > 1e-15 && l > 0.0) {
    mun = (r * mu + l) / rn;
    if (mun > 1.0) mun = 1.0; else if (mun < -1.0) mun = -1.0;
  } else {
    mun = mu;
  }
  double etan = eta;
  int evt = E_CENSUS;
  if (fabs(se - sb) < 1e-11) {
    evt = E_BOUNDARY;
    if (sb == s_rin) {
      rn = rin > 0.0 ? rin * (1.0 - 1e-12) : 1e-12;
      o->crossed = X_IM;
    } else if (sb == s_rout) {
      rn = rout * (1.0 + 1e-12);
      o->crossed = X_IP;
    } else if (sb == s_zlo) {
      zn = zlo;
      o->crossed = X_JM;
    } else {
      zn = zhi;
      o->crossed = X_JP;
    }
  } else if (se == s_scat) {
    evt = E_SCATTER;
    etan = 2.0 * orc_u01(st, 4 * ev + 2) - 1.0;
    mun = cos(2.0 * ORC_PI * orc_u01(st, 4 * ev + 3));
  }
  o->d1 = mun; o->d2 = etan; o->p1 = rn; o->p2 = zn; o->evt = evt;
  o->dcen = dcen - se;
  orc_absorb(o, w, sa, se);
}

/* dep, si: (nx,ny) C-order, zeroed by the caller.  stats[8]: events, boundary,
 * scatter, census, capture (always 0), floor kills, reflections, cap hits. */
void orc_transport_2d(int64_t n, double* w, double* d1, double* d2, const double* times, double* p1, double* p2,
                      int32_t* ci, int32_t* cj, const double* e1, const double* e2, int nx, int ny,
                      const double* sigma_a, const double* sigma_s, const double* vol, double dt,
                      const int32_t* reflect, int64_t max_events, int32_t geometry, double weight_floor,
                      uint64_t seed, uint32_t step, double* dep, double* si, double* bl_out, int64_t* stats_out,
                      int32_t nthreads) {
  int nt = 1;
#ifdef _OPENMP
  nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#endif
  int64_t ncell = (int64_t)nx * ny;
  double* depT = (double*)calloc((size_t)nt * (size_t)ncell, sizeof(double));
  double* siT = (double*)calloc((size_t)nt * (size_t)ncell, sizeof(double));
  double* blT = (double*)calloc((size_t)nt, sizeof(double));
  int64_t* stT = (int64_t*)calloc((size_t)nt * 8, sizeof(int64_t));

#pragma omp parallel for schedule(dynamic, 256) num_threads(nt)
  for (int64_t k = 0; k < n; ++k) {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    double* dp = depT + (size_t)tid * (size_t)ncell;
    double* sp = siT + (size_t)tid * (size_t)ncell;
    int64_t* stt = stT + (size_t)tid * 8;
    orc_stream st;
    orc_stream_init(&st, seed, ORC_TAG_TRANSPORT, (uint64_t)k, step);
    double dcen = (dt - times[k]) * ORC_C;
    int64_t events = 0;
    while (dcen > 0.0) {
      events += 1;
      stt[0] += 1;
      if (events > max_events) {
        stt[7] += 1;
        dcen = 0.0;
        break;
      }
      int i = ci[k], j = cj[k];
      if (i < 0 || i >= nx || j < 0 || j >= ny) break;
      int64_t c = (int64_t)i * ny + j;
      double sa = sigma_a[c], ss = sigma_s[c];
      orc_move_out o;
      uint32_t ev = (uint32_t)(events - 1);
      if (geometry == 0)
        orc_move_xy(&st, ev, w[k], d1[k], d2[k], p1[k], p2[k], i, j, e1, e2, sa, ss, dcen, &o);
      else
        orc_move_rz(&st, ev, w[k], d1[k], d2[k], p1[k], p2[k], i, j, e1, e2, sa, ss, dcen, &o);
      dcen = o.dcen;
      if (o.evt == E_BOUNDARY)
        stt[1] += 1;
      else if (o.evt == E_SCATTER)
        stt[2] += 1;
      else
        stt[3] += 1;
      double v = vol[c];
      dp[c] += o.dep_w / v;
      sp[c] += o.dep_i / (dt * v);
      w[k] = o.w;
      if (o.w <= weight_floor) {
        stt[5] += 1;
        break;
      }
      d1[k] = o.d1; d2[k] = o.d2; p1[k] = o.p1; p2[k] = o.p2;
      if (o.crossed != X_NONE) {
        if (o.crossed == X_IP) ci[k] += 1;
        else if (o.crossed == X_IM) ci[k] -= 1;
        else if (o.crossed == X_JP) cj[k] += 1;
        else cj[k] -= 1;
        int i2 = ci[k], j2 = cj[k];
        if (i2 < 0) {
          if (reflect[0]) { d1[k] = -d1[k]; ci[k] = 0; stt[6] += 1; }
          else { blT[tid] += w[k]; w[k] = 0.0; dcen = 0.0; }
        } else if (i2 >= nx) {
          if (reflect[1]) { d1[k] = -d1[k]; ci[k] = nx - 1; stt[6] += 1; }
          else { blT[tid] += w[k]; w[k] = 0.0; dcen = 0.0; }
        } else if (j2 < 0) {
          if (reflect[2]) { d2[k] = -d2[k]; cj[k] = 0; stt[6] += 1; }
          else { blT[tid] += w[k]; w[k] = 0.0; dcen = 0.0; }
        } else if (j2 >= ny) {
          if (reflect[3]) { d2[k] = -d2[k]; cj[k] = ny - 1; stt[6] += 1; }
          else { blT[tid] += w[k]; w[k] = 0.0; dcen = 0.0; }
        }
      }
      if (w[k] <= 1e-300) {
        w[k] = 0.0;
        dcen = 0.0;
      }
    }
  }
  double bl = 0.0;
  for (int t = 0; t < 8; ++t) stats_out[t] = 0;
  for (int th = 0; th < nt; ++th) {
    for (int64_t c = 0; c < ncell; ++c) {
      dep[c] += depT[(size_t)th * ncell + c];
      si[c] += siT[(size_t)th * ncell + c];
    }
    bl += blT[th];
    for (int t = 0; t < 8; ++t) stats_out[t] += stT[th * 8 + t];
  }
  *bl_out = bl;
  free(depT);
  free(siT);
  free(blT);
  free(stT);
}

/* ------------------------------------------------------------------------- */
/* Adjoint of the transport tallies (the differentiable step).                */
/* For adjoint weights lam (per cell, dL/dE_abs) it returns the derivatives   */
/* of  Lambda = sum over segments lam[cell] * dep  with respect to sigma_a    */
/* and sigma_s of every cell, by replaying each history twice from its input  */
/* state with the same streams:                                               */
/*  - continuous absorption (pathwise): segment j in cell c with length s_j   */
/*    gives gA[c] += -s_j (Lambda - P_j) + lam[c] w_{j+1} s_j, P_j = deposits */
/*    of segments <= j weighted by lam;                                       */
/*  - scatter sampling (score function): segment j gives                      */
/*    gS[c] += ([scatter] / sigma_s - s_j) (Lambda - P_{j-1}), zero where     */
/*    sigma_s <= 1e-12 or the opacity cap fixes the sampling rate;            */
/*  - emission weights proportional to sigma_a of the birth cell: particles   */
/*    k in [emit_lo, emit_hi) add Lambda / sigma_a[birth] to gA[birth].       */
/* The input arrays are the transport's input pool (unchanged).               */
/* ------------------------------------------------------------------------- */
static double orc_replay(int pass, int64_t k, double w0, double u1, double u2, const double* times, double x,
                         double y, int i, int j, const double* e1, const double* e2, int nx, int ny,
                         const double* sigma_a, const double* sigma_s, double dt, const int32_t* reflect,
                         int64_t max_events, int32_t geometry, double weight_floor, uint64_t seed, uint32_t step,
                         const double* lam, double Lam, double* gA, double* gS) {
  orc_stream st;
  orc_stream_init(&st, seed, ORC_TAG_TRANSPORT, (uint64_t)k, step);
  double dcen = (dt - times[k]) * ORC_C;
  double w = w0, P = 0.0;
  int64_t events = 0;
  while (dcen > 0.0) {
    events += 1;
    if (events > max_events) break;
    if (i < 0 || i >= nx || j < 0 || j >= ny) break;
    int64_t c = (int64_t)i * ny + j;
    double sa = sigma_a[c], ss = sigma_s[c];
    orc_move_out o;
    uint32_t ev = (uint32_t)(events - 1);
    if (geometry == 0)
      orc_move_xy(&st, ev, w, u1, u2, x, y, i, j, e1, e2, sa, ss, dcen, &o);
    else
      orc_move_rz(&st, ev, w, u1, u2, x, y, i, j, e1, e2, sa, ss, dcen, &o);
    dcen = o.dcen;
    double Pprev = P;
    P += lam[c] * o.dep_w;
    if (pass == 1) {
      double xl, xr, yl, yr;
      xl = e1[i]; xr = e1[i + 1]; yl = e2[j]; yr = e2[j + 1];
      double cl = (xr - xl) < (yr - yl) ? (xr - xl) : (yr - yl);
      gA[c] += -o.s * (Lam - P) + lam[c] * o.w * o.s;
      if (ss > 1e-12 && !(ss * cl > 10000.0))
        gS[c] += ((o.evt == E_SCATTER ? 1.0 / ss : 0.0) - o.s) * (Lam - Pprev);
    }
    w = o.w;
    if (o.w <= weight_floor) break;
    u1 = o.d1; u2 = o.d2; x = o.p1; y = o.p2;
    if (o.crossed != X_NONE) {
      if (o.crossed == X_IP) i += 1;
      else if (o.crossed == X_IM) i -= 1;
      else if (o.crossed == X_JP) j += 1;
      else j -= 1;
      if (i < 0) {
        if (reflect[0]) { u1 = -u1; i = 0; } else { w = 0.0; dcen = 0.0; }
      } else if (i >= nx) {
        if (reflect[1]) { u1 = -u1; i = nx - 1; } else { w = 0.0; dcen = 0.0; }
      } else if (j < 0) {
        if (reflect[2]) { u2 = -u2; j = 0; } else { w = 0.0; dcen = 0.0; }
      } else if (j >= ny) {
        if (reflect[3]) { u2 = -u2; j = ny - 1; } else { w = 0.0; dcen = 0.0; }
      }
    }
    if (w <= 1e-300) {
      w = 0.0;
      dcen = 0.0;
    }
  }
  return P;
}

void orc_transport_2d_adjoint(int64_t n, const double* w, const double* d1, const double* d2, const double* times,
                              const double* p1, const double* p2, const int32_t* ci, const int32_t* cj,
                              const double* e1, const double* e2, int nx, int ny, const double* sigma_a,
                              const double* sigma_s, double dt, const int32_t* reflect, int64_t max_events,
                              int32_t geometry, double weight_floor, uint64_t seed, uint32_t step, const double* lam,
                              int64_t emit_lo, int64_t emit_hi, double* gA, double* gS, double* Lam_out,
                              int32_t nthreads) {
  int nt = 1;
#ifdef _OPENMP
  nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#endif
  int64_t ncell = (int64_t)nx * ny;
  double* aT = (double*)calloc((size_t)nt * (size_t)ncell, sizeof(double));
  double* sT = (double*)calloc((size_t)nt * (size_t)ncell, sizeof(double));
#pragma omp parallel for schedule(dynamic, 256) num_threads(nt)
  for (int64_t k = 0; k < n; ++k) {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    double* ga = aT + (size_t)tid * (size_t)ncell;
    double* gs = sT + (size_t)tid * (size_t)ncell;
    double Lam = orc_replay(0, k, w[k], d1[k], d2[k], times, p1[k], p2[k], ci[k], cj[k], e1, e2, nx, ny, sigma_a,
                            sigma_s, dt, reflect, max_events, geometry, weight_floor, seed, step, lam, 0.0, ga, gs);
    orc_replay(1, k, w[k], d1[k], d2[k], times, p1[k], p2[k], ci[k], cj[k], e1, e2, nx, ny, sigma_a, sigma_s, dt,
               reflect, max_events, geometry, weight_floor, seed, step, lam, Lam, ga, gs);
    if (k >= emit_lo && k < emit_hi && ci[k] >= 0 && ci[k] < nx && cj[k] >= 0 && cj[k] < ny) {
      int64_t c = (int64_t)ci[k] * ny + cj[k];
      if (sigma_a[c] > 0.0) ga[c] += Lam / sigma_a[c];
    }
    if (Lam_out) Lam_out[k] = Lam;
  }
  for (int64_t c = 0; c < ncell; ++c) {
    double a = 0.0, b = 0.0;
    for (int th = 0; th < nt; ++th) {
      a += aT[(size_t)th * ncell + c];
      b += sT[(size_t)th * ncell + c];
    }
    gA[c] = a;
    gS[c] = b;
  }
  free(aT);
  free(sT);
}

/* ------------------------------------------------------------------------- */
/* Forward-mode likelihood-ratio gradient of the gray 2-D transport tallies   */
/* (reference for the replay adjoint).  For L = sum_c lam_c E_abs,c and a     */
/* direction (vA, vS) of per-cell perturbations of (sigma_a, sigma_s), one    */
/* forward pass per history accumulates                                       */
/*   X     = sum_j lam_{c_j} dep_j            (the history's contribution)    */
/*   dX/de = sum_j lam_{c_j} d dep_j / de     (paths held fixed: dep_j =      */
/*           w_{j-1} - w_j, w_j = w_{j-1} e^{-sigma_a s_j}, so d w_j / de =   */
/*           e^{-sigma_a s_j} d w_{j-1} / de - w_j s_j vA_{c_j})              */
/*   S     = sum_j vS_{c_j} ([scatter_j] / sigma_s - s_j), the derivative of  */
/*           the log density of the sampled path (segments whose scatter      */
/*           rate is not capped and sigma_s > 1e-12)                          */
/* with an emission-born history (k in [emit_lo, emit_hi)) starting from      */
/* d w_0 / de = w_0 vA_b / sigma_a,b (its weight is proportional to the birth */
/* cell's sigma_a).  E[sum_h dX/de + X S] is the directional derivative; the  */
/* two sums are returned separately (out[0] pathwise, out[1] score, out[2]    */
/* sum X).  The same histories and draw slots as orc_transport_2d.            */
/* ------------------------------------------------------------------------- */
static void orc_lr_history(int64_t k, double w0, double u1, double u2, const double* times, double x, double y,
                           int i, int j, const double* e1, const double* e2, int nx, int ny, const double* sigma_a,
                           const double* sigma_s, double dt, const int32_t* reflect, int64_t max_events,
                           int32_t geometry, double weight_floor, uint64_t seed, uint32_t step, const double* lam,
                           const double* vA, const double* vS, double dw0, double* X_out, double* dX_out,
                           double* S_out) {
  orc_stream st;
  orc_stream_init(&st, seed, ORC_TAG_TRANSPORT, (uint64_t)k, step);
  double dcen = (dt - times[k]) * ORC_C;
  double w = w0, dw = dw0, X = 0.0, dX = 0.0, S = 0.0;
  int64_t events = 0;
  while (dcen > 0.0) {
    events += 1;
    if (events > max_events) break;
    if (i < 0 || i >= nx || j < 0 || j >= ny) break;
    int64_t c = (int64_t)i * ny + j;
    double sa = sigma_a[c], ss = sigma_s[c];
    orc_move_out o;
    uint32_t ev = (uint32_t)(events - 1);
    if (geometry == 0)
      orc_move_xy(&st, ev, w, u1, u2, x, y, i, j, e1, e2, sa, ss, dcen, &o);
    else
      orc_move_rz(&st, ev, w, u1, u2, x, y, i, j, e1, e2, sa, ss, dcen, &o);
    dcen = o.dcen;
    /* weight derivative across the segment and the deposit's derivative */
    double dw_next = dw * exp(-sa * o.s) - o.w * o.s * vA[c];
    X += lam[c] * o.dep_w;
    dX += lam[c] * (dw - dw_next);
    double xl = e1[i], xr = e1[i + 1], yl = e2[j], yr = e2[j + 1];
    double cl = (xr - xl) < (yr - yl) ? (xr - xl) : (yr - yl);
    if (ss > 1e-12 && !(ss * cl > 10000.0)) S += vS[c] * ((o.evt == E_SCATTER ? 1.0 / ss : 0.0) - o.s);
    w = o.w;
    dw = dw_next;
    if (o.w <= weight_floor) break;
    u1 = o.d1; u2 = o.d2; x = o.p1; y = o.p2;
    if (o.crossed != X_NONE) {
      if (o.crossed == X_IP) i += 1;
      else if (o.crossed == X_IM) i -= 1;
      else if (o.crossed == X_JP) j += 1;
      else j -= 1;
      if (i < 0) {
        if (reflect[0]) { u1 = -u1; i = 0; } else { w = 0.0; dcen = 0.0; }
      } else if (i >= nx) {
        if (reflect[1]) { u1 = -u1; i = nx - 1; } else { w = 0.0; dcen = 0.0; }
      } else if (j < 0) {
        if (reflect[2]) { u2 = -u2; j = 0; } else { w = 0.0; dcen = 0.0; }
      } else if (j >= ny) {
        if (reflect[3]) { u2 = -u2; j = ny - 1; } else { w = 0.0; dcen = 0.0; }
      }
    }
    if (w <= 1e-300) {
      w = 0.0;
      dcen = 0.0;
    }
  }
  *X_out = X;
  *dX_out = dX;
  *S_out = S;
}

void orc_transport_2d_lr(int64_t n, const double* w, const double* d1, const double* d2, const double* times,
                         const double* p1, const double* p2, const int32_t* ci, const int32_t* cj, const double* e1,
                         const double* e2, int nx, int ny, const double* sigma_a, const double* sigma_s, double dt,
                         const int32_t* reflect, int64_t max_events, int32_t geometry, double weight_floor,
                         uint64_t seed, uint32_t step, const double* lam, const double* vA, const double* vS,
                         int64_t emit_lo, int64_t emit_hi, double* out, int32_t nthreads) {
  int nt = 1;
#ifdef _OPENMP
  nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#endif
  double path = 0.0, score = 0.0, xsum = 0.0;
#pragma omp parallel for schedule(dynamic, 256) num_threads(nt) reduction(+ : path, score, xsum)
  for (int64_t k = 0; k < n; ++k) {
    double dw0 = 0.0;
    if (k >= emit_lo && k < emit_hi && ci[k] >= 0 && ci[k] < nx && cj[k] >= 0 && cj[k] < ny) {
      int64_t b = (int64_t)ci[k] * ny + cj[k];
      if (sigma_a[b] > 0.0) dw0 = w[k] * vA[b] / sigma_a[b];
    }
    double X, dX, S;
    orc_lr_history(k, w[k], d1[k], d2[k], times, p1[k], p2[k], ci[k], cj[k], e1, e2, nx, ny, sigma_a, sigma_s, dt,
                   reflect, max_events, geometry, weight_floor, seed, step, lam, vA, vS, dw0, &X, &dX, &S);
    path += dX;
    score += X * S;
    xsum += X;
  }
  out[0] = path;
  out[1] = score;
  out[2] = xsum;
}

/* ------------------------------------------------------------------------- */
/* a18: per-cell stochastic comb, IMC/IMC2D.py:1448-1511, IMC/IMC1D.py:770-828 */
/* cell[k] is the flat cell id (any bijection; dead particles have w <= 0).   */
/* stream_base offsets the Philox index (problem << 32 in batched runs).      */
/* Fills ecen, ew (per cell) and copies (per particle); returns total copies. */
/* ------------------------------------------------------------------------- */
int64_t orc_comb_counts(int64_t n, const double* w, const int64_t* cell, int64_t ncell, int64_t Nmax,
                        uint64_t seed, uint32_t step, uint64_t stream_base, double* ecen, double* ew,
                        int64_t* copies, int32_t* passthrough) {
  *passthrough = 0;
  for (int64_t c = 0; c < ncell; ++c) ecen[c] = 0.0;
  for (int64_t k = 0; k < n; ++k)
    if (w[k] > 0.0) ecen[cell[k]] += w[k];
  double E = orc_canonical_sum(ecen, ncell);
  if (!(E > 0.0)) {
    for (int64_t c = 0; c < ncell; ++c) ew[c] = 0.0;
    int64_t tot = 0;
    for (int64_t k = 0; k < n; ++k) {
      copies[k] = w[k] > 0.0 ? 1 : 0;
      tot += copies[k];
    }
    *passthrough = 1; /* nothing to comb: survivors keep their weights */
    return tot;
  }
  for (int64_t c = 0; c < ncell; ++c) {
    if (ecen[c] > 0.0) {
      double d = nearbyint((double)Nmax * ecen[c] / E);
      if (d < 1.0) d = 1.0;
      ew[c] = ecen[c] / d;
    } else {
      ew[c] = 0.0;
    }
  }
  int64_t tot = 0;
  for (int64_t k = 0; k < n; ++k) {
    int64_t m = 0;
    if (w[k] > 0.0) {
      double et = ew[cell[k]];
      if (et > 0.0) {
        orc_stream s;
        orc_stream_init(&s, seed, ORC_TAG_COMB, stream_base + (uint64_t)k, step);
        m = (int64_t)(w[k] / et + orc_u01(&s, 0));
      }
    }
    copies[k] = m;
    tot += m;
  }
  return tot;
}

/* ------------------------------------------------------------------------- */
/* Carter-Forest variant, IMC/IMC2D_CarterForest.py.                          */
/* Emission times: the truncated exponential of 51-56, from slot 0 of the     */
/* particle's emission stream (the slot the fixed-source sampler uses for a   */
/* uniform time), so the whole emission is one sampler call plus this map.   */
/* Transport (103-508): analog absorption at the true sigma_a; an absorbed    */
/* particle is re-emitted isotropically after a delay -ln(xi)/(c sigma_a      */
/* beta) if that ends before census, else it is captured and its weight       */
/* deposited; every segment adds w s to the scalar-intensity tally; a face    */
/* crossing nudges the position 1e-12 cell widths past the face.              */
/* Draw slots for event e: 6e absorption distance, 6e+1 capped redraw, 6e+2   */
/* delay, 6e+3 fast-path capture test, 6e+4 and 6e+5 the new direction.       */
/* ------------------------------------------------------------------------- */
double orc_cf_time(double rate, double dt, double xi) {
  if (rate <= 1e-12) return xi * dt;
  return -log(1.0 - xi * (1.0 - exp(-rate * dt))) / rate;
}

void orc_cf_emission_times(int64_t n, const int32_t* ci, const int32_t* cj, int ny, const double* rate, double dt,
                           uint64_t seed, uint32_t step, int32_t tag, double* t) {
  for (int64_t k = 0; k < n; ++k) {
    orc_stream s;
    orc_stream_init(&s, seed, tag, (uint64_t)k, step);
    t[k] = orc_cf_time(rate[(int64_t)ci[k] * ny + cj[k]], dt, orc_u01(&s, 0));
  }
}

enum { CF_CENSUS = 0, CF_BOUNDARY = 1, CF_ABSORB = 2, CF_CAPTURE = 4 };

typedef struct {
  double w, d1, d2, p1, p2, t, dep_w, dep_i, dcen;
  int crossed, evt;
} orc_cf_out;

static double orc_cf_exp_dist(orc_stream* st, uint32_t slot, double rate) {
  double u = 1.0 - orc_u01(st, slot);
  return -log(u > 1e-16 ? u : 1e-16) / rate;
}

/* the absorption branch shared by both geometries (IMC/IMC2D_CarterForest.py:
 * 183-213, 317-343): o already holds the moved particle; returns with evt,
 * w, dep_w, dcen, t and (on re-emission) the new direction set */
static void orc_cf_absorb(orc_stream* st, uint32_t ev, int geom, double w, double sa, double beta, double se,
                          double tl, double eps_s, double fastpath, orc_cf_out* o) {
  if (beta > 1e-15 && sa > 1e-15) {
    double rate = ORC_C * sa * beta;
    double trem = o->dcen / ORC_C;
    int iso = 0;
    if (fastpath > 0.0 && trem > 1e-14 && rate * trem > fastpath) {
      double pcap = exp(-rate * trem);
      if (orc_u01(st, 6 * ev + 3) < pcap) {
        o->w = 0.0; o->dep_w = w; o->dcen = 0.0; o->t = tl; o->evt = CF_CAPTURE;
        return;
      }
      iso = 1;
      o->t = tl;
      o->evt = CF_ABSORB;
    } else {
      double xi = orc_u01(st, 6 * ev + 2);
      if (xi < 1e-15) xi = 1e-15;
      double delay = -log(xi) / rate;
      if (delay < trem - 1e-15) {
        iso = 1;
        o->t = tl + se / ORC_C + delay;
        double dc = (trem - delay) * ORC_C;
        o->dcen = dc > 0.0 ? dc : 0.0;
        if (o->dcen <= eps_s) {
          o->dcen = 0.0;
          o->evt = CF_CENSUS;
        } else {
          o->evt = CF_ABSORB;
        }
      }
    }
    if (iso) {
      o->w = w; o->dep_w = 0.0;
      if (geom == 0) {
        double uz = 2.0 * orc_u01(st, 6 * ev + 4) - 1.0;
        double phi = 2.0 * ORC_PI * orc_u01(st, 6 * ev + 5);
        double rxy = sqrt(fmax(0.0, 1.0 - uz * uz));
        o->d1 = rxy * cos(phi);
        o->d2 = rxy * sin(phi);
      } else {
        o->d2 = 2.0 * orc_u01(st, 6 * ev + 4) - 1.0;
        o->d1 = cos(2.0 * ORC_PI * orc_u01(st, 6 * ev + 5));
      }
      return;
    }
  }
  o->w = 0.0; o->dep_w = w; o->dcen = 0.0; o->t = tl; o->evt = CF_CAPTURE;
}

static void orc_cf_move_xy(orc_stream* st, uint32_t ev, double w, double ux, double uy, double x, double y, int i,
                           int j, const double* e1, const double* e2, double sa, double beta, double dcen, double tl,
                           double fastpath, orc_cf_out* o) {
  double xl = e1[i], xr = e1[i + 1], yl = e2[j], yr = e2[j + 1];
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
  double s_abs = sa > 1e-12 ? orc_cf_exp_dist(st, 6 * ev, sa) : 1e30;
  double cl = (xr - xl) < (yr - yl) ? (xr - xl) : (yr - yl);
  double eps_s = fmax(1e-14, 1e-12 * cl), eps_x = fmax(1e-14, 1e-12 * (xr - xl));
  double eps_y = fmax(1e-14, 1e-12 * (yr - yl));
  if (sa * cl > 10000.0) s_abs = orc_cf_exp_dist(st, 6 * ev + 1, 10000.0 / cl);
  double se = sb;
  if (s_abs < se) se = s_abs;
  if (dcen < se) se = dcen;
  if (fabs(se - sb) < 1e-11 && se < eps_s && dcen > eps_s) se = eps_s;
  o->w = w; o->d1 = ux; o->d2 = uy; o->p1 = x; o->p2 = y; o->t = tl; o->dep_w = 0.0; o->dep_i = 0.0;
  o->crossed = X_NONE;
  if (se <= 0.0) {
    o->dcen = 0.0; o->evt = CF_CENSUS;
    return;
  }
  o->p1 = x + ux * se;
  o->p2 = y + uy * se;
  o->dep_i = w * se;
  o->dcen = dcen - se;
  if (fabs(se - sb) < 1e-11) {
    if (sx <= sy) {
      if (ux > 0.0) { o->p1 = xr + eps_x; o->crossed = X_IP; } else { o->p1 = xl - eps_x; o->crossed = X_IM; }
    } else {
      if (uy > 0.0) { o->p2 = yr + eps_y; o->crossed = X_JP; } else { o->p2 = yl - eps_y; o->crossed = X_JM; }
    }
    o->evt = CF_BOUNDARY;
    return;
  }
  if (fabs(se - s_abs) < 1e-11) {
    orc_cf_absorb(st, ev, 0, w, sa, beta, se, tl, eps_s, fastpath, o);
    return;
  }
  o->dcen = 0.0; o->evt = CF_CENSUS;
}

static void orc_cf_move_rz(orc_stream* st, uint32_t ev, double w, double mu, double eta, double r, double z, int i,
                           int j, const double* e1, const double* e2, double sa, double beta, double dcen, double tl,
                           double fastpath, orc_cf_out* o) {
  double rin = e1[i], rout = e1[i + 1], zlo = e2[j], zhi = e2[j + 1];
  double s_rin = rin > 0.0 ? orc_dist_radial(r, mu, eta, rin) : 1e30;
  double s_rout = orc_dist_radial(r, mu, eta, rout);
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
  double s_abs = sa > 1e-12 ? orc_cf_exp_dist(st, 6 * ev, sa) : 1e30;
  double cl = (rout - rin) < (zhi - zlo) ? (rout - rin) : (zhi - zlo);
  double eps_s = fmax(1e-14, 1e-12 * cl), eps_r = fmax(1e-14, 1e-12 * (rout - rin));
  double eps_z = fmax(1e-14, 1e-12 * (zhi - zlo));
  if (sa * cl > 10000.0) s_abs = orc_cf_exp_dist(st, 6 * ev + 1, 10000.0 / cl);
  double se = sb;
  if (s_abs < se) se = s_abs;
  if (dcen < se) se = dcen;
  if (fabs(se - sb) < 1e-11 && se < eps_s && dcen > eps_s) se = eps_s;
  o->w = w; o->d1 = mu; o->d2 = eta; o->p1 = r; o->p2 = z; o->t = tl; o->dep_w = 0.0; o->dep_i = 0.0;
  o->crossed = X_NONE;
  if (se <= 0.0) {
    o->dcen = 0.0; o->evt = CF_CENSUS;
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
  o->d1 = mun; o->p1 = rn; o->p2 = zn;
  o->dep_i = w * se;
  o->dcen = dcen - se;
  if (fabs(se - sb) < 1e-11) {
    if (sb == s_rin) {
      double rr = rin - eps_r;
      o->p1 = rin > 0.0 ? (rr > 1e-12 ? rr : 1e-12) : 1e-12;
      o->crossed = X_IM;
    } else if (sb == s_rout) {
      o->p1 = rout + eps_r; o->crossed = X_IP;
    } else if (sb == s_zlo) {
      o->p2 = zlo - eps_z; o->crossed = X_JM;
    } else {
      o->p2 = zhi + eps_z; o->crossed = X_JP;
    }
    o->evt = CF_BOUNDARY;
    return;
  }
  if (fabs(se - s_abs) < 1e-11) {
    orc_cf_absorb(st, ev, 1, w, sa, beta, se, tl, eps_s, fastpath, o);
    return;
  }
  o->dcen = 0.0; o->evt = CF_CENSUS;
}

/* _transport_particles_2d_cf (IMC/IMC2D_CarterForest.py:346-508).  dep: raw
 * captured energy per cell; si: raw sum of w s per cell (the caller divides
 * by dt V); both (nx,ny) C-order, zeroed by the caller.  stats[8]: events,
 * boundary, absorb-continue, census, capture, floor kills, reflections, cap
 * hits.  A weight that falls to the floor is deposited in its cell (the
 * reference drops it unbooked).  times are updated in place. */
void orc_transport_2d_cf(int64_t n, double* w, double* d1, double* d2, double* times, double* p1, double* p2,
                         int32_t* ci, int32_t* cj, const double* e1, const double* e2, int nx, int ny,
                         const double* sigma_a, const double* beta, double dt, const int32_t* reflect,
                         int64_t max_events, int32_t geometry, double weight_floor, double fastpath,
                         double no_progress_tol, int64_t no_progress_max, uint64_t seed, uint32_t step, double* dep,
                         double* si, double* bl_out, int64_t* stats_out, int32_t nthreads) {
  int nt = 1;
#ifdef _OPENMP
  nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#endif
  int64_t ncell = (int64_t)nx * ny;
  double* depT = (double*)calloc((size_t)nt * (size_t)ncell, sizeof(double));
  double* siT = (double*)calloc((size_t)nt * (size_t)ncell, sizeof(double));
  double* blT = (double*)calloc((size_t)nt, sizeof(double));
  int64_t* stT = (int64_t*)calloc((size_t)nt * 8, sizeof(int64_t));

#pragma omp parallel for schedule(dynamic, 256) num_threads(nt)
  for (int64_t k = 0; k < n; ++k) {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    double* dp = depT + (size_t)tid * (size_t)ncell;
    double* sp = siT + (size_t)tid * (size_t)ncell;
    int64_t* stt = stT + (size_t)tid * 8;
    orc_stream st;
    orc_stream_init(&st, seed, ORC_TAG_TRANSPORT, (uint64_t)k, step);
    double dcen = (dt - times[k]) * ORC_C;
    int64_t events = 0, stuck = 0;
    while (dcen > 0.0 && w[k] > weight_floor) {
      double dcen_before = dcen;
      int i = ci[k], j = cj[k];
      if (i < 0 || i >= nx || j < 0 || j >= ny) break;
      events += 1;
      stt[0] += 1;
      if (max_events > 0 && events > max_events) {
        stt[7] += 1;
        break;
      }
      int64_t c = (int64_t)i * ny + j;
      orc_cf_out o;
      uint32_t ev = (uint32_t)(events - 1);
      if (geometry == 0)
        orc_cf_move_xy(&st, ev, w[k], d1[k], d2[k], p1[k], p2[k], i, j, e1, e2, sigma_a[c], beta[c], dcen, times[k],
                       fastpath, &o);
      else
        orc_cf_move_rz(&st, ev, w[k], d1[k], d2[k], p1[k], p2[k], i, j, e1, e2, sigma_a[c], beta[c], dcen, times[k],
                       fastpath, &o);
      dp[c] += o.dep_w;
      sp[c] += o.dep_i;
      w[k] = o.w; d1[k] = o.d1; d2[k] = o.d2; p1[k] = o.p1; p2[k] = o.p2; times[k] = o.t;
      dcen = o.dcen;
      if (o.evt == CF_BOUNDARY) stt[1] += 1;
      else if (o.evt == CF_ABSORB) stt[2] += 1;
      else if (o.evt == CF_CAPTURE) stt[4] += 1;
      else stt[3] += 1;
      if (dcen >= dcen_before * (1.0 - no_progress_tol)) {
        stuck += 1;
        if (stuck >= no_progress_max) break;
      } else {
        stuck = 0;
      }
      if (o.crossed != X_NONE) {
        if (o.crossed == X_IP) ci[k] += 1;
        else if (o.crossed == X_IM) ci[k] -= 1;
        else if (o.crossed == X_JP) cj[k] += 1;
        else cj[k] -= 1;
        int i2 = ci[k], j2 = cj[k];
        if (i2 < 0) {
          if (reflect[0]) { d1[k] = -d1[k]; ci[k] = 0; stt[6] += 1; }
          else { blT[tid] += w[k]; w[k] = 0.0; dcen = 0.0; }
        } else if (i2 >= nx) {
          if (reflect[1]) { d1[k] = -d1[k]; ci[k] = nx - 1; stt[6] += 1; }
          else { blT[tid] += w[k]; w[k] = 0.0; dcen = 0.0; }
        } else if (j2 < 0) {
          if (reflect[2]) { d2[k] = -d2[k]; cj[k] = 0; stt[6] += 1; }
          else { blT[tid] += w[k]; w[k] = 0.0; dcen = 0.0; }
        } else if (j2 >= ny) {
          if (reflect[3]) { d2[k] = -d2[k]; cj[k] = ny - 1; stt[6] += 1; }
          else { blT[tid] += w[k]; w[k] = 0.0; dcen = 0.0; }
        }
      }
      if (w[k] <= weight_floor && w[k] > 0.0) {
        stt[5] += 1;
        dp[(int64_t)ci[k] * ny + cj[k]] += w[k];
        w[k] = 0.0;
        dcen = 0.0;
      }
    }
  }
  double bl = 0.0;
  for (int t = 0; t < 8; ++t) stats_out[t] = 0;
  for (int th = 0; th < nt; ++th) {
    for (int64_t c = 0; c < ncell; ++c) {
      dep[c] += depT[(size_t)th * ncell + c];
      si[c] += siT[(size_t)th * ncell + c];
    }
    bl += blT[th];
    for (int t = 0; t < 8; ++t) stats_out[t] += stT[th * 8 + t];
  }
  *bl_out = bl;
  free(depT);
  free(siT);
  free(blT);
  free(stT);
}
