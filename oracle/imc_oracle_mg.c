/*
 * imc_oracle_mg.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C FP64 restatement of the reference's multigroup 2-D IMC particle
 * step, MG_IMC/MG_IMC2D.py: Planck mixture-of-gammas group sampling (99-101,
 * 202-254), group fractions b*_g and the Planck-mean Fleck factor
 * (1693-1719, 1879-1885), boundary sources (392-577), the equilibrium census
 * (992-1095), fixed sources (1098-1234), material emission (1866-1943),
 * the MG tracker (580-989) and the (cell, group) comb (1270-1378).
 *
 * Differences from the reference, all deliberate:
 *  - Philox streams (philox_ref.h) replace np.random; every multinomial
 *    np.random.choice becomes an i.i.d. inverse-CDF draw per particle
 *    (the same distribution).  The material emission's cells come from the
 *    order statistics of its N uniforms (exponential spacings), so they are
 *    drawn in non-decreasing cell order; the multiset of draws has the same
 *    distribution.
 *  - B_g is the Planck group integral computed here from its series; the
 *    reference calls the external planck_integrals module, absent here
 *    (its gray fallback would make b*_g = 1/G).
 *  - Right/top (xy) and rmax/zmax (rz) face particles start in the last
 *    cell; the reference's searchsorted places them one past the mesh
 *    (MG_IMC2D.py:175-178), where they are lost (xy) or dropped (rz).
 *  - A reflecting left/rmin face resets the cell index to 0 (the reference
 *    leaves it at -1 and loses the particle next event, 921-930).
 *  - Comb streams are indexed by position in the transported pool (escaped
 *    particles carry zero weight and receive no copies).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "philox_ref.h"

#define OM_C 29.98
#define OM_A 0.01372
#define OM_PI 3.14159265358979323846
#define OM_NMAX 200

/* ------------------------------------------------------------------------- */
/* Planck spectrum: mixture of gammas (MG_IMC2D.py:99-101, 202-254)           */
/* ------------------------------------------------------------------------- */
/* CDF of p(n) ~ 1/n^4, n = 1..200: sequential sums in double. */
void om_planck_cdf(double* cdf) {
  double inv[OM_NMAX];
  double s = 0.0;
  for (int n = 1; n <= OM_NMAX; ++n) {
    double d = (double)n;
    inv[n - 1] = 1.0 / (d * d * d * d);
    s += inv[n - 1];
  }
  double c = 0.0;
  for (int n = 0; n < OM_NMAX; ++n) {
    c += inv[n] / s;
    cdf[n] = c;
  }
}

/* Group of a photon drawn from the Planck spectrum at T.  Uses slots
 * slot0 (mixture index) and slot0+1..slot0+4 (four exponentials). */
int32_t om_planck_group(orc_stream* st, uint32_t slot0, double T, int G, const double* edges, const double* cdf) {
  if (!(T > 0.0)) return 0;
  double xi = orc_u01(st, slot0);
  int lo = 0, hi = OM_NMAX - 1; /* first i with xi <= cdf[i]; else n = 200 */
  if (!(xi <= cdf[OM_NMAX - 1])) {
    lo = OM_NMAX - 1;
  } else {
    while (lo < hi) {
      int mid = (lo + hi) >> 1;
      if (xi <= cdf[mid])
        hi = mid;
      else
        lo = mid + 1;
    }
  }
  double ns = (double)(lo + 1);
  double r1 = -log(orc_u01(st, slot0 + 1) + 1e-300);
  double r2 = -log(orc_u01(st, slot0 + 2) + 1e-300);
  double r3 = -log(orc_u01(st, slot0 + 3) + 1e-300);
  double r4 = -log(orc_u01(st, slot0 + 4) + 1e-300);
  double x = (r1 + r2 + r3 + r4) / ns;
  double nu = T * x;
  if (nu < edges[0]) return 0;
  if (nu >= edges[G]) return G - 1;
  int a = 0, b = G - 1; /* last g with edges[g] <= nu */
  while (a < b) {
    int mid = (a + b + 1) >> 1;
    if (edges[mid] <= nu)
      a = mid;
    else
      b = mid - 1;
  }
  return a;
}

/* ------------------------------------------------------------------------- */
/* Planck group integral B_g(T) = a c T^4/(4 pi) (15/pi^4) int x^3/(e^x-1)    */
/* over [e_lo/T, e_hi/T]: Bernoulli series below x = 1, exponential series    */
/* above.                                                                     */
/* ------------------------------------------------------------------------- */
static const double om_bern[10] = {
    /* B_2m / ((2m)! (2m+3)), m = 1..10 */
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

double om_planck_head(double x) { /* int_0^x t^3/(e^t-1) dt, 0 <= x < 1 */
  double x2 = x * x, x3 = x2 * x;
  double s = 0.0, p = x3 * x2; /* x^(2m+3), m = 1 */
  for (int m = 0; m < 10; ++m) {
    s += om_bern[m] * p;
    p *= x2;
  }
  return x3 / 3.0 - x3 * x / 8.0 + s;
}

double om_planck_tail(double x) { /* int_x^inf t^3/(e^t-1) dt, x >= 1 */
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

/* (15/pi^4) int_{xlo}^{xhi} t^3/(e^t-1) dt, 0 <= xlo <= xhi */
double om_planck_fraction(double xlo, double xhi) {
  const double full = OM_PI * OM_PI * OM_PI * OM_PI / 15.0;
  double v;
  if (xlo >= 1.0)
    v = om_planck_tail(xlo) - om_planck_tail(xhi);
  else if (xhi < 1.0)
    v = om_planck_head(xhi) - om_planck_head(xlo);
  else
    v = (full - om_planck_tail(xhi)) - om_planck_head(xlo);
  return v / full;
}

double om_bg(double T, double elo, double ehi) {
  if (!(T > 0.0)) return 0.0;
  double t4 = (T * T) * (T * T);
  return OM_A * OM_C / (4.0 * OM_PI) * t4 * om_planck_fraction(elo / T, ehi / T);
}

/* ------------------------------------------------------------------------- */
/* Fleck factor with the Planck-mean opacity (MG_IMC2D.py:1693-1719) and the  */
/* emission group fractions b*_g (1879-1885).  sigma: (G, ncell), c = i*ny+j. */
/* ------------------------------------------------------------------------- */
void om_fleck_mg(int64_t ncell, int G, const double* edges, const double* T, const double* sigma, const double* cv,
                 double theta, double dt, double* sa, double* ss, double* bstar, double* f_out) {
  double* B = (double*)malloc(sizeof(double) * (size_t)G);
  for (int64_t c = 0; c < ncell; ++c) {
    double bs = 0.0, num = 0.0;
    for (int g = 0; g < G; ++g) {
      B[g] = om_bg(T[c], edges[g], edges[g + 1]);
      bs += B[g];
      num += sigma[(int64_t)g * ncell + c] * B[g];
    }
    double sP = num / (bs + 1e-300);
    double t = T[c];
    double beta = 4.0 * OM_A * (t * t * t) / cv[c];
    double f = 1.0 / (1.0 + theta * beta * sP * OM_C * dt);
    if (f < 0.0) f = 0.0;
    if (f > 1.0) f = 1.0;
    if (f_out) f_out[c] = f;
    for (int g = 0; g < G; ++g) {
      double s = sigma[(int64_t)g * ncell + c];
      ss[(int64_t)g * ncell + c] = s * (1.0 - f);
      sa[(int64_t)g * ncell + c] = s * f;
      bstar[(int64_t)g * ncell + c] = B[g] / (bs + 1e-300);
    }
  }
  free(B);
}

/* Emission by group sigma_a,g b*_g a c T^4 dt V (MG_IMC2D.py:1888-1894); cells
 * below T_emit_floor emit nothing.  Returns per-cell totals in tot. */
void om_emission_mg(int64_t ncell, int G, const double* T, const double* sa, const double* bstar, const double* vol,
                    double dt, double T_emit_floor, double* E, double* tot) {
  for (int64_t c = 0; c < ncell; ++c) {
    double t = T[c];
    double t4 = (t * t) * (t * t);
    double s = 0.0;
    for (int g = 0; g < G; ++g) {
      int64_t q = (int64_t)g * ncell + c;
      double e = sa[q] * bstar[q] * OM_A * OM_C * t4 * dt * vol[c];
      if (T_emit_floor > 0.0 && t < T_emit_floor) e = 0.0;
      E[q] = e;
      s += e;
    }
    tot[c] = s;
  }
}

/* Canonical inclusive prefix: sequential within chunks of 1024, chunk
 * offsets summed sequentially; cdf[n-1] equals the canonical sum for
 * n <= 2^20. */
void om_prefix(const double* x, int64_t n, double* cdf) {
  double off = 0.0;
  for (int64_t c0 = 0; c0 < n; c0 += 1024) {
    int64_t c1 = c0 + 1024 < n ? c0 + 1024 : n;
    double s = 0.0;
    for (int64_t i = c0; i < c1; ++i) {
      s += x[i];
      cdf[i] = off + s;
    }
    off += s;
  }
}

/* Running sum in three levels, each left to right: 1024-element chunks,
 * chunk offsets within groups of 1024 chunks, group offsets;
 * S[i] = (group offset + chunk offset) + local prefix (the device's
 * prefix_three_level). */
void om_prefix3(const double* x, int64_t n, double* S) {
  const int64_t nch = (n + 1023) / 1024, ngroup = (nch + 1023) / 1024;
  double* tot = (double*)calloc((size_t)(nch > 0 ? nch : 1), sizeof(double));
  double* choff = (double*)calloc((size_t)(nch > 0 ? nch : 1), sizeof(double));
  double* gtot = (double*)calloc((size_t)(ngroup > 0 ? ngroup : 1), sizeof(double));
  if (!tot || !choff || !gtot) abort();
  for (int64_t b = 0; b < nch; ++b) {
    const int64_t c0 = b * 1024, c1 = c0 + 1024 < n ? c0 + 1024 : n;
    double s = 0.0;
    for (int64_t i = c0; i < c1; ++i) {
      s += x[i];
      S[i] = s;
    }
    tot[b] = s;
  }
  for (int64_t q = 0; q < ngroup; ++q) {
    const int64_t b0 = q * 1024, b1 = b0 + 1024 < nch ? b0 + 1024 : nch;
    double o = 0.0;
    for (int64_t b = b0; b < b1; ++b) {
      choff[b] = o;
      o += tot[b];
    }
    gtot[q] = o;
  }
  for (int64_t i = 0; i < n; ++i) {
    const int64_t b = i >> 10, q = b >> 10;
    double go = 0.0;
    for (int64_t r = 0; r < q; ++r) go += gtot[r];
    S[i] = (go + choff[b]) + S[i];
  }
  free(tot);
  free(choff);
  free(gtot);
}

/* first index with cdf[index] > u (np.random.choice), clamped to n-1 */
static int64_t om_pick(const double* cdf, int64_t n, double u) {
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

/* searchsorted(e, x, 'right') - 1 clamped to [0, n-1] */
static int om_locate(const double* e, int n, double x) {
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

/* Normalised cumulative group probabilities of _sample_group_piecewise_constant
 * (MG_IMC2D.py:307-338): sequential sums, last entry forced to 1. */
void om_group_cdf(int G, const double* tot, double* gcdf) {
  double s = 0.0;
  for (int g = 0; g < G; ++g) s += tot[g];
  double c = 0.0;
  for (int g = 0; g < G; ++g) {
    c += s > 0.0 ? tot[g] / s : 1.0 / (double)G;
    gcdf[g] = c;
  }
  gcdf[G - 1] = 1.0;
}

static void om_iso(int geometry, double ua, double ub, double* d1, double* d2) {
  if (geometry == 0) { /* _sample_isotropic_xy, MG_IMC2D.py:183-190 */
    double uz = 2.0 * ua - 1.0;
    double phi = 2.0 * OM_PI * ub;
    double r = sqrt(fmax(0.0, 1.0 - uz * uz));
    *d1 = r * cos(phi);
    *d2 = r * sin(phi);
  } else { /* _sample_isotropic_rz, 193-198 */
    *d2 = 2.0 * ua - 1.0;
    *d1 = cos(2.0 * OM_PI * ub);
  }
}

static void om_pos(int geometry, const double* e1, const double* e2, int i, int j, double u1, double u2, double* p1,
                   double* p2) {
  if (geometry == 0) {
    *p1 = e1[i] + u1 * (e1[i + 1] - e1[i]);
  } else {
    double rl = e1[i], rr = e1[i + 1];
    *p1 = sqrt(rl * rl + u1 * (rr * rr - rl * rl));
  }
  *p2 = e2[j] + u2 * (e2[j + 1] - e2[j]);
}

/* ------------------------------------------------------------------------- */
/* Material emission (MG_IMC2D.py:1896-1943): N particles of weight E/N;     */
/* cell from the cell-total CDF, group from the cell's group energies.        */
/* Slots: 0 cell, 1 group, 2 pos1, 3 pos2, 4/5 direction, 6 time.             */
/* ------------------------------------------------------------------------- */
/* ln x for x in (0, 1] from frexp and the atanh series to s^23, in plain
 * FP64 arithmetic (built without FMA contraction): the device draws its
 * exponential spacings with the same operations, so both get the same bits. */
static double om_ln01(double x) {
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

/* Material emission: N particles whose cells are the order statistics of N
 * uniforms, u_k = S_k / S_N with S the running sum (om_prefix3 order) of the
 * N + 1 spacings -ln(1 - u0_k), u0_k slot 0 of stream k; the group, position,
 * direction and time from slots 1..6 of stream k. */
void om_sample_emission(int geometry, int nx, int ny, int G, const double* e1, const double* e2, const double* E,
                        const double* cdf, int64_t N, double dt, uint64_t seed, uint32_t step, double* w, double* d1,
                        double* d2, double* t, double* p1, double* p2, int32_t* ci, int32_t* cj, int32_t* grp) {
  int64_t ncell = (int64_t)nx * ny;
  double Etot = cdf[ncell - 1];
  if (N <= 0) return;
  double* x = (double*)calloc((size_t)(N + 1), sizeof(double));
  double* S = (double*)calloc((size_t)(N + 1), sizeof(double));
  if (!x || !S) abort();
  for (int64_t k = 0; k <= N; ++k) {
    orc_stream s;
    orc_stream_init(&s, seed, ORC_TAG_EMISSION, (uint64_t)k, step);
    x[k] = -om_ln01(1.0 - orc_u01(&s, 0));
  }
  om_prefix3(x, N + 1, S);
  for (int64_t k = 0; k < N; ++k) {
    orc_stream s;
    orc_stream_init(&s, seed, ORC_TAG_EMISSION, (uint64_t)k, step);
    double u = S[k] / S[N];
    if (!(u < 1.0)) u = nextafter(1.0, 0.0);
    int64_t c = om_pick(cdf, ncell, u * Etot);
    int i = (int)(c / ny), j = (int)(c % ny);
    double tot = 0.0;
    for (int g = 0; g < G; ++g) tot += E[(int64_t)g * ncell + c];
    int32_t gs = 0;
    if (tot > 0.0) {
      double xi = orc_u01(&s, 1) * tot, cum = 0.0;
      for (int g = 0; g < G; ++g) {
        cum += E[(int64_t)g * ncell + c];
        if (xi <= cum) {
          gs = g;
          break;
        }
      }
    }
    om_pos(geometry, e1, e2, i, j, orc_u01(&s, 2), orc_u01(&s, 3), &p1[k], &p2[k]);
    om_iso(geometry, orc_u01(&s, 4), orc_u01(&s, 5), &d1[k], &d2[k]);
    t[k] = dt * orc_u01(&s, 6);
    w[k] = Etot / (double)N;
    ci[k] = i;
    cj[k] = j;
    grp[k] = gs;
  }
  free(x);
  free(S);
}

/* Equilibrium census (MG_IMC2D.py:992-1095): energy a Tr^4 V per cell (cdf),
 * N particles of weight E/N, t = 0, group from the Planck spectrum at the
 * cell's Tr.  Slots: 0 cell, 1 pos1, 2 pos2, 3/4 direction, 5..9 Planck. */
void om_sample_ic(int geometry, int nx, int ny, int G, const double* e1, const double* e2, const double* edges,
                  const double* Tr, const double* cdf, const double* pcdf, int64_t N, uint64_t seed, double* w,
                  double* d1, double* d2, double* t, double* p1, double* p2, int32_t* ci, int32_t* cj,
                  int32_t* grp) {
  int64_t ncell = (int64_t)nx * ny;
  double Etot = cdf[ncell - 1];
  for (int64_t k = 0; k < N; ++k) {
    orc_stream s;
    orc_stream_init(&s, seed, ORC_TAG_IC, (uint64_t)k, 0);
    int64_t c = om_pick(cdf, ncell, orc_u01(&s, 0) * Etot);
    int i = (int)(c / ny), j = (int)(c % ny);
    om_pos(geometry, e1, e2, i, j, orc_u01(&s, 1), orc_u01(&s, 2), &p1[k], &p2[k]);
    om_iso(geometry, orc_u01(&s, 3), orc_u01(&s, 4), &d1[k], &d2[k]);
    t[k] = 0.0;
    w[k] = Etot / (double)N;
    ci[k] = i;
    cj[k] = j;
    grp[k] = om_planck_group(&s, 5, Tr[c], G, edges, pcdf);
  }
}

/* Face source (MG_IMC2D.py:392-577).  side 0 left/rmin, 1 right/rmax,
 * 2 bottom/zmin, 3 top/zmax.  prof (nullable): running sum of per-face-cell
 * energy with face-cell temperatures Tface; otherwise uniform at Tb.
 * Slots: 0 time, 1 mu_n, 2 phi (xy) / eta or phi (rz), 3 along the face,
 * 4 face cell, 5..9 Planck. */
void om_sample_boundary(int geometry, int side, int64_t n, double E_total, double dt, double Tb, int nx, int ny,
                        const double* e1, const double* e2, int G, const double* edges, const double* pcdf,
                        const double* prof, const double* Tface, uint64_t seed, uint32_t step, double* w, double* d1,
                        double* d2, double* t, double* p1, double* p2, int32_t* ci, int32_t* cj, int32_t* grp) {
  int nface = side < 2 ? ny : nx;
  const double* fe = side < 2 ? e2 : e1;
  double x0 = e1[0], x1 = e1[nx], y0 = e2[0], y1 = e2[ny];
  for (int64_t k = 0; k < n; ++k) {
    orc_stream s;
    orc_stream_init(&s, seed, ORC_TAG_BOUNDARY0 + side, (uint64_t)k, step);
    double a = side < 2 ? y0 : x0, b = side < 2 ? y1 : x1, Tk = Tb;
    int m = -1;
    if (prof) {
      double u = orc_u01(&s, 4) * prof[nface - 1];
      m = 0;
      while (m < nface - 1 && !(u < prof[m])) ++m;
      while (m > 0 && prof[m] == prof[m - 1]) --m;
      a = fe[m];
      b = fe[m + 1];
      Tk = Tface[m];
    }
    t[k] = dt * orc_u01(&s, 0);
    double along = a + (b - a) * orc_u01(&s, 3);
    double mun = sqrt(orc_u01(&s, 1));
    double q1, q2, u1, u2;
    if (geometry == 0) {
      double mut = sqrt(fmax(0.0, 1.0 - mun * mun));
      double ct = mut * cos(2.0 * OM_PI * orc_u01(&s, 2));
      if (side == 0) { u1 = mun; u2 = ct; q1 = x0; q2 = along; }
      else if (side == 1) { u1 = -mun; u2 = ct; q1 = x1; q2 = along; }
      else if (side == 2) { u2 = mun; u1 = ct; q2 = y0; q1 = along; }
      else { u2 = -mun; u1 = ct; q2 = y1; q1 = along; }
    } else {
      if (side < 2) {
        u2 = 2.0 * orc_u01(&s, 2) - 1.0;
        u1 = side == 0 ? mun : -mun;
        q1 = side == 0 ? x0 : x1;
        q2 = along;
        if (prof) q1 = side == 0 ? x0 + 1e-12 : x1 - 1e-12;
      } else {
        u1 = cos(2.0 * OM_PI * orc_u01(&s, 2));
        u2 = side == 2 ? mun : -mun;
        q1 = sqrt(a * a + orc_u01(&s, 3) * (b * b - a * a));
        q2 = side == 2 ? y0 : y1;
        if (prof) q2 = side == 2 ? y0 + 1e-12 : y1 - 1e-12;
      }
    }
    /* cell: searchsorted(edges, pos, 'right') - 1, faces kept inside */
    int i = om_locate(e1, nx, q1), j = om_locate(e2, ny, q2);
    w[k] = E_total / (double)n;
    d1[k] = u1;
    d2[k] = u2;
    p1[k] = q1;
    p2[k] = q2;
    ci[k] = i;
    cj[k] = j;
    grp[k] = om_planck_group(&s, 5, Tk, G, edges, pcdf);
  }
}

/* Fixed source (MG_IMC2D.py:1098-1234): S (G, ncell) energy per cell this
 * step (source x V x dt); group from the group totals, cell from the
 * group's cell CDF.  cdfg: (G, ncell) canonical prefixes of S[g]; gcdf:
 * normalised cumulative group totals (last = 1).  Slots: 0 group, 1 cell,
 * 2 pos1, 3 pos2, 4/5 direction, 6 time. */
void om_sample_source(int geometry, int nx, int ny, int G, const double* e1, const double* e2, const double* cdfg,
                      const double* gcdf, double total, int64_t N, double dt, uint64_t seed, uint32_t step, double* w,
                      double* d1, double* d2, double* t, double* p1, double* p2, int32_t* ci, int32_t* cj,
                      int32_t* grp) {
  int64_t ncell = (int64_t)nx * ny;
  for (int64_t k = 0; k < N; ++k) {
    orc_stream s;
    orc_stream_init(&s, seed, ORC_TAG_SOURCE, (uint64_t)k, step);
    double xi = orc_u01(&s, 0);
    int g = G - 1;
    for (int q = 0; q < G; ++q)
      if (xi <= gcdf[q]) {
        g = q;
        break;
      }
    const double* cg = cdfg + (int64_t)g * ncell;
    int64_t c = 0;
    if (cg[ncell - 1] > 0.0) c = om_pick(cg, ncell, orc_u01(&s, 1) * cg[ncell - 1]);
    int i = (int)(c / ny), j = (int)(c % ny);
    om_pos(geometry, e1, e2, i, j, orc_u01(&s, 2), orc_u01(&s, 3), &p1[k], &p2[k]);
    om_iso(geometry, orc_u01(&s, 4), orc_u01(&s, 5), &d1[k], &d2[k]);
    t[k] = dt * orc_u01(&s, 6);
    w[k] = total / (double)N;
    ci[k] = i;
    cj[k] = j;
    grp[k] = g;
  }
}

/* ------------------------------------------------------------------------- */
/* Tracker (MG_IMC2D.py:580-989).  Event e uses slots 4e (distance to        */
/* scatter) and 4e+2, 4e+3 (new direction).                                   */
/* ------------------------------------------------------------------------- */
static double om_dist_radial(double r, double mu, double eta, double R) { /* 647-682 */
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

enum { OM_CENSUS = 0, OM_BOUNDARY = 1, OM_SCATTER = 2 };
enum { OM_NONE = 0, OM_IP = 1, OM_IM = 2, OM_JP = 3, OM_JM = 4 };

/* stats[8]: events, boundary crossings, scatters, census, captures (0),
 * floor kills, reflections, event-cap hits.  dep_raw, path_raw: (G, ncell)
 * sums of deposited energy and, where sigma_a <= 1e-12, of w*s.
 * bl[0] total, bl[1..G] by group, bl[G+1..G+4] by side. */
void om_transport(int geometry, int64_t n, double* w, double* d1, double* d2, double* times, double* p1, double* p2,
                  int32_t* ci, int32_t* cj, const int32_t* grp, int nx, int ny, int G, const double* e1,
                  const double* e2, const double* sa, const double* ss, double dt, const int32_t* refl,
                  int64_t max_events, double weight_floor, uint64_t seed, uint32_t step, double* dep_raw,
                  double* path_raw, double* bl, int64_t* stats) {
  int64_t ncell = (int64_t)nx * ny;
  for (int64_t k = 0; k < n; ++k) {
    double wt = w[k], u1 = d1[k], u2 = d2[k], t = times[k], x = p1[k], y = p2[k];
    int i = ci[k], j = cj[k], g = grp[k];
    if (wt <= weight_floor) continue;
    orc_stream st;
    orc_stream_init(&st, seed, ORC_TAG_TRANSPORT, (uint64_t)k, step);
    uint32_t ev = 0;
    for (int64_t e = 0; e < max_events; ++e) {
      if (i < 0 || i >= nx || j < 0 || j >= ny) {
        int side = i < 0 ? 0 : (i >= nx ? 1 : (j < 0 ? 2 : 3));
        bl[0] += wt;
        bl[1 + g] += wt;
        bl[1 + G + side] += wt;
        stats[1] += 1;
        wt = 0.0;
        break;
      }
      if (wt <= weight_floor) {
        stats[5] += 1;
        break;
      }
      double dcen = OM_C * (dt - t);
      int64_t q = (int64_t)g * ncell + (int64_t)i * ny + j;
      double sga = sa[q], sgs = ss[q];
      double s, xn, yn, un1 = u1;
      int evt = OM_CENSUS, cross = OM_NONE;
      double s_scat = 1e30;
      if (sgs > 1e-12) s_scat = -log(1.0 - orc_u01(&st, 4 * ev)) / sgs;
      if (geometry == 0) {
        double xl = e1[i], xr = e1[i + 1], yl = e2[j], yr = e2[j + 1];
        double sx = 1e30, sy = 1e30;
        if (u1 > 0.0) sx = (xr - x) / (u1 + 1e-300);
        else if (u1 < 0.0) sx = (xl - x) / (u1 - 1e-300);
        if (u2 > 0.0) sy = (yr - y) / (u2 + 1e-300);
        else if (u2 < 0.0) sy = (yl - y) / (u2 - 1e-300);
        s = sx;
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
        xn = x + u1 * s;
        yn = y + u2 * s;
        if (s == dcen) evt = OM_CENSUS;
        else if (s == s_scat) evt = OM_SCATTER;
        else if (s == sx) { cross = u1 > 0.0 ? OM_IP : OM_IM; evt = OM_BOUNDARY; }
        else if (s == sy) { cross = u2 > 0.0 ? OM_JP : OM_JM; evt = OM_BOUNDARY; }
      } else {
        double rl = e1[i], rr = e1[i + 1], zl = e2[j], zr = e2[j + 1];
        double srl = rl > 0.0 ? om_dist_radial(x, u1, u2, rl) : 1e30;
        double srr = om_dist_radial(x, u1, u2, rr);
        double sr = srl < srr ? srl : srr;
        double sz = 1e30;
        if (fabs(u2) > 1e-12) sz = u2 > 0.0 ? (zr - y) / (u2 + 1e-300) : (zl - y) / (u2 - 1e-300);
        s = sr;
        if (sz < s) s = sz;
        if (s_scat < s) s = s_scat;
        if (dcen < s) s = dcen;
        if (s < 1e-14) s = 1e-14;
        double ome2 = fmax(0.0, 1.0 - u2 * u2);
        double l = s * sqrt(ome2);
        double rn2 = x * x + 2.0 * x * u1 * l + l * l;
        if (rn2 < 0.0) rn2 = 0.0;
        xn = sqrt(rn2);
        yn = y + u2 * s;
        if (xn > 1e-15 && l > 0.0) {
          un1 = (x * u1 + l) / xn;
          if (un1 > 1.0) un1 = 1.0;
          else if (un1 < -1.0) un1 = -1.0;
        }
        if (s == dcen) evt = OM_CENSUS;
        else if (s == s_scat) evt = OM_SCATTER;
        else if (s == sr) { cross = fabs(srl - sr) < 1e-14 ? OM_IM : OM_IP; evt = OM_BOUNDARY; }
        else if (s == sz) { cross = u2 > 0.0 ? OM_JP : OM_JM; evt = OM_BOUNDARY; }
      }
      double wn = wt, dep = 0.0;
      if (sga > 1e-12) {
        wn = wt * exp(-sga * s);
        dep = wt - wn;
      }
      dep_raw[q] += dep;
      if (!(sga > 1e-12)) path_raw[q] += wt * s;
      x = xn;
      y = yn;
      u1 = un1;
      wt = wn;
      t = t + s / OM_C;
      stats[0] += 1;
      uint32_t evs = ev;
      ev++;
      if (evt == OM_CENSUS) {
        stats[3] += 1;
        break;
      } else if (evt == OM_SCATTER) {
        om_iso(geometry, orc_u01(&st, 4 * evs + 2), orc_u01(&st, 4 * evs + 3), &u1, &u2);
        stats[2] += 1;
      } else {
        stats[1] += 1;
        if (cross == OM_IP) ++i;
        else if (cross == OM_IM) --i;
        else if (cross == OM_JP) ++j;
        else if (cross == OM_JM) --j;
        int side = -1;
        if (i < 0) side = 0;
        else if (i >= nx) side = 1;
        else if (j < 0) side = 2;
        else if (j >= ny) side = 3;
        if (side >= 0) {
          if (refl[side]) {
            if (side < 2) { u1 = -u1; i = side == 0 ? 0 : nx - 1; }
            else { u2 = -u2; j = side == 2 ? 0 : ny - 1; }
            stats[6] += 1;
          } else {
            bl[0] += wt;
            bl[1 + g] += wt;
            bl[1 + G + side] += wt;
            wt = 0.0;
            break;
          }
        }
      }
      if (e == max_events - 1) stats[7] += 1;
    }
    w[k] = wt;
    d1[k] = u1;
    d2[k] = u2;
    times[k] = t;
    p1[k] = x;
    p2[k] = y;
    ci[k] = i;
    cj[k] = j;
  }
}

/* (cell, group) comb (MG_IMC2D.py:1270-1378): bins b = g*ncell + c.  Returns
 * the number of copies; passthrough (no census energy): survivors keep
 * their weights. */
int64_t om_comb_counts(int64_t n, const double* w, const int32_t* ci, const int32_t* cj, const int32_t* grp, int nx,
                       int ny, int G, int64_t Nmax, uint64_t seed, uint32_t step, double* ecen, double* ew,
                       int64_t* copies, int32_t* passthrough) {
  int64_t ncell = (int64_t)nx * ny, nb = ncell * G;
  extern double orc_canonical_sum(const double* x, int64_t n);
  for (int64_t b = 0; b < nb; ++b) ecen[b] = 0.0;
  for (int64_t k = 0; k < n; ++k)
    if (w[k] > 0.0 && ci[k] >= 0 && ci[k] < nx && cj[k] >= 0 && cj[k] < ny)
      ecen[(int64_t)grp[k] * ncell + (int64_t)ci[k] * ny + cj[k]] += w[k];
  double E = orc_canonical_sum(ecen, nb);
  *passthrough = 0;
  int64_t tot = 0;
  if (!(E > 0.0)) {
    *passthrough = 1;
    for (int64_t b = 0; b < nb; ++b) ew[b] = 0.0;
    for (int64_t k = 0; k < n; ++k) {
      copies[k] = w[k] > 0.0 ? 1 : 0;
      tot += copies[k];
    }
    return tot;
  }
  for (int64_t b = 0; b < nb; ++b) {
    if (ecen[b] > 0.0) {
      double d = nearbyint((double)Nmax * ecen[b] / E);
      if (d < 1.0) d = 1.0;
      ew[b] = ecen[b] / d;
    } else {
      ew[b] = 0.0;
    }
  }
  for (int64_t k = 0; k < n; ++k) {
    int64_t m = 0;
    if (w[k] > 0.0 && ci[k] >= 0 && ci[k] < nx && cj[k] >= 0 && cj[k] < ny) {
      double et = ew[(int64_t)grp[k] * ncell + (int64_t)ci[k] * ny + cj[k]];
      if (et > 0.0) {
        orc_stream s;
        orc_stream_init(&s, seed, ORC_TAG_COMB, (uint64_t)k, step);
        m = (int64_t)(w[k] / et + orc_u01(&s, 0));
      }
    }
    copies[k] = m;
    tot += m;
  }
  return tot;
}
