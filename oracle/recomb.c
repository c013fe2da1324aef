/*
 * recomb.c - the oracle's restatement of src/core/recombinations.c (after
 * 21cmFAST): the MHR00 density-PDF parameters A, C, beta (:159-300), the
 * self-shielded photoionisation rate (:103-108), the equilibrium neutral
 * fraction (:321-352), the recombination rate per baryon integrated over the
 * PDF in ln D (:110-156) and its table in (z, ln Gamma12) (init_MHR :56-82),
 * read back by splined_recombination_rate (:30-54).  Also dtdz
 * (XRayHeatingFunctions.c:377-392), which the N_rec update uses.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  It shares no code with the
 * library.  GSL's routines are restated from their published algorithms: the
 * natural cubic spline (gsl_interp_cspline; evaluation outside the knots is a
 * GSL domain error, returned here as NaN) and QUADPACK's QAG with the 61-point
 * Gauss-Kronrod rule, whose nodes and weights the caller installs first
 * (orc_set_gk61; oracle/ts_host.py computes them from their definition).
 * Build with -ffp-contract=off.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define R_HUBBLE 3.2407789e-18
#define R_GRAVITY 6.672e-8
#define R_PROTONMASS 1.6726e-24
#define R_TINY 1e-30
#define R_ALPHAB_10K 2.59e-13
#define R_FRACT_FLOAT_ERR 1e-7
#define R_A_NPTS 60
#define R_C_NPTS 12
#define R_BETA_NPTS 5
#define R_Z_NPTS 300
#define R_DEL_Z ((float)0.2)
#define R_LNG_NPTS 150
#define R_LNG_MIN (-10.0)
#define R_DEL_LNG ((float)0.1)

/* ------------------------------------------------------------ GK61 + QAG */

static double g_xgk[31], g_wgk[31], g_wg[15];
static int g_rule_set = 0;

void orc_set_gk61(const double* xgk, const double* wgk, const double* wg)
{
  memcpy(g_xgk, xgk, sizeof g_xgk);
  memcpy(g_wgk, wgk, sizeof g_wgk);
  memcpy(g_wg, wg, sizeof g_wg);
  g_rule_set = 1;
}

static double r_rescale(double err, double res_abs, double res_asc)
{
  err = fabs(err);
  if (res_asc != 0 && err != 0) {
    const double sc = pow(200 * err / res_asc, 1.5);
    err = sc < 1 ? res_asc * sc : res_asc;
  }
  if (res_abs > DBL_MIN / (50 * DBL_EPSILON)) {
    const double m = 50 * DBL_EPSILON * res_abs;
    if (m > err) err = m;
  }
  return err;
}

static double r_gk61(double (*f)(double, void*), void* par, double a, double b, double* abserr, double* resabs,
                     double* resasc)
{
  const double c = 0.5 * (a + b), h = 0.5 * (b - a);
  double f1[30], f2[30];
  const double fc = f(c, par);
  for (int j = 0; j < 30; j++) {
    f1[j] = f(c - h * g_xgk[j], par);
    f2[j] = f(c + h * g_xgk[j], par);
  }
  double rg = 0, rk = fc * g_wgk[30], ra = fabs(rk);
  for (int j = 0; j < 15; j++) {
    const int t = 2 * j + 1;
    const double s = f1[t] + f2[t];
    rg += g_wg[j] * s;
    rk += g_wgk[t] * s;
    ra += g_wgk[t] * (fabs(f1[t]) + fabs(f2[t]));
  }
  for (int j = 0; j < 15; j++) {
    const int t = 2 * j;
    rk += g_wgk[t] * (f1[t] + f2[t]);
    ra += g_wgk[t] * (fabs(f1[t]) + fabs(f2[t]));
  }
  const double mean = rk * 0.5;
  double rs = g_wgk[30] * fabs(fc - mean);
  for (int j = 0; j < 30; j++) rs += g_wgk[j] * (fabs(f1[j] - mean) + fabs(f2[j] - mean));
  const double err = (rk - rg) * h;
  rk *= h;
  ra *= fabs(h);
  rs *= fabs(h);
  *resabs = ra;
  *resasc = rs;
  *abserr = r_rescale(err, ra, rs);
  return rk;
}

/* QAG, absolute tolerance 0: bisect the interval of largest error estimate
 * (its lower half keeps the slot unless the upper half has the larger error)
 * until the summed error is within epsrel |result|, roundoff is detected, an
 * interval gets too small or `limit` intervals exist; the result is the sum
 * of the intervals in storage order. */
double orc_qag61(double (*f)(double, void*), void* par, double a, double b, double epsrel, int limit,
                 double* abserr)
{
  enum { MAXI = 1000 };
  double al[MAXI], bl[MAXI], rl[MAXI], el[MAXI];
  double resabs, resasc, e0;
  if (!g_rule_set) return NAN;
  if (limit > MAXI) limit = MAXI;
  const double r0 = r_gk61(f, par, a, b, &e0, &resabs, &resasc);
  double tol = epsrel * fabs(r0);
  if ((e0 <= 50 * DBL_EPSILON * resabs && e0 > tol) || (e0 <= tol && e0 != resasc) || e0 == 0.0 || limit == 1) {
    if (abserr) *abserr = e0;
    return r0;
  }
  int n = 1;
  al[0] = a;
  bl[0] = b;
  rl[0] = r0;
  el[0] = e0;
  double area = r0, errsum = e0;
  int rt1 = 0, rt2 = 0, stop = 0;
  for (int it = 1; it < limit && !stop; it++) {
    int k = 0;
    for (int i = 1; i < n; i++)
      if (el[i] > el[k]) k = i;
    const double a1 = al[k], b1 = 0.5 * (al[k] + bl[k]), a2 = b1, b2 = bl[k];
    double e1, e2, s1, s2, x1, x2;
    const double r1 = r_gk61(f, par, a1, b1, &e1, &x1, &s1);
    const double r2 = r_gk61(f, par, a2, b2, &e2, &x2, &s2);
    const double area12 = r1 + r2, error12 = e1 + e2;
    errsum += error12 - el[k];
    area += area12 - rl[k];
    if (s1 != e1 && s2 != e2) {
      if (fabs(rl[k] - area12) <= 1.0e-5 * fabs(area12) && error12 >= 0.99 * el[k]) rt1++;
      if (it >= 10 && error12 > el[k]) rt2++;
    }
    tol = epsrel * fabs(area);
    if (errsum > tol) {
      if (rt1 >= 6 || rt2 >= 20) stop = 1;
      const double lim = (1 + 100 * DBL_EPSILON) * (fabs(a2) + 1000 * DBL_MIN);
      if (fabs(a1) <= lim && fabs(b2) <= lim) stop = 1;
    }
    if (e2 > e1) {
      al[k] = a2;
      rl[k] = r2;
      el[k] = e2;
      al[n] = a1;
      bl[n] = b1;
      rl[n] = r1;
      el[n] = e1;
    } else {
      bl[k] = b1;
      rl[k] = r1;
      el[k] = e1;
      al[n] = a2;
      bl[n] = b2;
      rl[n] = r2;
      el[n] = e2;
    }
    n++;
    if (errsum <= tol) break;
  }
  double res = 0;
  for (int i = 0; i < n; i++) res += rl[i];
  if (abserr) *abserr = errsum;
  return res;
}

/* --------------------------------------------------- natural cubic spline */

typedef struct {
  int n;
  double x[R_LNG_NPTS], y[R_LNG_NPTS], c[R_LNG_NPTS];
} r_spline;

static void r_spline_init(r_spline* s, const double* x, const double* y, int n)
{
  s->n = n;
  memcpy(s->x, x, sizeof(double) * n);
  memcpy(s->y, y, sizeof(double) * n);
  s->c[0] = s->c[n - 1] = 0.0;
  /* tridiagonal system of the interior second-derivative coefficients */
  double cp[R_LNG_NPTS], dp[R_LNG_NPTS];
  const int m = n - 2;
  for (int i = 0; i < m; i++) {
    const double h0 = x[i + 1] - x[i], h1 = x[i + 2] - x[i + 1];
    const double diag = 2.0 * (h0 + h1);
    const double rhs = 3.0 * ((y[i + 2] - y[i + 1]) / h1 - (y[i + 1] - y[i]) / h0);
    const double sub = i > 0 ? h0 : 0.0;
    const double den = diag - (i > 0 ? sub * cp[i - 1] : 0.0);
    cp[i] = (i < m - 1) ? h1 / den : 0.0;
    dp[i] = (rhs - (i > 0 ? sub * dp[i - 1] : 0.0)) / den;
  }
  for (int i = m - 1; i >= 0; i--) s->c[i + 1] = dp[i] - (i < m - 1 ? cp[i] * s->c[i + 2] : 0.0);
}

static double r_spline_eval(const r_spline* s, double xv)
{
  if (xv < s->x[0] || xv > s->x[s->n - 1]) return NAN; /* GSL_EDOM */
  int lo = 0, hi = s->n - 1;
  while (hi > lo + 1) {
    const int i = (hi + lo) / 2;
    if (s->x[i] > xv)
      hi = i;
    else
      lo = i;
  }
  const double h = s->x[lo + 1] - s->x[lo], dx = xv - s->x[lo], dy = s->y[lo + 1] - s->y[lo];
  const double b = dy / h - h * (s->c[lo + 1] + 2.0 * s->c[lo]) / 3.0;
  const double d = (s->c[lo + 1] - s->c[lo]) / (3.0 * h);
  return s->y[lo] + dx * (b + dx * (s->c[lo] + dx * d));
}

/* ------------------------------------------------------------- physics */

double orc_dtdz(double omega_l, double omega_m, double hubble_h, float z)
{
  const double x = sqrt(omega_l / omega_m) * pow(1 + z, -3.0 / 2.0); /* 1 + z: float arithmetic */
  const double dxdz = sqrt(omega_l / omega_m) * pow(1 + z, -5.0 / 2.0) * (-3.0 / 2.0);
  const double const1 = 2 * sqrt(1 + omega_m / omega_l) / (3.0 * R_HUBBLE * hubble_h);
  const double numer = dxdz * (1 + x * pow(pow(x, 2) + 1, -0.5));
  const double denom = x + sqrt(pow(x, 2) + 1);
  return const1 * numer / denom;
}

static double r_alpha_B(double T) { return R_ALPHAB_10K * pow(T / 1.0e4, -0.75); }

double orc_rr_Gamma_SS(double Gamma_bg, double Delta, double T_4, double z)
{
  const double D_ss = 26.7 * pow(T_4, 0.17) * pow((1 + z) / 10.0, -3) * pow(Gamma_bg, 2.0 / 3.0);
  return Gamma_bg * (0.98 * pow((1.0 + pow(Delta / D_ss, 1.64)), -2.28) + 0.02 * pow(1.0 + Delta / D_ss, -0.84));
}

/* case B only: every call site in the reference passes usecaseB = 1 */
double orc_rr_neutral_fraction(double y_he, double density, double T4, double gamma12)
{
  const double corr_He = 1.0 / (4.0 / y_he - 3);
  const double alpha = r_alpha_B(T4 * 1e4);
  gamma12 *= 1e-12;
  double chi = (1 + corr_He) * density * alpha / gamma12;
  if (chi < R_TINY) return 0;
  if (chi < 1e-5) return chi;
  const double b = -2 - gamma12 / (density * (1 + corr_He) * alpha);
  chi = (-b - sqrt(b * b - 4)) / 2.0;
  return chi;
}

typedef struct {
  double No, y_he;
  r_spline A, C, beta;
  double lnG[R_LNG_NPTS];
  r_spline row[R_Z_NPTS];
  int built;
} r_tables;

static r_tables* g_rr = NULL;

double orc_rr_C(double z)
{
  if (z >= 13.0) return 1.0;
  if (z <= 2.0) return 0.558;
  return r_spline_eval(&g_rr->C, z);
}

double orc_rr_beta(double z)
{
  if (z >= 6.0) return -2.50;
  if (z <= 2.0) return -2.23;
  return r_spline_eval(&g_rr->beta, z);
}

double orc_rr_A(double z)
{
  if (z >= 2.0 + (float)R_A_NPTS) return r_spline_eval(&g_rr->A, 2.0 + (float)R_A_NPTS);
  if (z <= 2.0) return r_spline_eval(&g_rr->A, 2.0);
  return r_spline_eval(&g_rr->A, z);
}

static double r_aux(double D, void* vp)
{
  const double z = *(const double*)vp;
  const double s = 2.0 * 7.61 / (3.0 * (1.0 + z));
  return exp(-(pow(D, -2.0 / 3.0) - orc_rr_C(z)) * (pow(D, -2.0 / 3.0) - orc_rr_C(z)) / (2.0 * s * s)) *
         pow(D, orc_rr_beta(z));
}

double orc_rr_A_aux_integral(double z) { return orc_qag61(r_aux, &z, 1e-25, 1e25, 0.001, 1000, NULL); }

typedef struct {
  double z, gamma12_bg, T4, A, C_0, beta, avenH, y_he;
} r_par;

static double r_mhr(double lnD, void* vp)
{
  const r_par* p = (const r_par*)vp;
  const double D = exp(lnD);
  const double gamma = orc_rr_Gamma_SS(p->gamma12_bg, D, p->T4, p->z);
  const double n_H = p->avenH * D;
  const double x_e = 1.0 - orc_rr_neutral_fraction(p->y_he, n_H, p->T4, gamma);
  const double PDelta =
    p->A * exp(-0.5 * pow((pow(D, -2.0 / 3.0) - p->C_0) / ((2.0 * 7.61 / (3.0 * (1.0 + p->z)))), 2)) * pow(D, p->beta);
  const double alpha = r_alpha_B(p->T4 * 1e4);
  return n_H * PDelta * alpha * x_e * x_e * D * D;
}

double orc_rr_recombination_rate(double z_eff, double gamma12_bg, double T4)
{
  r_par p = {z_eff, gamma12_bg, T4, orc_rr_A(z_eff), orc_rr_C(z_eff), orc_rr_beta(z_eff),
             g_rr->No * pow(1 + z_eff, 3), g_rr->y_he};
  return orc_qag61(r_mhr, &p, log(0.01), log(200), 0.01, 1000, NULL);
}

/* init_MHR for the cosmology's No and Y_He (reionization.h:17-27); 0 on success */
int orc_rr_build(double hubble_h, double omega_m, double baryon_frac, double y_he)
{
  if (!g_rule_set) return MRX_EINVAL;
  const double rhoc = 3.0 * R_HUBBLE * R_HUBBLE * hubble_h * hubble_h / (8.0 * M_PI * R_GRAVITY);
  const double No = rhoc * baryon_frac * omega_m * (1 - y_he) / R_PROTONMASS;
  if (!g_rr) g_rr = (r_tables*)calloc(1, sizeof(r_tables));
  if (!g_rr) return MRX_ENOMEM;
  if (g_rr->built && g_rr->No == No && g_rr->y_he == y_he) return MRX_OK;
  g_rr->No = No;
  g_rr->y_he = y_he;
  double x[R_A_NPTS], y[R_A_NPTS];
  const double Ct[R_C_NPTS] = {0.558, 0.599, 0.611, 0.769, 0.868, 0.930, 0.964, 0.983, 0.993, 0.998, 0.999, 1.00};
  const double Bt[R_BETA_NPTS] = {-2.23, -2.35, -2.48, -2.49, -2.50};
  for (int i = 0; i < R_C_NPTS; i++) x[i] = (float)i + 2.0;
  r_spline_init(&g_rr->C, x, Ct, R_C_NPTS);
  r_spline_init(&g_rr->beta, x, Bt, R_BETA_NPTS);
  for (int i = 0; i < R_A_NPTS; i++) {
    x[i] = 2.0 + (float)i;
    y[i] = 1.0 / orc_rr_A_aux_integral(2.0 + (float)i);
  }
  r_spline_init(&g_rr->A, x, y, R_A_NPTS);
  for (int g = 0; g < R_LNG_NPTS; g++) g_rr->lnG[g] = R_LNG_MIN + (float)(g * R_DEL_LNG);
#pragma omp parallel for schedule(dynamic)
  for (int zc = 0; zc < R_Z_NPTS; zc++) {
    const float z = zc * R_DEL_Z;
    double row[R_LNG_NPTS];
    for (int g = 0; g < R_LNG_NPTS; g++) {
      const float gamma = (float)exp(g_rr->lnG[g]);
      row[g] = orc_rr_recombination_rate(z, gamma, 1);
    }
    r_spline_init(&g_rr->row[zc], g_rr->lnG, row, R_LNG_NPTS);
  }
  g_rr->built = 1;
  return MRX_OK;
}

/* splined_recombination_rate (recombinations.c:30-54) */
double orc_rr_rate(double z_eff, double gamma12_bg)
{
  int z_ct = (int)(z_eff / R_DEL_Z + 0.5);
  double lnGamma = log(gamma12_bg);
  if (z_ct < 0)
    z_ct = 0;
  else if (z_ct >= R_Z_NPTS)
    z_ct = R_Z_NPTS - 1;
  if (lnGamma < R_LNG_MIN) return 0;
  if (lnGamma >= (R_LNG_MIN + R_DEL_LNG * R_LNG_NPTS))
    lnGamma = R_LNG_MIN + R_DEL_LNG * R_LNG_NPTS - R_FRACT_FLOAT_ERR;
  return r_spline_eval(&g_rr->row[z_ct], lnGamma);
}

/* a knot value of the table, for the tests */
double orc_rr_table(int z_ct, int g) { return g_rr->row[z_ct].y[g]; }
double orc_rr_lnG(int g) { return g_rr->lnG[g]; }
