/*
 * ts_tables.c - host-side scalar physics of the spin-temperature step.
 *
 * Restates, in double precision and in the operation order that fixes the
 * result, the per-snapshot (not per-cell) parts of
 *   src/core/ComputeTs.c:32-76 (set_sfr_history), :270-376 (shell ladder and
 *     SFR-history bookkeeping), :505-716 (frequency-integral tables, sum_lyn,
 *     luminosity conversion and prefactor);
 *   src/core/XRayHeatingFunctions.c: T_RECFAST/xion_RECFAST :105-223,
 *     nu_tau_one :238-298, tauX :301-375, dtdz/hubble/drdz :377-473, zmax,
 *     nu_n, frecycle :496-583, spectral_emissivity :586-671, the three
 *     frequency integrands and integrate_over_nu :673-800, the x_int tables
 *     and interpolation :805-1220, dicke/ddicke_dz :1469-1511, kappa_10* :1602-1840.
 * GSL is replaced by this file's own natural cubic spline, adaptive
 * Gauss-Kronrod 61-point integrator (QAG) and Brent root finder, with the
 * reference's tolerances.  Compile with -ffp-contract=off.
 */
#include "ts_tables.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "gk61_rule.h"

/* src/meraxes.h:21-44 */
#define K_GRAVITY 6.672e-8
#define K_SOLAR_MASS 1.989e33
#define K_BOLTZMANN 1.3806e-16
#define K_SPEED_OF_LIGHT 2.9979e10
#define K_PLANCK 6.6262e-27
#define K_PROTONMASS 1.6726e-24
#define K_HUBBLE 3.2407789e-18
#define K_SEC_PER_YEAR 3.155e7
#define K_MPC 3.086e24
#define K_TCMB 2.728
/* XRayHeatingFunctions.h:27-40, ComputeTs.h:6, reionization.h:38 */
#define K_Ly_alpha_HZ 2.46606727e15
#define K_NU_over_EV (1.60217646e-12 / K_PLANCK)
#define K_NUIONIZATION (13.60 * K_NU_over_EV)
#define K_HeI_NUIONIZATION (24.59 * K_NU_over_EV)
#define K_HeII_NUIONIZATION (K_NUIONIZATION * 4)
#define K_SIGMA_HI 6.3e-18
#define K_TINY 1e-30
#define K_FRACT_FLOAT_ERR 1e-7
#define K_R_XLy_MAX ((float)(500))
#define K_L_FACTOR 0.620350491

/* ------------------------------------------------------------ spline */

int mrx_spline_init(mrx_spline* s, const double* x, const double* y, int n)
{
  if (n < 3 || n > MRX_RECFAST_NPTS) return -1;
  s->n = n;
  for (int i = 0; i < n; i++) {
    s->x[i] = x[i];
    s->y[i] = y[i];
  }
  /* natural end conditions: second derivative zero at both ends; solve the
   * tridiagonal system for c_i = y''_i / 2 by forward elimination */
  double g[MRX_RECFAST_NPTS], diag[MRX_RECFAST_NPTS], rhs[MRX_RECFAST_NPTS];
  const int m = n - 2;
  for (int i = 0; i < m; i++) {
    const double h0 = x[i + 1] - x[i], h1 = x[i + 2] - x[i + 1];
    diag[i] = 2.0 * (h0 + h1);
    g[i] = h1;
    rhs[i] = 3.0 * ((y[i + 2] - y[i + 1]) / h1 - (y[i + 1] - y[i]) / h0);
  }
  for (int i = 1; i < m; i++) {
    const double w = g[i - 1] / diag[i - 1];
    diag[i] -= w * g[i - 1];
    rhs[i] -= w * rhs[i - 1];
  }
  s->c[0] = 0.0;
  s->c[n - 1] = 0.0;
  if (m > 0) {
    s->c[m] = rhs[m - 1] / diag[m - 1];
    for (int i = m - 2; i >= 0; i--) s->c[i + 1] = (rhs[i] - g[i] * s->c[i + 2]) / diag[i];
  }
  for (int i = 0; i < n - 1; i++) {
    const double h = x[i + 1] - x[i];
    s->b[i] = (y[i + 1] - y[i]) / h - h * (s->c[i + 1] + 2.0 * s->c[i]) / 3.0;
    s->d[i] = (s->c[i + 1] - s->c[i]) / (3.0 * h);
  }
  s->b[n - 1] = s->d[n - 1] = 0.0;
  return 0;
}

double mrx_spline_eval(const mrx_spline* s, double xv)
{
  int lo = 0, hi = s->n - 1;
  if (xv <= s->x[0]) {
    lo = 0;
  } else if (xv >= s->x[hi]) {
    lo = hi - 1;
  } else {
    while (hi - lo > 1) {
      int mid = (lo + hi) / 2;
      if (s->x[mid] > xv)
        hi = mid;
      else
        lo = mid;
    }
  }
  const double dx = xv - s->x[lo];
  return s->y[lo] + dx * (s->b[lo] + dx * (s->c[lo] + dx * s->d[lo]));
}

/* ------------------------------------------------------------ table files */

static FILE* open_table(const char* dir, const char* name)
{
  char path[1024];
  snprintf(path, sizeof path, "%s/%s", dir, name);
  return fopen(path, "r");
}

static int read_recfast(mrx_ts_data* d, const char* dir)
{
  FILE* F = open_table(dir, "recfast_LCDM.dat");
  if (!F) return -1;
  double zt[MRX_RECFAST_NPTS], TK[MRX_RECFAST_NPTS], xe[MRX_RECFAST_NPTS];
  /* the file runs z = 500 .. 0; stored ascending (XRayHeatingFunctions.c:127-131) */
  for (int i = MRX_RECFAST_NPTS - 1; i >= 0; i--) {
    float currz, currxion, trash, currTK;
    if (fscanf(F, "%f %E %E %E", &currz, &currxion, &trash, &currTK) != 4) {
      fclose(F);
      return -1;
    }
    zt[i] = currz;
    xe[i] = currxion;
    TK[i] = currTK;
  }
  fclose(F);
  d->recfast_zmax = zt[MRX_RECFAST_NPTS - 1];
  if (mrx_spline_init(&d->recfast_T, zt, TK, MRX_RECFAST_NPTS)) return -1;
  if (mrx_spline_init(&d->recfast_xe, zt, xe, MRX_RECFAST_NPTS)) return -1;
  return 0;
}

static int read_kappa_file(mrx_spline* s, const char* dir, const char* name, int n)
{
  FILE* F = open_table(dir, name);
  if (!F) return -1;
  double t[32], k[32];
  for (int i = 0; i < n; i++) {
    float curr_TK, curr_kappa;
    if (fscanf(F, "%f %e", &curr_TK, &curr_kappa) != 2) {
      fclose(F);
      return -1;
    }
    t[i] = log((double)curr_TK);
    k[i] = log((double)curr_kappa);
  }
  fclose(F);
  return mrx_spline_init(s, t, k, n);
}

static int init_kappa_HH(mrx_ts_data* d)
{
  /* spin de-excitation rate for H-H collisions, Zygelman (2005) table 2
   * column 4, as tabulated at XRayHeatingFunctions.c:1612-1665 */
  static const double tk[27] = {1.0,  2.0,  4.0,  6.0,  8.0,   10.0,  15.0,  20.0,   25.0,
                                30.0, 40.0, 50.0, 60.0, 70.0,  80.0,  90.0,  100.0,  200.0,
                                300.0, 501.0, 701.0, 1000.0, 2000.0, 3000.0, 5000.0, 7000.0, 10000.0};
  static const double ka[27] = {1.38e-13, 1.43e-13, 2.71e-13, 6.60e-13, 1.47e-12, 2.88e-12, 9.10e-12,
                                1.78e-11, 2.73e-11, 3.67e-11, 5.38e-11, 6.86e-11, 8.14e-11, 9.25e-11,
                                1.02e-10, 1.11e-10, 1.19e-10, 1.75e-10, 2.09e-10, 2.565e-10, 2.91e-10,
                                3.31e-10, 4.27e-10, 4.97e-10, 6.03e-10, 6.87e-10, 7.87e-10};
  double lt[27], lk[27];
  for (int i = 0; i < 27; i++) {
    lt[i] = log(tk[i]);
    lk[i] = log(ka[i]);
  }
  return mrx_spline_init(&d->kappa_HH, lt, lk, 27);
}

static int read_spectra(mrx_ts_data* d, const char* dir, double nion)
{
  FILE* F = open_table(dir, "stellar_spectra.dat");
  if (!F) return -1;
  memset(d->spec_n, 0, sizeof d->spec_n);
  for (int i = 1; i < MRX_NSPEC_MAX; i++) {
    if (fscanf(F, "%i %e %e %e %e", &d->spec_n[i], &d->spec_N0_2[i], &d->spec_alpha_2[i], &d->spec_N0_3[i],
               &d->spec_alpha_3[i]) != 5) {
      fclose(F);
      return -1;
    }
  }
  fclose(F);
  for (int i = 1; i < MRX_NSPEC_MAX; i++) d->spec_nu[i] = (float)(4.0 / 3.0 * (1.0 - 1.0 / pow(d->spec_n[i], 2.0)));
  /* normalisation (XRayHeatingFunctions.c:633-638); the Pop III slot is
   * normalised with the same ReionNionPhotPerBary, minihalos being off */
  for (int i = 1; i < (MRX_NSPEC_MAX - 1); i++) {
    double n0_fac = (pow(d->spec_nu[i + 1], d->spec_alpha_2[i] + 1) - pow(d->spec_nu[i], d->spec_alpha_2[i] + 1));
    d->spec_N0_2[i] *= (d->spec_alpha_2[i] + 1) / n0_fac * nion;
    n0_fac = (pow(d->spec_nu[i + 1], d->spec_alpha_3[i] + 1) - pow(d->spec_nu[i], d->spec_alpha_3[i] + 1));
    d->spec_N0_3[i] *= (d->spec_alpha_3[i] + 1) / n0_fac * nion;
  }
  return 0;
}

static int read_xint(mrx_ts_data* d, const char* dir)
{
  static const float xhii[MRX_X_INT_NXHII] = {1.0e-4f, 2.318e-4f, 4.677e-4f, 1.0e-3f, 2.318e-3f,
                                              4.677e-3f, 1.0e-2f, 2.318e-2f, 4.677e-2f, 1.0e-1f,
                                              0.5f,     0.9f,     0.99f,    0.999f};
  for (int n = 0; n < MRX_X_INT_NXHII; n++) {
    d->xi_XHII[n] = xhii[n];
    char name[256];
    if (xhii[n] < 0.3)
      snprintf(name, sizeof name, "x_int_tables/log_xi_%1.1f.dat", log10(xhii[n]));
    else
      snprintf(name, sizeof name, "x_int_tables/xi_%1.3f.dat", xhii[n]);
    FILE* F = open_table(dir, name);
    if (!F) return -1;
    char label[128];
    float a, b, c, zz, T, trash;
    for (int i = 0; i < 5; i++)
      if (fscanf(F, "%127s", label) != 1) goto bad;
    if (fscanf(F, "%g %g %g %g %g", &a, &b, &c, &zz, &T) != 5) goto bad;
    for (int i = 0; i < 11; i++)
      if (fscanf(F, "%127s", label) != 1) goto bad;
    for (int i = 0; i < MRX_X_INT_NENERGY; i++) {
      if (fscanf(F, "%g %g %g %g %g %g %g %g %g", &d->xi_E[i], &trash, &d->xi_fheat[n][i], &trash,
                 &d->xi_nLya[n][i], &d->xi_nHI[n][i], &d->xi_nHeI[n][i], &d->xi_nHeII[n][i], &trash) != 9)
        goto bad;
    }
    fclose(F);
    continue;
  bad:
    fclose(F);
    return -1;
  }
  return 0;
}

int mrx_ts_data_init(mrx_ts_data* d, const mrx_ts_params* p)
{
  memset(d, 0, sizeof *d);
  const char* dir = p->tables_dir;
  if (!dir) return MRX_EINVAL;
  if (read_recfast(d, dir) || read_kappa_file(&d->kappa_eH, dir, "kappa_eH_table.dat", 20) ||
      read_kappa_file(&d->kappa_pH, dir, "kappa_pH_table.dat", 17) || init_kappa_HH(d) ||
      read_spectra(d, dir, p->nion_phot_per_bary) || read_xint(d, dir))
    return MRX_EINVAL;
  d->hubble_h = p->hubble_h;
  d->omega_m = p->omega_m;
  d->omega_l = p->omega_lambda;
  d->omega_r = p->omega_r;
  d->omega_k = p->omega_k;
  d->w_lambda = p->w_lambda;
  d->reion_efficiency = p->reion_efficiency;
  /* reionization.h:17-34 */
  const double h = p->hubble_h;
  const double rhocrit_cgs = 3.0 * K_HUBBLE * K_HUBBLE * h * h / (8.0 * M_PI * K_GRAVITY);
  const double OMb = p->baryon_frac * p->omega_m;
  d->No = rhocrit_cgs * OMb * (1 - p->y_he) / K_PROTONMASS;
  d->He_No = rhocrit_cgs * OMb * p->y_he / (4.0 * K_PROTONMASS);
  d->f_H = d->No / (d->No + d->He_No);
  d->f_He = d->He_No / (d->No + d->He_No);
  d->N_b0 = d->No + d->He_No;
  return MRX_OK;
}

/* ------------------------------------------------------- cosmology etc. */

double mrx_cosmo_dtdz(double omega_l, double omega_m, double hubble_h, float z)
{
  const double OMl = omega_l, OMm = omega_m;
  const double x = sqrt(OMl / OMm) * pow(1 + z, -3.0 / 2.0);
  const double dxdz = sqrt(OMl / OMm) * pow(1 + z, -5.0 / 2.0) * (-3.0 / 2.0);
  const double const1 = 2 * sqrt(1 + OMm / OMl) / (3.0 * K_HUBBLE * hubble_h);
  const double numer = dxdz * (1 + x * pow(pow(x, 2) + 1, -0.5));
  const double denom = x + sqrt(pow(x, 2) + 1);
  return (const1 * numer / denom);
}

double mrx_ts_dtdz(const mrx_ts_data* d, float z) { return mrx_cosmo_dtdz(d->omega_l, d->omega_m, d->hubble_h, z); }

double mrx_ts_hubble(const mrx_ts_data* d, float z)
{
  return K_HUBBLE * d->hubble_h * sqrt(d->omega_m * pow(1 + z, 3) + d->omega_r * pow(1 + z, 4) + d->omega_l);
}

double mrx_ts_drdz(const mrx_ts_data* d, float z) { return (1.0 + z) * K_SPEED_OF_LIGHT * mrx_ts_dtdz(d, z); }

double mrx_ts_dicke(const mrx_ts_data* d, double z)
{
  const double tiny = 1e-4;
  const double OMm = d->omega_m, OMl = d->omega_l, OMr = d->omega_r, OMk = d->omega_k, wl = d->w_lambda;
  const double OMtot = OMm + OMl + OMr + OMk;
  if (fabs(OMm - 1.0) < tiny) return 1.0 / (1.0 + z);
  if ((OMl > (-tiny)) && (fabs(OMl + OMm + OMr - 1.0) < 0.01) && (fabs(wl + 1.0) < tiny)) {
    const double omegaM_z = OMm * pow(1 + z, 3) / (OMl + OMm * pow(1 + z, 3) + OMr * pow(1 + z, 4));
    const double dick_z =
      2.5 * omegaM_z / (1.0 / 70.0 + omegaM_z * (209 - omegaM_z) / 140.0 + pow(omegaM_z, 4.0 / 7.0));
    const double dick_0 = 2.5 * OMm / (1.0 / 70.0 + OMm * (209 - OMm) / 140.0 + pow(OMm, 4.0 / 7.0));
    return dick_z / (dick_0 * (1.0 + z));
  }
  if ((OMtot < (1 + tiny)) && (fabs(OMl) < tiny)) {
    const double x_0 = 1.0 / (OMm + 0.0) - 1.0;
    const double dick_0 = 1 + 3.0 / x_0 + 3 * log(sqrt(1 + x_0) - sqrt(x_0)) * sqrt(1 + x_0) / pow(x_0, 1.5);
    const double x = fabs(1.0 / (OMm + 0.0) - 1.0) / (1 + z);
    const double dick_z = 1 + 3.0 / x + 3 * log(sqrt(1 + x) - sqrt(x)) * sqrt(1 + x) / pow(x, 1.5);
    return dick_z / dick_0;
  }
  return -1;
}

static double ddicke_dz(const mrx_ts_data* d, double z)
{
  const float dz = 1e-10; /* a float step, as XRayHeatingFunctions.c:1508 declares it */
  return (mrx_ts_dicke(d, z + dz) - mrx_ts_dicke(d, z)) / dz;
}

double mrx_ts_T_RECFAST(const mrx_ts_data* d, float z)
{
  if (z > d->recfast_zmax) return -1;
  return mrx_spline_eval(&d->recfast_T, z);
}

double mrx_ts_xion_RECFAST(const mrx_ts_data* d, float z)
{
  if (z > d->recfast_zmax) return -1;
  return mrx_spline_eval(&d->recfast_xe, z);
}

double mrx_ts_kappa_10(const mrx_ts_data* d, double TK)
{
  const mrx_spline* s = &d->kappa_HH;
  double ans;
  if (log(TK) < s->x[0])
    ans = s->y[0];
  else if (log(TK) > s->x[s->n - 1])
    ans = log(exp(s->y[s->n - 1]) * pow(TK / exp(s->x[s->n - 1]), 0.381));
  else
    ans = mrx_spline_eval(s, log(TK));
  return exp(ans);
}

static double kappa_10_file(const mrx_spline* s, double T)
{
  const int n = s->n;
  double ans;
  T = log(T);
  if (T < s->x[0])
    ans = s->y[0];
  else if (T > s->x[n - 1])
    ans = s->y[n - 1] + ((s->y[n - 1] - s->y[n - 2]) / (s->x[n - 1] - s->x[n - 2]) * (T - s->x[n - 1]));
  else
    ans = mrx_spline_eval(s, T);
  return exp(ans);
}

double mrx_ts_kappa_10_elec(const mrx_ts_data* d, double T) { return kappa_10_file(&d->kappa_eH, T); }
double mrx_ts_kappa_10_pH(const mrx_ts_data* d, double T) { return kappa_10_file(&d->kappa_pH, T); }

double mrx_ts_alpha_A(double T)
{
  const double logT = log(T / (double)1.1604505e4);
  return exp(-28.6130338 - 0.72411256 * logT - 2.02604473e-2 * pow(logT, 2) - 2.38086188e-3 * pow(logT, 3) -
             3.21260521e-4 * pow(logT, 4) - 1.42150291e-5 * pow(logT, 5) + 4.98910892e-6 * pow(logT, 6) +
             5.75561414e-7 * pow(logT, 7) - 1.85676704e-8 * pow(logT, 8) - 3.07113524e-9 * pow(logT, 9));
}

static float zmax(float z, int n)
{
  const double num = 1 - pow(n + 1, -2);
  const double denom = 1 - pow(n, -2);
  return (float)((1 + z) * num / denom - 1);
}

static double nu_n(int n) { return (1.0 - pow(n, -2.0)) / 0.75; }

static double frecycle(int n)
{
  /* fraction of Ly-n photons converted to Ly-alpha (XRayHeatingFunctions.c:516-583) */
  static const double f[31] = {1,      1,      1,      0,      0.2609, 0.3078, 0.3259, 0.3353,
                               0.3410, 0.3448, 0.3476, 0.3496, 0.3512, 0.3524, 0.3535, 0.3543,
                               0.3550, 0.3556, 0.3561, 0.3565, 0.3569, 0.3572, 0.3575, 0.3578,
                               0.3580, 0.3582, 0.3584, 0.3586, 0.3587, 0.3589, 0.3590};
  return (n >= 0 && n <= 30) ? f[n] : 0;
}

/* Pop II Ly-series spectrum at nu_norm (in Ly-alpha units), the flag-0 branch
 * of spectral_emissivity the Ly-n sum calls (XRayHeatingFunctions.c:655-670);
 * above the last tabulated line the un-normalised last segment applies. */
double mrx_ts_spectral_emissivity(const mrx_ts_data* d, double nu_norm)
{
  for (int i = 1; i < (MRX_NSPEC_MAX - 1); i++)
    if ((nu_norm >= d->spec_nu[i]) && (nu_norm < d->spec_nu[i + 1]))
      return d->spec_N0_2[i] * pow(nu_norm, d->spec_alpha_2[i]) / K_Ly_alpha_HZ;
  const int i = MRX_NSPEC_MAX - 1;
  return d->spec_N0_2[i] * pow(nu_norm, d->spec_alpha_2[i]) / K_Ly_alpha_HZ;
}

/* ---------------------------------------------------- x_int tables */

static int locate_energy_index(float En)
{
  if (En < 1008.88) return (int)(log(En / 10.0) / 1.98026273e-2);
  return 232 + (int)(log(En / 1008.88) / 9.53101798e-2);
}

int mrx_ts_locate_xHII_index(const mrx_ts_data* d, float xHII_call)
{
  int m = MRX_X_INT_NXHII - 1;
  while (xHII_call < d->xi_XHII[m]) m--;
  return m;
}

/* The energy side of the x_int interpolation: the clamped energy and its bin,
 * or -1 under the lowest tabulated energy (XRayHeatingFunctions.c:914-1192). */
static int xint_energy_bin(const mrx_ts_data* d, float* En)
{
  const float* E = d->xi_E;
  if (*En > 0.999 * E[MRX_X_INT_NENERGY - 1])
    *En = (float)(E[MRX_X_INT_NENERGY - 1] * 0.999);
  else if (*En < E[0])
    return -1;
  return locate_energy_index(*En);
}

/* The x_e side: bilinear interpolation in float at a clamped energy En in bin
 * n_low, as the interp_* functions do. */
static float xint_at(const mrx_ts_data* d, const float (*tab)[MRX_X_INT_NENERGY], float En, int n_low,
                     float xHII_call)
{
  const float* E = d->xi_E;
  if (xHII_call > d->xi_XHII[MRX_X_INT_NXHII - 1] * 0.999)
    xHII_call = (float)(d->xi_XHII[MRX_X_INT_NXHII - 1] * 0.999);
  else if (xHII_call < d->xi_XHII[0])
    xHII_call = (float)(1.001 * d->xi_XHII[0]);
  const int n_high = n_low + 1;
  const int m_low = mrx_ts_locate_xHII_index(d, xHII_call), m_high = m_low + 1;
  float elow = ((tab[m_low][n_high] - tab[m_low][n_low]) / (E[n_high] - E[n_low]));
  elow *= (En - E[n_low]);
  elow += tab[m_low][n_low];
  float ehigh = ((tab[m_high][n_high] - tab[m_high][n_low]) / (E[n_high] - E[n_low]));
  ehigh *= (En - E[n_low]);
  ehigh += tab[m_high][n_low];
  float r = (ehigh - elow) / (d->xi_XHII[m_high] - d->xi_XHII[m_low]);
  r *= (xHII_call - d->xi_XHII[m_low]);
  r += elow;
  return r;
}

/* bilinear interpolation of one x_int table in float, as the interp_*
 * functions do (XRayHeatingFunctions.c:914-1192); `below` is returned under
 * the lowest tabulated energy (1 for f_heat, 0 otherwise) */
float mrx_ts_interp_xint(const mrx_ts_data* d, const float (*tab)[MRX_X_INT_NENERGY], float En, float xHII_call,
                         float below)
{
  const int n_low = xint_energy_bin(d, &En);
  return (n_low < 0) ? below : xint_at(d, tab, En, n_low, xHII_call);
}

/* ---------------------------------------------------- cross sections */

static double HI_ion_crosssec(double nu)
{
  if (nu < K_NUIONIZATION) return 0;
  if (nu == K_NUIONIZATION) nu += K_TINY;
  const double Z = 1, eps = sqrt(nu / K_NUIONIZATION - 1);
  return K_SIGMA_HI / Z / Z * pow(K_NUIONIZATION / nu, 4) * exp(4 - (4 * atan(eps) / eps)) /
         (1 - exp(-2 * M_PI / eps));
}

static double HeII_ion_crosssec(double nu)
{
  if (nu < K_HeII_NUIONIZATION) return 0;
  if (nu == K_HeII_NUIONIZATION) nu += K_TINY;
  const double Z = 2, eps = sqrt(nu / K_HeII_NUIONIZATION - 1);
  return (6.3e-18) / Z / Z * pow(K_HeII_NUIONIZATION / nu, 4) * exp(4 - (4 * atan(eps) / eps)) /
         (1 - exp(-2 * M_PI / eps));
}

static double HeI_ion_crosssec(double nu)
{
  if (nu < K_HeI_NUIONIZATION) return 0;
  const double x = nu / K_NU_over_EV / 13.61 - 0.4434;
  const double y = sqrt(x * x + pow(2.136, 2));
  return 9.492e-16 * ((x - 1) * (x - 1) + 2.039 * 2.039) * pow(y, (0.5 * 3.188 - 5.5)) *
         pow(1.0 + sqrt(y / 1.469), -3.188);
}

static double species_weighted_xsec(const mrx_ts_data* d, double nu, double x_e)
{
  const double HI = d->f_H * (1 - x_e) * HI_ion_crosssec(nu);
  const double HeI = d->f_He * (1 - x_e) * HeI_ion_crosssec(nu);
  const double HeII = d->f_He * x_e * HeII_ion_crosssec(nu);
  return HI + HeI + HeII;
}

/* ------------------------------------------------------ integration */

static double rescale_error(double err, double result_abs, double result_asc)
{
  err = fabs(err);
  if (result_asc != 0 && err != 0) {
    const double scale = pow((200 * err / result_asc), 1.5);
    err = (scale < 1) ? result_asc * scale : result_asc;
  }
  if (result_abs > DBL_MIN / (50 * DBL_EPSILON)) {
    const double min_err = 50 * DBL_EPSILON * result_abs;
    if (min_err > err) err = min_err;
  }
  return err;
}

/* 61-point Gauss-Kronrod on [a, b]: result, error estimate, |f| and |f - mean| integrals */
static double gk61(double (*f)(double, void*), void* par, double a, double b, double* abserr, double* resabs,
                   double* resasc)
{
  const double center = 0.5 * (a + b), half = 0.5 * (b - a), abs_half = fabs(half);
  double fv1[31], fv2[31];
  const double fc = f(center, par);
  double res_g = 0, res_k = fc * mrx_gk61_wgk[30], res_abs = fabs(res_k);
  for (int j = 0; j < 15; j++) {
    const int jtw = 2 * j + 1;
    const double ab = half * mrx_gk61_xgk[jtw];
    const double f1 = f(center - ab, par), f2 = f(center + ab, par);
    fv1[jtw] = f1;
    fv2[jtw] = f2;
    res_g += mrx_gk61_wg[j] * (f1 + f2);
    res_k += mrx_gk61_wgk[jtw] * (f1 + f2);
    res_abs += mrx_gk61_wgk[jtw] * (fabs(f1) + fabs(f2));
  }
  for (int j = 0; j < 15; j++) {
    const int jtwm1 = 2 * j;
    const double ab = half * mrx_gk61_xgk[jtwm1];
    const double f1 = f(center - ab, par), f2 = f(center + ab, par);
    fv1[jtwm1] = f1;
    fv2[jtwm1] = f2;
    res_k += mrx_gk61_wgk[jtwm1] * (f1 + f2);
    res_abs += mrx_gk61_wgk[jtwm1] * (fabs(f1) + fabs(f2));
  }
  const double mean = res_k * 0.5;
  double res_asc = mrx_gk61_wgk[30] * fabs(fc - mean);
  for (int j = 0; j < 30; j++) res_asc += mrx_gk61_wgk[j] * (fabs(fv1[j] - mean) + fabs(fv2[j] - mean));
  const double err = (res_k - res_g) * half;
  res_k *= half;
  res_abs *= abs_half;
  res_asc *= abs_half;
  *resabs = res_abs;
  *resasc = res_asc;
  *abserr = rescale_error(err, res_abs, res_asc);
  return res_k;
}

/* Adaptive bisection of the interval with the largest error estimate until
 * the summed error is below epsrel*|result| (the QAG strategy, absolute
 * tolerance 0 as every call site in the reference uses). */
double mrx_ts_qag61(double (*f)(double, void*), void* par, double a, double b, double epsrel, int limit,
                    double* abserr)
{
  enum { MAXI = 1000 };
  if (limit > MAXI) limit = MAXI;
  static __thread double ia[MAXI], ib[MAXI], ir[MAXI], ie[MAXI];
  double resabs, resasc, e0;
  const double r0 = gk61(f, par, a, b, &e0, &resabs, &resasc);
  double tolerance = epsrel * fabs(r0);
  const double round_off = 50 * DBL_EPSILON * resabs;
  if ((e0 <= round_off && e0 > tolerance) || (e0 <= tolerance && e0 != resasc) || e0 == 0.0 || limit == 1) {
    if (abserr) *abserr = e0;
    return r0;
  }
  int n = 1;
  ia[0] = a;
  ib[0] = b;
  ir[0] = r0;
  ie[0] = e0;
  double area = r0, errsum = e0;
  int roundoff1 = 0, roundoff2 = 0;
  for (int iteration = 1; iteration < limit; iteration++) {
    int k = 0;
    for (int i = 1; i < n; i++)
      if (ie[i] > ie[k]) k = i;
    const double a1 = ia[k], b1 = 0.5 * (ia[k] + ib[k]), a2 = b1, b2 = ib[k];
    double e1, e2, ra1, rs1, ra2, rs2;
    const double r1 = gk61(f, par, a1, b1, &e1, &ra1, &rs1);
    const double r2 = gk61(f, par, a2, b2, &e2, &ra2, &rs2);
    const double area12 = r1 + r2, error12 = e1 + e2;
    errsum += (error12 - ie[k]);
    area += area12 - ir[k];
    if (rs1 != e1 && rs2 != e2) {
      const double delta = ir[k] - area12;
      if (fabs(delta) <= 1.0e-5 * fabs(area12) && error12 >= 0.99 * ie[k]) roundoff1++;
      if (iteration >= 10 && error12 > ie[k]) roundoff2++;
    }
    tolerance = epsrel * fabs(area);
    ib[k] = b1;
    ir[k] = r1;
    ie[k] = e1;
    ia[n] = a2;
    ib[n] = b2;
    ir[n] = r2;
    ie[n] = e2;
    n++;
    if (errsum <= tolerance) break;
    if (roundoff1 >= 6 || roundoff2 >= 20) break;
    if (fmax(fabs(a1), fabs(b2)) <= (1 + 100 * DBL_EPSILON) * (fabs(a2) + 1000 * DBL_MIN)) break;
  }
  double result = 0;
  for (int i = 0; i < n; i++) result += ir[i];
  if (abserr) *abserr = errsum;
  return result;
}

/* Brent's method on a bracketing interval until the bracket width is within
 * relerr of the smaller end (gsl_root_test_interval with epsabs = 0). */
double mrx_ts_brent(double (*f)(double, void*), void* par, double lo, double hi, double relerr, int max_iter)
{
  double a = lo, b = hi, fa = f(a, par), fb = f(b, par);
  double c = b, fc = fb, dd = b - a, e = dd;
  for (int it = 0; it < max_iter; it++) {
    if ((fb < 0 && fc < 0) || (fb > 0 && fc > 0)) {
      c = a;
      fc = fa;
      e = dd = b - a;
    }
    if (fabs(fc) < fabs(fb)) {
      a = b;
      b = c;
      c = a;
      fa = fb;
      fb = fc;
      fc = fa;
    }
    const double tol = 0.5 * DBL_EPSILON * fabs(b);
    const double m = 0.5 * (c - b);
    if (fb == 0) return b;
    if (fabs(m) <= tol) return b;
    if (fabs(e) < tol || fabs(fa) <= fabs(fb)) {
      dd = m;
      e = m;
    } else {
      double pp, q, r;
      const double s = fb / fa;
      if (a == c) {
        pp = 2 * m * s;
        q = 1 - s;
      } else {
        q = fa / fc;
        r = fb / fc;
        pp = s * (2 * m * q * (q - r) - (b - a) * (r - 1));
        q = (q - 1) * (r - 1) * (s - 1);
      }
      if (pp > 0)
        q = -q;
      else
        pp = -pp;
      if (2 * pp < fmin(3 * m * q - fabs(tol * q), fabs(e * q))) {
        e = dd;
        dd = pp / q;
      } else {
        dd = m;
        e = m;
      }
    }
    a = b;
    fa = fb;
    if (fabs(dd) > tol)
      b += dd;
    else
      b += (m > 0 ? +tol : -tol);
    fb = f(b, par);
    /* The bracket this iteration reports: the new point b and the end of
     * opposite sign - c, or a (the previous b) when f(b) has c's sign.  The
     * search stops when it is within relerr of its smaller-magnitude end. */
    const double cb = ((fb < 0 && fc < 0) || (fb > 0 && fc > 0)) ? a : c;
    const double xl = fmin(b, cb), xu = fmax(b, cb);
    const double min_abs = (xl > 0 && xu > 0) || (xl < 0 && xu < 0) ? fmin(fabs(xl), fabs(xu)) : 0;
    if (fabs(xu - xl) < relerr * min_abs) return b;
  }
  return b;
}

/* ------------------------------------------------------- tauX, nu_tau_one */

typedef struct {
  const mrx_ts_data* d;
  const mrx_ts_params* p;
  double nu_0, x_e, x_e_ave, ion_eff;
} tauX_par;

static double interpolate_fcoll(const mrx_ts_params* p, double redshift)
{
  const int s = p->snapshot;
  const double* Z = p->redshifts;
  const double* F = p->stored_fcoll;
  double v;
  if (s == 0)
    v = F[s];
  else
    v = F[s - 1] + (redshift - Z[s - 1]) * (F[s] - F[s - 1]) / (Z[s] - Z[s - 1]);
  return v < 0.0 ? 0.0 : v;
}

static double tauX_integrand(double zhat, void* vp)
{
  const tauX_par* q = (const tauX_par*)vp;
  const double drpropdz = K_SPEED_OF_LIGHT * mrx_ts_dtdz(q->d, (float)zhat);
  const double n = q->d->N_b0 * pow(1 + zhat, 3);
  const double nuhat = q->nu_0 * (1 + zhat);
  const double fcoll = interpolate_fcoll(q->p, zhat);
  double ff;
  if (fcoll < 1e-20)
    ff = 1;
  else
    ff = 1 - (q->ion_eff * fcoll) / (1.0 - q->x_e_ave);
  if (ff < 1e-4) ff = 1e-4;
  return drpropdz * n * ff * species_weighted_xsec(q->d, nuhat, q->x_e);
}

double mrx_ts_tauX(const mrx_ts_data* d, const mrx_ts_params* p, double nu, double x_e, double x_e_ave, double zp,
                   double zpp)
{
  tauX_par q = {d, p, nu / (1 + zp), x_e, x_e_ave, d->reion_efficiency};
  return mrx_ts_qag61(tauX_integrand, &q, zpp, zp, 0.005, 1000, NULL);
}

typedef struct {
  const mrx_ts_data* d;
  const mrx_ts_params* p;
  double x_e, x_e_ave, zp, zpp;
} ntau_par;

static double nu_tau_one_helper(double nu, void* vp)
{
  const ntau_par* q = (const ntau_par*)vp;
  return mrx_ts_tauX(q->d, q->p, nu, q->x_e, q->x_e_ave, q->zp, q->zpp) - 1;
}

double mrx_ts_nu_tau_one(const mrx_ts_data* d, const mrx_ts_params* p, double zp, double zpp, double x_e,
                         double x_e_ave)
{
  if (x_e > 0.9999) return -1;
  if (mrx_ts_tauX(d, p, K_HeI_NUIONIZATION, x_e, x_e_ave, zp, zpp) < 1) return K_HeI_NUIONIZATION;
  ntau_par q = {d, p, x_e, x_e_ave, zp, zpp};
  return mrx_ts_brent(nu_tau_one_helper, &q, K_HeI_NUIONIZATION, 1e6 * K_NU_over_EV, 0.02, 100);
}

/* ------------------------------------------------- frequency integrals */

/* What the frequency integrands need at a node nu that does not depend on
 * x_e: the HI, HeI and HeII cross sections, the spectrum factor, and per
 * species threshold the interpolation energy and its bin (-1: below). */
typedef struct {
  double nu;
  double xs[3], pw;
  float En[3];
  int bin[3];
} nu_node;

static const double k_nu_th[3] = {K_NUIONIZATION, K_HeI_NUIONIZATION, K_HeII_NUIONIZATION};

static void nu_node_eval(const mrx_ts_data* d, double thresh, double spec_index, double nu, nu_node* v)
{
  v->nu = nu;
  v->xs[0] = HI_ion_crosssec(nu);
  v->xs[1] = HeI_ion_crosssec(nu);
  v->xs[2] = HeII_ion_crosssec(nu);
  v->pw = pow(nu / (thresh * K_NU_over_EV), -spec_index - 1);
  for (int i = 0; i < 3; i++) {
    v->En[i] = (float)((nu - k_nu_th[i]) / K_NU_over_EV);
    v->bin[i] = xint_energy_bin(d, &v->En[i]);
  }
}

/* The integrals of one shell share their interval [lower, nu_max], and QAG
 * bisects it the same way for every x_e, so they evaluate mostly the same
 * nodes: the node values are kept in an open-addressing table keyed by the
 * bits of nu (the same doubles as evaluating them again). */
enum { NU_CACHE_BITS = 14, NU_CACHE_SIZE = 1 << NU_CACHE_BITS, NU_CACHE_FILL = NU_CACHE_SIZE / 2 };
typedef struct {
  int count;
  unsigned char used[NU_CACHE_SIZE];
  nu_node e[NU_CACHE_SIZE];
} nu_cache;

static void nu_cache_clear(nu_cache* c)
{
  c->count = 0;
  memset(c->used, 0, sizeof c->used);
}

typedef struct {
  const mrx_ts_data* d;
  double x_e, thresh, spec_index;
  nu_cache* cache; /* NULL: every node evaluated */
  nu_node scratch;
} nu_par;

static const nu_node* node_at(nu_par* q, double nu)
{
  nu_cache* c = q->cache;
  if (!c) {
    nu_node_eval(q->d, q->thresh, q->spec_index, nu, &q->scratch);
    return &q->scratch;
  }
  uint64_t key;
  memcpy(&key, &nu, sizeof key);
  unsigned h = (unsigned)((key * 0x9E3779B97F4A7C15ull) >> (64 - NU_CACHE_BITS));
  for (;;) {
    if (!c->used[h]) break;
    if (memcmp(&c->e[h].nu, &nu, sizeof nu) == 0) return &c->e[h];
    h = (h + 1) & (NU_CACHE_SIZE - 1);
  }
  if (c->count >= NU_CACHE_FILL) {
    nu_cache_clear(c);
    h = (unsigned)((key * 0x9E3779B97F4A7C15ull) >> (64 - NU_CACHE_BITS));
  }
  c->used[h] = 1;
  c->count++;
  nu_node_eval(q->d, q->thresh, q->spec_index, nu, &c->e[h]);
  return &c->e[h];
}

static float xint_node(const mrx_ts_data* d, const float (*tab)[MRX_X_INT_NENERGY], const nu_node* v, int i,
                       float xe, float below)
{
  return (v->bin[i] < 0) ? below : xint_at(d, tab, v->En[i], v->bin[i], xe);
}

static double nu_heat(double nu, void* vp)
{
  nu_par* q = (nu_par*)vp;
  const mrx_ts_data* d = q->d;
  const nu_node* v = node_at(q, nu);
  const float xe = (float)q->x_e;
  double s = xint_node(d, d->xi_fheat, v, 0, xe, 1.0f) * K_PLANCK * (nu - K_NUIONIZATION) * d->f_H * (1 - q->x_e) *
             v->xs[0];
  s += xint_node(d, d->xi_fheat, v, 1, xe, 1.0f) * K_PLANCK * (nu - K_HeI_NUIONIZATION) * d->f_He * (1 - q->x_e) *
       v->xs[1];
  s += xint_node(d, d->xi_fheat, v, 2, xe, 1.0f) * K_PLANCK * (nu - K_HeII_NUIONIZATION) * d->f_He * q->x_e *
       v->xs[2];
  return s * v->pw;
}

static double nion_sum(const mrx_ts_data* d, const nu_node* v, int i, float xe)
{
  return xint_node(d, d->xi_nHI, v, i, xe, 0.0f) + xint_node(d, d->xi_nHeI, v, i, xe, 0.0f) +
         xint_node(d, d->xi_nHeII, v, i, xe, 0.0f) + 1;
}

static double nu_ion(double nu, void* vp)
{
  nu_par* q = (nu_par*)vp;
  const mrx_ts_data* d = q->d;
  const nu_node* v = node_at(q, nu);
  const float xe = (float)q->x_e;
  double F_i = nion_sum(d, v, 0, xe);
  double s = F_i * d->f_H * (1 - q->x_e) * v->xs[0];
  F_i = nion_sum(d, v, 1, xe);
  s += F_i * d->f_He * (1 - q->x_e) * v->xs[1];
  F_i = nion_sum(d, v, 2, xe);
  s += F_i * d->f_He * q->x_e * v->xs[2];
  return s * v->pw;
}

static double nu_lya(double nu, void* vp)
{
  nu_par* q = (nu_par*)vp;
  const mrx_ts_data* d = q->d;
  const nu_node* v = node_at(q, nu);
  const float xe = (float)q->x_e;
  double s = xint_node(d, d->xi_nLya, v, 0, xe, 0.0f) * d->f_H * (1 - q->x_e) * v->xs[0];
  s += xint_node(d, d->xi_nLya, v, 1, xe, 0.0f) * d->f_He * (1 - q->x_e) * v->xs[1];
  s += xint_node(d, d->xi_nLya, v, 2, xe, 0.0f) * d->f_He * q->x_e * v->xs[2];
  return s * v->pw;
}

static double integrate_over_nu(const mrx_ts_data* d, const mrx_ts_params* p, double zp, double local_x_e,
                                double lower_int_limit, int flag, nu_cache* cache)
{
  nu_par q = {d, local_x_e, p->nu_xray_gal_threshold, p->spec_index_xray_gal, cache, {0}};
  double (*f)(double, void*) = (flag == 0) ? nu_heat : (flag == 1 ? nu_ion : nu_lya);
  const double result = mrx_ts_qag61(f, &q, lower_int_limit, p->nu_xray_max * K_NU_over_EV, 0.01, 1000, NULL);
  if (flag == 2) return result * K_SPEED_OF_LIGHT / (4 * M_PI) / K_Ly_alpha_HZ / mrx_ts_hubble(d, (float)zp);
  return result;
}

double mrx_ts_integrate_over_nu(const mrx_ts_data* d, const mrx_ts_params* p, double zp, double local_x_e,
                                double lower_int_limit, int flag)
{
  return integrate_over_nu(d, p, zp, local_x_e, lower_int_limit, flag, NULL);
}

/* -------------------------------------------------- per-snapshot plan */

static void push_op(mrx_ts_snap* s, int ct, int kind, int slot, double w)
{
  if (s->n_ops[ct] < MRX_TS_MAX_SHELL_OPS) {
    mrx_sfr_op* o = &s->ops[ct][s->n_ops[ct]++];
    o->kind = kind;
    o->slot = slot;
    o->weight = w;
  } else {
    s->ops_overflow = 1;
  }
  if ((kind == MRX_SFR_LOAD || kind == MRX_SFR_ADD) && slot > s->max_slot) s->max_slot = slot;
}

int mrx_ts_plan(mrx_ts_snap* s, const mrx_ts_data* d, const mrx_ts_params* p)
{
  memset(s, 0, sizeof *s);
  const int nsteps = p->num_filter_steps;
  if (nsteps < 1 || nsteps > MRX_TS_MAX_STEPS || p->snapshot < 0 || !p->redshifts) return MRX_EINVAL;
  const int snapshot = p->snapshot;
  const double* ZZ = p->redshifts;
  s->nsteps = nsteps;
  s->max_slot = -1;
  const double redshift = ZZ[snapshot];
  const double prev_redshift = (snapshot == 0) ? ZZ[snapshot] : ZZ[snapshot - 1];
  s->zp = redshift;
  s->dzp = s->zp - prev_redshift;
  s->dt_dzp = mrx_ts_dtdz(d, (float)(float)s->zp);
  s->homogeneous = ((s->zp - p->max_heating_redshift) >= -0.0001);
  if (s->homogeneous) return MRX_OK;

  const double box_size = p->box_size / p->hubble_h; /* Mpc */
  const int N = p->grid_dim;
  double R = K_L_FACTOR * box_size / (float)N;
  const double R_factor = pow(K_R_XLy_MAX / R, 1 / (float)nsteps);
  int scb[MRX_TS_MAX_STEPS];
  scb[0] = 1;
  double weight = 1;
  const double zp = s->zp;
  for (int ct = 0; ct < nsteps; ct++) {
    double prev_zpp, prev_R;
    s->R[ct] = R;
    if (ct == 0) {
      prev_zpp = zp;
      prev_R = 0;
    } else {
      prev_zpp = s->zpp_edge[ct - 1];
      prev_R = s->R[ct - 1];
      scb[ct] = scb[ct - 1];
    }
    s->zpp_edge[ct] = prev_zpp - (R - prev_R) * K_MPC / (mrx_ts_drdz(d, (float)prev_zpp));
    const double zpp = (s->zpp_edge[ct] + prev_zpp) * 0.5;
    const double total_weight = (prev_zpp - s->zpp_edge[ct]) * mrx_ts_dtdz(d, (float)zpp);
    if (snapshot - scb[ct] >= 0) {
      if (snapshot - scb[ct] == 0 && weight < 0) {
        push_op(s, ct, MRX_SFR_ZERO_HALF, 0, 0);
        scb[ct] += 1;
      } else if (s->zpp_edge[ct] > ZZ[snapshot - scb[ct]]) {
        if (ct == 0) {
          weight = s->dt_dzp * s->dzp / total_weight;
          push_op(s, ct, MRX_SFR_SCALE_HALF, 0, weight);
        } else {
          if (ZZ[snapshot - scb[ct]] > s->zpp_edge[ct - 1]) {
            weight = -mrx_ts_dtdz(d, (float)(0.5 * (ZZ[snapshot - scb[ct]] + s->zpp_edge[ct - 1])));
            weight *= (ZZ[snapshot - scb[ct]] - s->zpp_edge[ct - 1]);
            weight /= total_weight;
            push_op(s, ct, MRX_SFR_LOAD, scb[ct] - 1, weight);
          } else {
            push_op(s, ct, MRX_SFR_ZERO_HALF, 0, 0);
          }
        }
        while (s->zpp_edge[ct] > ZZ[snapshot - scb[ct]]) {
          if (snapshot - scb[ct] == 0) {
            weight = -1;
            break;
          }
          if (s->zpp_edge[ct] < ZZ[snapshot - scb[ct] - 1]) {
            /* ComputeTs.c:345: the ternary's condition is zpp_edge[R_ct-1] itself,
             * non-zero for every shell, so zedge is the snapshot redshift; for
             * R_ct = 0 the reference reads zpp_edge[-1], taken here as non-zero. */
            const double prev_edge = (ct > 0) ? s->zpp_edge[ct - 1] : 1.0;
            const double zedge = prev_edge ? ZZ[snapshot - scb[ct]] : (double)(prev_edge > ZZ[snapshot - scb[ct]]);
            weight = s->zpp_edge[ct] - zedge;
            weight *= -mrx_ts_dtdz(d, (float)(0.5 * (s->zpp_edge[ct] + zedge)));
          } else {
            weight = ZZ[snapshot - scb[ct] - 1] - ZZ[snapshot - scb[ct]];
            weight *= -mrx_ts_dtdz(d, (float)(0.5 * (ZZ[snapshot - scb[ct] - 1] + ZZ[snapshot - scb[ct]])));
          }
          weight /= total_weight;
          push_op(s, ct, MRX_SFR_ADD, scb[ct], weight);
          scb[ct] += 1;
        }
      } else {
        if (ct > 0 && weight != 1) {
          weight = 1;
          push_op(s, ct, MRX_SFR_LOAD, scb[ct] - 1, weight);
        }
      }
    }
    R *= R_factor;
  }
  s->final_weight = weight;
  if (s->ops_overflow) return MRX_EUNSUPPORTED;
  /* per-shell factors of evolveInt (XRayHeatingFunctions.c:1291-1303) */
  for (int ct = 0; ct < nsteps; ct++) {
    if (ct == 0) {
      s->zpp[ct] = (s->zpp_edge[0] + (float)zp) * 0.5;
      s->dzpp[ct] = (float)zp - s->zpp_edge[0];
    } else {
      s->zpp[ct] = (s->zpp_edge[ct] + s->zpp_edge[ct - 1]) * 0.5;
      s->dzpp[ct] = s->zpp_edge[ct - 1] - s->zpp_edge[ct];
    }
    s->dt_dzpp[ct] = mrx_ts_dtdz(d, (float)s->zpp[ct]);
    s->zpp_pow[ct] = pow(1 + s->zpp[ct], -p->spec_index_xray_gal);
  }
  return MRX_OK;
}

/* The shells' tables are independent of one another: one work item is a
 * shell, its lower frequency limit, raw Ly-n sum and 3 x 14 frequency
 * integrals (which share one table of node values); items are handed out one
 * at a time to up to 16 host threads.  Each item is computed exactly as the
 * serial loop would, so the tables do not depend on the thread count. */
typedef struct {
  mrx_ts_snap* s;
  const mrx_ts_data* d;
  const mrx_ts_params* p;
  int nitems;
  int next; /* next unclaimed work item */
} tab_job;

static void tab_item(tab_job* j, int ct, nu_cache* cache)
{
  mrx_ts_snap* s = j->s;
  const mrx_ts_data* d = j->d;
  const mrx_ts_params* p = j->p;
  const double zp = s->zp;
  const double prev_zpp = (ct == 0) ? zp : s->zpp_edge[ct - 1];
  const double zpp = (s->zpp_edge[ct] + prev_zpp) * 0.5;
  const double lower =
    fmax(mrx_ts_nu_tau_one(d, p, zp, zpp, s->x_e_ave, s->x_e_ave), p->nu_xray_gal_threshold * K_NU_over_EV);
  s->nu_lower[ct] = lower;
  double sum = 0;
  for (int n_ct = MRX_NSPEC_MAX; n_ct >= 2; n_ct--) {
    if (zpp > zmax((float)zp, n_ct)) continue;
    const double nuprime = nu_n(n_ct) * (1 + zpp) / (1.0 + zp);
    sum += frecycle(n_ct) * mrx_ts_spectral_emissivity(d, nuprime);
  }
  s->sum_lyn[ct] = sum;
  if (cache) nu_cache_clear(cache);
  for (int x = 0; x < MRX_X_INT_NXHII; x++) {
    s->freq_heat[x][ct] = integrate_over_nu(d, p, zp, d->xi_XHII[x], lower, 0, cache);
    s->freq_ion[x][ct] = integrate_over_nu(d, p, zp, d->xi_XHII[x], lower, 1, cache);
    s->freq_lya[x][ct] = integrate_over_nu(d, p, zp, d->xi_XHII[x], lower, 2, cache);
  }
}

static void* tab_worker(void* vp)
{
  tab_job* j = (tab_job*)vp;
  /* without the table (allocation failed) every node is evaluated: the same values */
  nu_cache* cache = (nu_cache*)malloc(sizeof(nu_cache));
  for (;;) {
    const int it = __atomic_fetch_add(&j->next, 1, __ATOMIC_RELAXED);
    if (it >= j->nitems) break;
    tab_item(j, it, cache);
  }
  free(cache);
  return NULL;
}

static void tab_run(tab_job* j)
{
  enum { MAXT = 16 };
  long nt = sysconf(_SC_NPROCESSORS_ONLN);
  if (nt > MAXT) nt = MAXT;
  if (nt > j->nitems) nt = j->nitems;
  pthread_t th[MAXT];
  int started = 0;
  for (int t = 1; t < nt; t++)
    if (pthread_create(&th[started], NULL, tab_worker, j) == 0) started++;
  tab_worker(j); /* the calling thread works too, and alone if no thread started */
  for (int t = 0; t < started; t++) pthread_join(th[t], NULL);
}

int mrx_ts_build_tables(mrx_ts_snap* s, const mrx_ts_data* d, const mrx_ts_params* p)
{
  if (s->homogeneous) return MRX_OK;
  const double zp = s->zp;
  const double collapse_fraction = s->collapse_fraction;
  if (p->stored_fcoll) p->stored_fcoll[p->snapshot] = collapse_fraction;
  s->no_light = (collapse_fraction > 0.0) ? 0 : 1;
  int first_radii = 1, first_zero = 1;
  const int n_pts_radii = 1000;
  double weight = s->final_weight; /* the shell loop's last weight (one variable in the reference) */
  tab_job job = {s, d, p, s->nsteps, 0};
  tab_run(&job);
  for (int ct = 0; ct < s->nsteps; ct++) {
    double prev_zpp, prev_R;
    if (ct == 0) {
      prev_zpp = zp;
      prev_R = 0;
    } else {
      prev_zpp = s->zpp_edge[ct - 1];
      prev_R = s->R[ct - 1];
    }
    /* smooth the kink where the Ly-n contribution first vanishes (ComputeTs.c:625-663) */
    if (ct > 2 && s->sum_lyn[ct] == 0.0 && s->sum_lyn[ct - 1] > 0. && first_radii) {
      const double trial_zpp_max =
        (prev_zpp - (s->R[ct] - prev_R) * K_MPC / mrx_ts_drdz(d, (float)prev_zpp) + prev_zpp) * 0.5;
      const double trial_zpp_min = (s->zpp_edge[ct - 2] -
                                    (s->R[ct - 1] - s->R[ct - 2]) * K_MPC / mrx_ts_drdz(d, (float)s->zpp_edge[ct - 2]) +
                                    s->zpp_edge[ct - 2]) *
                                   0.5;
      for (int ii = 0; ii < n_pts_radii; ii++) {
        const double trial_zpp =
          trial_zpp_min + (trial_zpp_max - trial_zpp_min) * (float)ii / ((float)n_pts_radii - 1.);
        int counter = 0;
        for (int n_ct = MRX_NSPEC_MAX; n_ct >= 2; n_ct--) {
          if (trial_zpp > zmax(zp, n_ct)) continue;
          counter += 1;
        }
        if (counter == 0 && first_zero) {
          first_zero = 0;
          weight = (float)ii / (float)n_pts_radii;
        }
      }
      s->sum_lyn[ct] = weight * s->sum_lyn[ct - 1];
      first_radii = 0;
    }
  }
  s->growth_factor_zp = mrx_ts_dicke(d, zp);
  s->dgrowth_factor_dzp = ddicke_dz(d, zp);
  s->dt_dzp = mrx_ts_dtdz(d, (float)zp);
  /* soft-band luminosity -> X-ray photon number (ComputeTs.c:683-713) */
  double conv;
  const double th = p->nu_xray_gal_threshold, al = p->spec_index_xray_gal;
  if (fabs(al - 1.0) < 0.000001) {
    conv = (th * K_NU_over_EV) * log(p->nu_xray_soft_cut / th);
    conv = 1. / conv;
  } else {
    conv = pow(p->nu_xray_soft_cut * K_NU_over_EV, 1. - al) - pow(th * K_NU_over_EV, 1. - al);
    conv = 1. / conv;
    conv *= pow(th * K_NU_over_EV, -al) * (1 - al);
  }
  conv *= (K_SEC_PER_YEAR) / (K_PLANCK);
  s->const_zp_prefactor = (p->l_xray_gal * conv) / (th * K_NU_over_EV) * K_SPEED_OF_LIGHT * pow(1 + zp, al + 3);
  return MRX_OK;
}
