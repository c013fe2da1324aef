/*
 * recomb_tables.c - host tables of the inhomogeneous recombinations, restating
 * src/core/recombinations.c (itself from 21cmFAST): the MHR00 density PDF
 * P(D) = A exp(-(D^(-2/3) - C)^2 / (2 (2*7.61/(3(1+z)))^2)) D^beta with A
 * from its normalisation (:159-200) and C, beta splined from fixed knots
 * (:219-300); the self-shielded photoionisation rate Gamma_SS (:103-108,
 * Rahmati et al. 2013); the neutral fraction of a cell in ionisation
 * equilibrium (:321-352); the recombination rate per baryon integrated over
 * the PDF in ln D (:115-156) and its table in (z, ln Gamma12) with a natural
 * cubic spline per redshift (:56-82); and splined_recombination_rate
 * (:30-54).  GSL's QAG-61 and cspline are this build's mrx_ts_qag61 and
 * mrx_spline (ts_tables.c).  Compile with -ffp-contract=off.
 */
#include "recomb_tables.h"

#include <math.h>
#include <pthread.h>
#include <string.h>
#include <unistd.h>

#define K_HUBBLE 3.2407789e-18
#define K_GRAVITY 6.672e-8
#define K_PROTONMASS 1.6726e-24
#define K_ALPHA_B_10K 2.59e-13 /* alphaB_10k, reionization.h:41 */
#define K_TINY 1e-30
#define A_NPTS 60 /* recombinations.h */
#define C_NPTS 12
#define BETA_NPTS 5

double mrx_rr_No(double hubble_h, double omega_m, double baryon_frac, double y_he)
{
  const double rhocrit_cgs = 3.0 * K_HUBBLE * K_HUBBLE * hubble_h * hubble_h / (8.0 * M_PI * K_GRAVITY);
  const double OMb = baryon_frac * omega_m;
  return rhocrit_cgs * OMb * (1 - y_he) / K_PROTONMASS;
}

/* ------------------------------------------------------ MHR parameters */

double mrx_rr_C_MHR(const mrx_rr_tables* t, double z)
{
  if (z >= 13.0) return 1.0;
  if (z <= 2.0) return 0.558;
  return mrx_spline_eval(&t->C, z);
}

double mrx_rr_beta_MHR(const mrx_rr_tables* t, double z)
{
  if (z >= 6.0) return -2.50;
  if (z <= 2.0) return -2.23;
  return mrx_spline_eval(&t->beta, z);
}

double mrx_rr_A_MHR(const mrx_rr_tables* t, double z)
{
  if (z >= 2.0 + (float)A_NPTS) return mrx_spline_eval(&t->A, 2.0 + (float)A_NPTS);
  if (z <= 2.0) return mrx_spline_eval(&t->A, 2.0);
  return mrx_spline_eval(&t->A, z);
}

typedef struct {
  const mrx_rr_tables* t;
  double z;
} aux_par;

/* the unnormalised PDF in D (:165-175) */
static double aux_function(double D, void* vp)
{
  const aux_par* q = (const aux_par*)vp;
  const double z = q->z;
  const double Cz = mrx_rr_C_MHR(q->t, z);
  return exp(-(pow(D, -2.0 / 3.0) - Cz) * (pow(D, -2.0 / 3.0) - Cz) /
             (2.0 * (2.0 * 7.61 / (3.0 * (1.0 + z))) * (2.0 * 7.61 / (3.0 * (1.0 + z))))) *
         pow(D, mrx_rr_beta_MHR(q->t, z));
}

double mrx_rr_A_aux_integral(const mrx_rr_tables* t, double z)
{
  aux_par q = {t, z};
  return mrx_ts_qag61(aux_function, &q, 1e-25, 1e25, 0.001, 1000, NULL);
}

/* --------------------------------------------------- rate integrand */

double mrx_rr_Gamma_SS(double Gamma_bg, double Delta, double T_4, double z)
{
  const double D_ss = 26.7 * pow(T_4, 0.17) * pow((1 + z) / 10.0, -3) * pow(Gamma_bg, 2.0 / 3.0);
  return Gamma_bg * (0.98 * pow((1.0 + pow(Delta / D_ss, 1.64)), -2.28) + 0.02 * pow(1.0 + Delta / D_ss, -0.84));
}

static double alpha_B(double T) { return K_ALPHA_B_10K * pow(T / 1.0e4, -0.75); }

double mrx_rr_neutral_fraction(double y_he, double density, double T4, double gamma12, int usecaseB)
{
  const double corr_He = 1.0 / (4.0 / y_he - 3);
  const double alpha = usecaseB ? alpha_B(T4 * 1e4) : mrx_ts_alpha_A(T4 * 1e4);
  gamma12 *= 1e-12;
  double chi = (1 + corr_He) * density * alpha / gamma12;
  if (chi < K_TINY) return 0;
  if (chi < 1e-5) return chi;
  const double b = -2 - gamma12 / (density * (1 + corr_He) * alpha);
  chi = (-b - sqrt(b * b - 4)) / 2.0;
  return chi;
}

typedef struct {
  double z, gamma12_bg, T4, A, C_0, beta, avenH, y_he;
  int usecaseB;
} rr_par;

static double MHR_rr(double lnD, void* vp)
{
  const rr_par* p = (const rr_par*)vp;
  const double D = exp(lnD);
  const double z = p->z;
  const double gamma = mrx_rr_Gamma_SS(p->gamma12_bg, D, p->T4, z);
  const double n_H = p->avenH * D;
  const double x_e = 1.0 - mrx_rr_neutral_fraction(p->y_he, n_H, p->T4, gamma, p->usecaseB);
  const double PDelta = p->A * exp(-0.5 * pow((pow(D, -2.0 / 3.0) - p->C_0) / ((2.0 * 7.61 / (3.0 * (1.0 + z)))), 2)) *
                        pow(D, p->beta);
  const double alpha = p->usecaseB ? alpha_B(p->T4 * 1e4) : mrx_ts_alpha_A(p->T4 * 1e4);
  return n_H * PDelta * alpha * x_e * x_e * D * D; /* the extra D of the integral over ln D */
}

double mrx_rr_recombination_rate(const mrx_rr_tables* t, double z_eff, double gamma12_bg, double T4, int usecaseB)
{
  rr_par p = {z_eff,
              gamma12_bg,
              T4,
              mrx_rr_A_MHR(t, z_eff),
              mrx_rr_C_MHR(t, z_eff),
              mrx_rr_beta_MHR(t, z_eff),
              t->No * pow(1 + z_eff, 3),
              t->y_he,
              usecaseB};
  return mrx_ts_qag61(MHR_rr, &p, log(0.01), log(200), 0.01, 1000, NULL);
}

/* ------------------------------------------------------------ tables */

typedef struct {
  mrx_rr_tables* t;
  int next;
} rr_job;

static void* rr_worker(void* vp)
{
  rr_job* j = (rr_job*)vp;
  mrx_rr_tables* t = j->t;
  for (;;) {
    const int it = __atomic_fetch_add(&j->next, 1, __ATOMIC_RELAXED);
    if (it >= MRX_RR_Z_NPTS * MRX_RR_LNG_NPTS) return NULL;
    const int z_ct = it / MRX_RR_LNG_NPTS, g = it % MRX_RR_LNG_NPTS;
    const float z = z_ct * MRX_RR_DEL_Z;
    const float gamma = (float)exp(t->lnG[g]);
    t->y[z_ct][g] = mrx_rr_recombination_rate(t, z, gamma, 1, 1);
  }
}

int mrx_rr_init(mrx_rr_tables* t, double No, double y_he)
{
  memset(t, 0, sizeof *t);
  t->No = No;
  t->y_he = y_he;
  /* C and beta knots (:219-236, :259-276) */
  static const double Ct[C_NPTS] = {0.558, 0.599, 0.611, 0.769, 0.868, 0.930, 0.964, 0.983, 0.993, 0.998, 0.999, 1.00};
  static const double Bt[BETA_NPTS] = {-2.23, -2.35, -2.48, -2.49, -2.50};
  double x[A_NPTS], y[A_NPTS];
  for (int i = 0; i < C_NPTS; i++) x[i] = (float)i + 2.0;
  if (mrx_spline_init(&t->C, x, Ct, C_NPTS)) return -1;
  if (mrx_spline_init(&t->beta, x, Bt, BETA_NPTS)) return -1;
  /* A: 1 / the PDF's normalisation integral (:187-201); the integrand reads C
   * and beta only */
  for (int i = 0; i < A_NPTS; i++) {
    x[i] = 2.0 + (float)i;
    y[i] = 1.0 / mrx_rr_A_aux_integral(t, 2.0 + (float)i);
  }
  if (mrx_spline_init(&t->A, x, y, A_NPTS)) return -1;
  for (int g = 0; g < MRX_RR_LNG_NPTS; g++) t->lnG[g] = MRX_RR_LNG_MIN + g * MRX_RR_DEL_LNG;
  /* the rate table, item by item on up to 16 threads */
  rr_job job = {t, 0};
  enum { MAXT = 16 };
  long nt = sysconf(_SC_NPROCESSORS_ONLN);
  if (nt > MAXT) nt = MAXT;
  if (nt < 1) nt = 1;
  pthread_t th[MAXT];
  int started = 0;
  for (int k = 1; k < nt; k++)
    if (pthread_create(&th[started], NULL, rr_worker, &job) == 0) started++;
  rr_worker(&job);
  for (int k = 0; k < started; k++) pthread_join(th[k], NULL);
  /* a natural cubic spline in ln Gamma12 per redshift (:73-76) */
  static __thread mrx_spline s;
  for (int z_ct = 0; z_ct < MRX_RR_Z_NPTS; z_ct++) {
    if (mrx_spline_init(&s, t->lnG, t->y[z_ct], MRX_RR_LNG_NPTS)) return -1;
    for (int g = 0; g < MRX_RR_LNG_NPTS; g++) {
      t->b[z_ct][g] = s.b[g];
      t->c[z_ct][g] = s.c[g];
      t->d[z_ct][g] = s.d[g];
    }
  }
  return 0;
}

double mrx_rr_rate(const mrx_rr_tables* t, float z_eff, double gamma12_bg)
{
  int z_ct = (int)(z_eff / MRX_RR_DEL_Z + 0.5);
  double lnGamma = log(gamma12_bg);
  if (z_ct < 0)
    z_ct = 0;
  else if (z_ct >= MRX_RR_Z_NPTS)
    z_ct = MRX_RR_Z_NPTS - 1;
  if (lnGamma < MRX_RR_LNG_MIN) return 0;
  if (lnGamma >= (MRX_RR_LNG_MIN + MRX_RR_DEL_LNG * MRX_RR_LNG_NPTS))
    lnGamma = MRX_RR_LNG_MIN + MRX_RR_DEL_LNG * MRX_RR_LNG_NPTS - 1e-7;
  /* gsl_spline_eval: the knot interval by bisection, then the cubic */
  int lo = 0, hi = MRX_RR_LNG_NPTS - 1;
  if (lnGamma >= t->lnG[hi]) {
    lo = hi - 1;
  } else {
    while (hi - lo > 1) {
      const int mid = (lo + hi) / 2;
      if (t->lnG[mid] > lnGamma)
        hi = mid;
      else
        lo = mid;
    }
  }
  const double dx = lnGamma - t->lnG[lo];
  return t->y[z_ct][lo] + dx * (t->b[z_ct][lo] + dx * (t->c[z_ct][lo] + dx * t->d[z_ct][lo]));
}
