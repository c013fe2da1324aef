/*
 * ts.c - the oracle's restatement of _ComputeTs (src/core/ComputeTs.c:78-1020)
 * with minihalos, Lyman-Werner and peculiar velocities off.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The per-snapshot host tables (shell radii, SFR-history operations,
 * frequency integrals, sum_lyn, prefactors, spline coefficients) arrive in an
 * orc_ts_tab built by the oracle's own restatement oracle/ts_host.py; this
 * file shares no code with the library.  Everything per cell is restated here:
 *   shell loop: ComputeTs.c:280-503 (working SFR grid ops, R2C, /N^3 in float,
 *     filter for R_ct > 0 with R in Mpc and BoxSize in Mpc/h, C2R, max(v,0));
 *   collapse fraction and <x_e> of the first shell: :225-257, :413-475;
 *   per-cell update: :743-900 with evolveInt (XRayHeatingFunctions.c:1241-1435),
 *     get_Ts and helpers (:1438-1882), kappa splines (:1602-1845);
 *   homogeneous branch: :189-262; averages: :940-985.
 * Build with -ffp-contract=off.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define O_TCMB 2.728
#define O_T21 0.0628
#define O_A10 2.85e-15
#define O_HUBBLE 3.2407789e-18
#define O_BOLTZMANN 1.3806e-16
#define O_SPEED_OF_LIGHT 2.9979e10
#define O_PROTONMASS 1.6726e-24
#define O_SOLAR_MASS 1.989e33
#define O_CLUMPING 2.0
#define O_FRACT_FLOAT_ERR 1e-7
#define O_MAX_TK ((float)5e4)
#define O_L_FACTOR 0.620350491

/* ------------------------------------------------------ per-cell physics */

static double o_hubble(const mrx_ts_params* p, float z)
{
  return O_HUBBLE * p->hubble_h * sqrt(p->omega_m * pow(1 + z, 3) + p->omega_r * pow(1 + z, 4) + p->omega_lambda);
}

/* gsl_spline_eval of a natural cubic spline (interval by bisection) */
static double o_spline(const orc_spline* s, double xv)
{
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

static double o_kappa_HH(const orc_ts_tab* t, double TK)
{
  const orc_spline* s = &t->kappa_HH;
  double ans;
  if (log(TK) < s->x[0])
    ans = s->y[0];
  else if (log(TK) > s->x[s->n - 1])
    ans = log(exp(s->y[s->n - 1]) * pow(TK / exp(s->x[s->n - 1]), 0.381));
  else
    ans = o_spline(s, log(TK));
  return exp(ans);
}

static double o_kappa_file(const orc_spline* s, double T)
{
  const int n = s->n;
  double ans;
  T = log(T);
  if (T < s->x[0])
    ans = s->y[0];
  else if (T > s->x[n - 1])
    ans = s->y[n - 1] + ((s->y[n - 1] - s->y[n - 2]) / (s->x[n - 1] - s->x[n - 2]) * (T - s->x[n - 1]));
  else
    ans = o_spline(s, T);
  return exp(ans);
}

/* alpha_A, XRayHeatingFunctions.c:1450-1458 */
static double o_alpha_A(double T)
{
  const double logT = log(T / (double)1.1604505e4);
  return exp(-28.6130338 - 0.72411256 * logT - 2.02604473e-2 * pow(logT, 2) - 2.38086188e-3 * pow(logT, 3) -
             3.21260521e-4 * pow(logT, 4) - 1.42150291e-5 * pow(logT, 5) + 4.98910892e-6 * pow(logT, 6) +
             5.75561414e-7 * pow(logT, 7) - 1.85676704e-8 * pow(logT, 8) - 3.07113524e-9 * pow(logT, 9));
}

/* locate_xHII_index, XRayHeatingFunctions.c:1210-1220 */
static int o_locate_xHII(const orc_ts_tab* t, float x)
{
  int m = ORC_XHII - 1;
  while (x < t->xHII[m]) m--;
  return m;
}

static double o_xcoll(const orc_ts_tab* d, double z, double TK, double delta, double xe)
{
  const double Trad = O_TCMB * (1.0 + z);
  const double nH = (1.0 - xe) * d->No * pow(1.0 + z, 3.0) * (1.0 + delta);
  const double x_HI = O_T21 / Trad * nH * o_kappa_HH(d, TK) / O_A10;
  const double ne = xe * d->N_b0 * pow(1.0 + z, 3.0) * (1.0 + delta);
  const double x_e = O_T21 / Trad * ne * o_kappa_file(&d->kappa_eH, TK) / O_A10;
  const double np = xe * d->No * pow(1.0 + z, 3.0) * (1.0 + delta);
  const double x_p = O_T21 / Trad * np * o_kappa_file(&d->kappa_pH, TK) / O_A10;
  return x_HI + x_e + x_p;
}

static double o_xalpha_tilde(const mrx_ts_params* p, const orc_ts_tab* d, double z, double Jalpha, double TK,
                             double TS, double delta, double xe)
{
  const double tgp = 1.342881e-7 / o_hubble(p, (float)z) * d->No * pow(1 + z, 3) * (1.0 + delta) * (1.0 - xe);
  const double xi = pow(1.0e-7 * tgp / TK / TK, 1.0 / 3.0);
  double S = 1.0 - 0.0631789 / TK + 0.115995 / TK / TK - 0.401403 / TS / TK;
  S += 0.336463 / TS / TK / TK;
  S /= 1.0 + 2.98394 * xi + 1.53583 * xi * xi + 3.85289 * xi * xi * xi;
  return 1.66e11 / (1.0 + z) * S * Jalpha;
}

static double o_Tc_eff(double TK, double TS)
{
  double ans = 1.0 / TK + 0.405535 / TK * (1.0 / TS - 1.0 / TK);
  return 1.0 / ans;
}

static float o_get_Ts(const mrx_ts_params* p, const orc_ts_tab* d, float z, float delta, float TK, float xe,
                      float Jalpha, float* xa_out)
{
  const double Trad = O_TCMB * (1.0 + z);
  const double xc = o_xcoll(d, z, TK, delta, xe);
  double TS;
  if (Jalpha > 1.0e-20) {
    double TSold = 0.0, xa = 0.0;
    TS = Trad;
    while (fabs(TS - TSold) / TS > 1.0e-3) {
      TSold = TS;
      xa = o_xalpha_tilde(p, d, z, Jalpha, TK, TS, delta, xe);
      const double Tceff = o_Tc_eff(TK, TS);
      const double TSinv = (1.0 / Trad + xa / Tceff + xc / TK) / (1.0 + xa + xc);
      TS = 1.0 / TSinv;
    }
    *xa_out = (float)xa;
  } else {
    const double TSinv = (1.0 / Trad + xc / TK) / (1.0 + xc);
    TS = 1.0 / TSinv;
    *xa_out = 0;
  }
  if (TS < 0.) TS = fabs(TS);
  return (float)TS;
}

static double o_dT_comp(const mrx_ts_params* p, const orc_ts_tab* d, double z, double TK, double xe)
{
  const double Trad = O_TCMB * (1.0 + z);
  double ans = (-1.51e-4) * (xe / (1.0 + xe + d->f_He)) / (o_hubble(p, (float)z) / (O_HUBBLE * p->hubble_h)) /
               p->hubble_h * pow(Trad, 4.0) / (1.0 + z);
  ans *= Trad - TK;
  return ans;
}

/* evolveInt without minihalos; deriv[0..4] as the reference fills them */
static void o_evolve(const mrx_ts_params* p, const orc_ts_tab* s, float zp, float curr_delNL0, const double* SFR,
                     const double* fh, const double* fi, const double* fl, const double* y, double* deriv)
{
  const double Conversion_factor = (O_SPEED_OF_LIGHT / (4. * M_PI)) / (O_PROTONMASS / O_SOLAR_MASS);
  const double x_e = y[0], T = y[1];
  const double n_b = s->N_b0 * pow(1 + zp, 3) * (1 + curr_delNL0);
  double dxheat = 0, dxion = 0, dxlya = 0, dstarlya = 0;
  if (!s->no_light) {
    for (int ct = 0; ct < s->nsteps; ct++) {
      const double dt_dzpp = s->dt_dzpp[ct], dzpp = s->dzpp[ct];
      const double zpp_integrand = SFR[ct] * s->zpp_pow[ct];
      dxheat += dt_dzpp * dzpp * zpp_integrand * fh[ct];
      dxion += dt_dzpp * dzpp * zpp_integrand * fi[ct];
      dxlya += dt_dzpp * dzpp * zpp_integrand * fl[ct];
      dstarlya += SFR[ct] * pow(1 + zp, 2) * (1 + s->zpp[ct]) * s->sum_lyn[ct] * dt_dzpp * dzpp;
    }
    dxheat *= s->const_zp_prefactor;
    dxion *= s->const_zp_prefactor;
    dxlya *= s->const_zp_prefactor * n_b;
    dstarlya *= Conversion_factor;
  }
  const double sink = o_alpha_A(T) * O_CLUMPING * x_e * x_e * s->f_H * n_b;
  const double dxe_dzp = s->dt_dzp * (dxion - sink);
  deriv[0] = dxe_dzp;
  double dadia = 3 / (1.0 + zp);
  if (fabs(curr_delNL0) > O_FRACT_FLOAT_ERR) dadia += s->dgrowth_factor_dzp / (1.0 / curr_delNL0 + s->growth_factor_zp);
  dadia *= (2.0 / 3.0) * T;
  const double dspec = -dxe_dzp * T / (1 + x_e);
  const double dcomp = o_dT_comp(p, s, zp, T, x_e);
  const double dxheat_dzp = dxheat * s->dt_dzp * 2.0 / 3.0 / O_BOLTZMANN / (1.0 + x_e);
  deriv[1] = dxheat_dzp + dcomp + dspec + dadia;
  deriv[2] = dxlya + dstarlya;
  deriv[3] = dxheat_dzp;
  deriv[4] = s->dt_dzp * dxion;
}

/* ------------------------------------------------------------ driver */

static double o_RtoM(const mrx_ts_params* p, double R)
{
  if (p->rtom_filter_type == 0) return (4.0 / 3.0) * M_PI * pow(R, 3) * (p->omega_m * p->rho_crit);
  return pow(2 * M_PI, 1.5) * p->omega_m * p->rho_crit * pow(R, 3);
}


/* collapse fraction in one cell at radius R (ComputeTs.c:236-242, :447-451) */
static double o_cell_fcoll(const mrx_ts_params* p, double R, double pixel_volume, float stars, float deltax)
{
  const double h = p->hubble_h;
  const double dom = 1.0 + deltax;
  double c = stars / (o_RtoM(p, R) * h * h * dom) * (4.0 / 3.0) * M_PI * pow(R, 3.0) / pixel_volume;
  return c > 1.0 ? 1.0 : c;
}

int orc_ts_first_shell_means(const mrx_ts_params* p, double R, const float* deltax, const float* stars,
                             const float* x_e_box_prev, double* collapse_fraction, double* x_e_ave)
{
  const int N = p->grid_dim;
  if (N < 2 || (N & 1) || !deltax || !stars || !x_e_box_prev) return MRX_EINVAL;
  const int PR = 2 * (N / 2 + 1);
  const double pixel_volume = pow(p->box_size / p->hubble_h / (double)N, 3);
  const double total_n_cells = pow((double)N, 3);
  double cf = 0.0, xe = 0.0;
  for (long r = 0; r < (long)N * N; r++)
    for (int z = 0; z < N; z++) {
      const long ip = r * PR + z;
      cf += o_cell_fcoll(p, R, pixel_volume, stars[ip], deltax[ip]);
      xe += x_e_box_prev[ip];
    }
  if (collapse_fraction) *collapse_fraction = cf / total_n_cells;
  if (x_e_ave) *x_e_ave = xe / total_n_cells;
  return MRX_OK;
}

int orc_compute_ts(const mrx_ts_params* p, const orc_ts_tab* t, double collapse_fraction, double x_e_ave,
                   const float* deltax, const float* stars, const float* sfr, const float* hist, float* x_e_box,
                   float* x_e_box_prev, float* Tk_box, float* TS_box, mrx_ts_globals* out, float* shells_out)
{
  const int N = p->grid_dim;
  if (N < 2 || (N & 1) || !t || !deltax || !stars || !sfr || !x_e_box || !x_e_box_prev || !Tk_box || !TS_box)
    return MRX_EINVAL;
  if (t->nsteps < 1 || t->nsteps > ORC_TS_MAX_STEPS || t->n_ops < 0 || t->n_ops > ORC_TS_MAX_OPS) return MRX_EINVAL;
  for (int k = 0; k < t->n_ops; k++)
    if (t->op_kind[k] >= 2 && (!hist || t->op_slot[k] < 0 || t->op_slot[k] >= p->n_history)) return MRX_EINVAL;
  const int PR = 2 * (N / 2 + 1);
  const long nreal = (long)N * N * N;
  const size_t npad = (size_t)N * N * PR;
  const size_t half = (size_t)N * N * (N / 2 + 1); /* slab_n_complex floats */
  const double box_size = p->box_size / p->hubble_h; /* Mpc */
  const double pixel_volume = pow(box_size / (double)N, 3);
  const double total_n_cells = pow((double)N, 3);
  const float zp = (float)t->zp;
  const double unit_fac = (p->unit_mass_g / p->unit_time_s);
  const double len_fac = pow(p->unit_length_cm, -3.);
  double J_alpha_ave = 0, xalpha_ave = 0, Xheat_ave = 0, Xion_ave = 0;

  if (t->homogeneous) {
    const float xe0 = (float)t->xe_recfast;
    const float T0 = (float)t->T_recfast;
    for (long r = 0; r < (long)N * N; r++)
      for (int z = 0; z < N; z++) {
        const long ir = r * N + z, ip = r * PR + z;
        float xa;
        x_e_box_prev[ip] = xe0;
        Tk_box[ir] = T0;
        TS_box[ir] = o_get_Ts(p, t, zp, deltax[ip], Tk_box[ir], x_e_box_prev[ip], 0, &xa);
      }
  } else {
    float* work = (float*)malloc(sizeof(float) * npad);
    float* filt = (float*)malloc(sizeof(float) * npad);
    float* spec = (float*)malloc(sizeof(float) * npad);
    float* shells = (float*)malloc(sizeof(float) * (size_t)t->nsteps * nreal);
    if (!work || !filt || !spec || !shells) {
      free(work);
      free(filt);
      free(spec);
      free(shells);
      return MRX_ENOMEM;
    }
    memcpy(work, sfr, sizeof(float) * npad);
    int k = 0;
    for (int ct = 0; ct < t->nsteps; ct++) {
      for (; k < t->n_ops && t->op_shell[k] == ct; k++) {
        const int kind = t->op_kind[k];
        if (kind == 0) {
          for (size_t i = 0; i < half; i++) work[i] = 0;
        } else if (kind == 1) {
          for (size_t i = 0; i < half; i++) work[i] *= t->op_weight[k];
        } else {
          /* load_reion_sfr_grids (reionization.c:1756-1784): float weight,
           * slot read with INDEX_REAL, written with INDEX_PADDED */
          const float w = (float)t->op_weight[k];
          const float* hs = hist + (size_t)t->op_slot[k] * npad;
          for (long r = 0; r < (long)N * N; r++)
            for (int z = 0; z < N; z++) {
              const float v = hs[r * N + z] * w;
              if (kind == 2)
                work[r * PR + z] = v;
              else
                work[r * PR + z] += v;
            }
        }
      }
      orc_r2c3d_f32(N, work, spec);
      const float fN3 = (float)total_n_cells;
      for (size_t i = 0; i < npad; i++) spec[i] /= fN3;
      if (ct > 0) orc_filter_f32(spec, 0, N, N, (float)t->R[ct], p->heating_filter_type, p->box_size);
      orc_c2r3d_f32(N, spec, filt);
      for (long r = 0; r < (long)N * N; r++)
        for (int z = 0; z < N; z++) shells[(size_t)ct * nreal + r * N + z] = fmaxf(filt[r * PR + z], 0.0);
    }

    const float* XH = t->xHII;
    for (long r = 0; r < (long)N * N; r++)
      for (int z = 0; z < N; z++) {
        const long ir = r * N + z, ip = r * PR + z;
        double SFR[ORC_TS_MAX_STEPS], fh[ORC_TS_MAX_STEPS], fi[ORC_TS_MAX_STEPS], fl[ORC_TS_MAX_STEPS];
        double ans[2], der[5];
        ans[0] = x_e_box_prev[ip];
        ans[1] = Tk_box[ir];
        for (int ct = 0; ct < t->nsteps; ct++) {
          SFR[ct] = ((double)shells[(size_t)ct * nreal + ir] / pixel_volume) * unit_fac * len_fac / O_SOLAR_MASS;
          double xHII_call = x_e_box_prev[ip];
          if (xHII_call > XH[ORC_XHII - 1] * 0.999)
            xHII_call = XH[ORC_XHII - 1] * 0.999;
          else if (xHII_call < XH[0])
            xHII_call = 1.001 * XH[0];
          const int lo = o_locate_xHII(t, (float)xHII_call), hi = lo + 1;
          fh[ct] = (t->freq_heat[hi][ct] - t->freq_heat[lo][ct]) / (XH[hi] - XH[lo]);
          fh[ct] *= (xHII_call - XH[lo]);
          fh[ct] += t->freq_heat[lo][ct];
          fi[ct] = (t->freq_ion[hi][ct] - t->freq_ion[lo][ct]) / (XH[hi] - XH[lo]);
          fi[ct] *= (xHII_call - XH[lo]);
          fi[ct] += t->freq_ion[lo][ct];
          fl[ct] = (t->freq_lya[hi][ct] - t->freq_lya[lo][ct]) / (XH[hi] - XH[lo]);
          fl[ct] *= (xHII_call - XH[lo]);
          fl[ct] += t->freq_lya[lo][ct];
        }
        o_evolve(p, t, zp, deltax[ip], SFR, fh, fi, fl, ans, der);
        x_e_box_prev[ip] += der[0] * t->dzp;
        if (x_e_box_prev[ip] > 1)
          x_e_box_prev[ip] = (float)(1 - O_FRACT_FLOAT_ERR);
        else if (x_e_box_prev[ip] < 0)
          x_e_box_prev[ip] = 0;
        if (Tk_box[ir] < O_MAX_TK) Tk_box[ir] += der[1] * t->dzp;
        if (Tk_box[ir] < 0) Tk_box[ir] = (float)(O_TCMB * (1 + t->zp)); /* zp is double here */
        float xa;
        TS_box[ir] = o_get_Ts(p, t, zp, deltax[ip], Tk_box[ir], x_e_box_prev[ip], (float)der[2], &xa);
        J_alpha_ave += der[2];
        xalpha_ave += xa;
        Xheat_ave += der[3];
        Xion_ave += der[4];
      }
    if (shells_out) memcpy(shells_out, shells, sizeof(float) * (size_t)t->nsteps * nreal);
    free(work);
    free(filt);
    free(spec);
    free(shells);
    J_alpha_ave /= total_n_cells;
    xalpha_ave /= total_n_cells;
    Xheat_ave /= total_n_cells;
    Xion_ave /= total_n_cells;
  }
  memcpy(x_e_box, x_e_box_prev, sizeof(float) * npad);
  double Ave_Ts = 0, Ave_Tk = 0, Ave_xe = 0;
  for (long r = 0; r < (long)N * N; r++)
    for (int z = 0; z < N; z++) {
      const long ir = r * N + z, ip = r * PR + z;
      Ave_Ts += (double)TS_box[ir];
      Ave_Tk += (double)Tk_box[ir];
      Ave_xe += (double)x_e_box_prev[ip];
    }
  if (p->reproduce_quirks) {
    /* ComputeTs.c:957 sums x_e_box_prev at the stale index of the previous loop: the last cell */
    const long ip_last = ((long)N * N - 1) * PR + (N - 1);
    Ave_xe = (double)x_e_box_prev[ip_last] * total_n_cells;
  }
  if (out) {
    out->volume_ave_J_alpha = J_alpha_ave;
    out->volume_ave_xalpha = xalpha_ave;
    out->volume_ave_Xheat = Xheat_ave;
    out->volume_ave_Xion = Xion_ave;
    out->volume_ave_TS = Ave_Ts / total_n_cells;
    out->volume_ave_TK = Ave_Tk / total_n_cells;
    out->volume_ave_xe = Ave_xe / total_n_cells;
    out->collapse_fraction = collapse_fraction;
    out->x_e_ave = x_e_ave;
    out->no_light = t->homogeneous ? 0 : t->no_light;
    out->homogeneous = t->homogeneous;
  }
  return MRX_OK;
}
