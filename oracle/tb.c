/*
 * tb.c - the oracle's restatement of ComputeBrightnessTemperatureBox for
 * Flag_IncludePecVelsFor21cm = 0 (BrightnessTemperature.c:24-124, :512-545).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Build with -ffp-contract=off.
 */
#include <math.h>

#include "oracle.h"

#define ORC_TCMB 2.728

int orc_brightness_temp(const mrx_tb_params* p, const float* xH, const float* deltax, const float* TS_box,
                        float* delta_T, double* ave_Tb)
{
  const int N = p->grid_dim;
  if (N < 2 || (N & 1) || !xH || !deltax || !delta_T) return MRX_EINVAL;
  if (p->include_spin_temp && !TS_box) return MRX_EINVAL;
  const int PR = 2 * (N / 2 + 1);
  const double total_n_cells = pow((double)N, 3);

  /* :46-52 - every factor is rounded to float where the reference declares float */
  float redshift = (float)(float)p->redshift;
  float Hubble_h = (float)(float)p->hubble_h;
  float OmegaM = (float)(float)p->omega_m;
  float OmegaB = (float)(OmegaM * p->baryon_frac);
  float const_factor = (float)(27.0 * (OmegaB * Hubble_h * Hubble_h / 0.023) *
                               sqrt((0.15 / OmegaM / Hubble_h / Hubble_h) * (1.0 + redshift) / 10.0));
  double T_rad = ORC_TCMB * (1. + redshift);

#pragma omp parallel for schedule(static)
  for (long r = 0; r < (long)N * N; r++)
    for (int z = 0; z < N; z++) {
      long i_real = r * N + z, i_padded = r * PR + z;
      double pixel_deltax = deltax[i_padded];
      double pixel_x_HI = xH[i_real];
      delta_T[i_real] = (float)(const_factor * pixel_x_HI * (1.0 + pixel_deltax));
      if (p->include_spin_temp) {
        double pixel_Ts_factor = (1 - T_rad / TS_box[i_real]);
        delta_T[i_real] *= pixel_Ts_factor;
      }
    }

  /* :512-535: sequential double sum, divided by N^3 */
  double Ave_Tb = 0.0;
  for (long i = 0; i < (long)N * N * N; i++) Ave_Tb += (double)delta_T[i];
  Ave_Tb /= total_n_cells;
  if (ave_Tb) *ave_Tb = Ave_Tb;
  return MRX_OK;
}

/* ------------------------------------------------------------------------
 * Peculiar-velocity branch (BrightnessTemperature.c:126-510) for
 * include_spin_temp = 1 and pec_vels > 0; otherwise orc_brightness_temp.
 */
#include <stdlib.h>

#define ORC_N_RSD 50          /* N_RSD_STEPS, reionization.h:34 */
#define ORC_MAX_DVDR 0.2f     /* MAX_DVDR, reionization.h:39 */
#define ORC_FRACT 1e-7        /* FRACT_FLOAT_ERR, reionization.h:31 */
#define ORC_MPC 3.086e24
#define ORC_HUBBLE 3.2407789e-18

/* hubble(z) of XRayHeatingFunctions.c:394-404 (z is a float there) */
static double orc_hubble_z(const mrx_tb_vel_params* p, float z)
{
  return ORC_HUBBLE * p->tb.hubble_h * sqrt(p->tb.omega_m * pow(1 + z, 3) + p->omega_r * pow(1 + z, 4) +
                                             p->omega_lambda);
}

/* Adds `v` (a double share) to a float bin the way `bin += double` rounds. */
static inline void add_d(float* bin, double v) { *bin = (float)((double)*bin + v); }

/* One line of sight (fixed x, y) of the sub-cell re-binning (:312-496).  dt,
 * xh: the column's delta_T and xH; vel: the column's converted velocities;
 * out: the re-binned column.  Indices that the reference's single +-N
 * correction would leave outside the column are wrapped periodically. */
static void rsd_column(int N, const float* dt, const float* xh, const float* vel, float H_z, const float* x_pos,
                       const float* x_off, double cell, double sub_w, float* out)
{
  for (int k = 0; k < N; k++) out[k] = 0.0f;
  const double half = (sub_w / 2.) / cell, frac_w = sub_w / cell;
  for (int k = 0; k < N; k++) {
    if (!(fabs(dt[k]) >= ORC_FRACT && xh[k] >= ORC_FRACT)) continue;
    const int km = (k == 0) ? N - 1 : k - 1, kp = (k == N - 1) ? 0 : k + 1;
    const double d1_low = vel[km] / H_z, d2_low = vel[k] / H_z;   /* float quotients */
    const double d1_high = vel[k] / H_z, d2_high = vel[kp] / H_z;
    const float share = dt[k] / (float)ORC_N_RSD;
#define AT(i) (out + ((((i) % N) + N) % N))
    for (int s = 0; s < ORC_N_RSD; s++) {
      const double disp = (x_pos[s] <= 0.5) ? d1_low + ((x_pos[s] + 0.5) - 0.) * (d2_low - d1_low) / (1. - 0.)
                                            : d1_high + ((x_pos[s] - 0.5) - 0.) * (d2_high - d1_high) / (1. - 0.);
      const double pos = (x_off[s] + disp) / cell;
      const double lo = pos - half, hi = pos + half;
      if (lo >= 0.0 && hi < 1.0) {
        *AT(k) += share; /* stays */
      } else if (lo < 0.0 && hi < 0.0) { /* wholly towards the observer */
        const double cd = ceil(fabs(lo)) - 1.;
        const int c = (int)cd;
        if (fabs(hi) > cd) {
          *AT(k < c + 1 ? k - (c + 1) + N : k - (c + 1)) += share;
        } else {
          const double fo = (fabs(lo) - cd) / frac_w, fw = 1. - fo;
          add_d(AT(k < c ? k - c + N : k - c), fw * dt[k] / (float)ORC_N_RSD);
          add_d(AT(k < c + 1 ? k - (c + 1) + N : k - (c + 1)), fo * dt[k] / (float)ORC_N_RSD);
        }
      } else if (lo < 0.0 && (hi > 0.0 && hi < 1.0)) { /* partly into the cell before */
        const double fw = hi / frac_w, fo = 1. - fw;
        add_d(AT(k == 0 ? N - 1 : k - 1), fo * dt[k] / (float)ORC_N_RSD);
        add_d(AT(k), fw * dt[k] / (float)ORC_N_RSD);
      } else if ((lo >= 0.0 && lo < 1.0) && hi >= 1.0) { /* partly into the cell after */
        const double fo = (hi - 1.) / frac_w, fw = 1. - fo;
        add_d(AT(k), fw * dt[k] / (float)ORC_N_RSD);
        add_d(AT(k == N - 1 ? 0 : k + 1), fo * dt[k] / (float)ORC_N_RSD);
      } else { /* wholly away from the observer */
        const double cd = floor(fabs(hi));
        const int c = (int)cd;
        if (lo >= cd) {
          *AT(k > N - 1 - c ? k + c - N : k + c) += share;
        } else {
          const double fo = (hi - cd) / frac_w, fw = 1. - fo;
          add_d(AT(k > N - 1 - (c - 1) ? k + c - 1 - N : k + c - 1), fw * dt[k] / (float)ORC_N_RSD);
          add_d(AT(k > N - 1 - c ? k + c - N : k + c), fo * dt[k] / (float)ORC_N_RSD);
        }
      }
    }
#undef AT
  }
}

/* The sub-cell re-binning alone (:312-496) on a delta_T grid, with vel the
 * converted (comoving) velocities; delta_T is re-binned in place. */
int orc_rsd_los(const mrx_tb_vel_params* p, const float* xH, const float* vel, float* delta_T)
{
  const int N = p->tb.grid_dim;
  const int PR = 2 * (N / 2 + 1);
  const float redshift = (float)(float)p->tb.redshift;
  const float H_z = (float)(float)orc_hubble_z(p, redshift);
  const double box_mpc = p->box_size / p->tb.hubble_h;
  const double sub_w = (box_mpc / ((double)N)) / (double)ORC_N_RSD;
  const double cell = box_mpc / (float)N;
  float x_off[ORC_N_RSD], x_pos[ORC_N_RSD];
  for (int s = 0; s < ORC_N_RSD; s++) {
    x_off[s] = (float)(sub_w * (float)s + sub_w / 2.);
    x_pos[s] = (float)(x_off[s] / (box_mpc / (float)N));
  }
#pragma omp parallel
  {
    float* col = (float*)malloc(sizeof(float) * (size_t)N);
    float* vrow = (float*)malloc(sizeof(float) * (size_t)N);
#pragma omp for schedule(static)
    for (long r = 0; r < (long)N * N; r++) {
      for (int z = 0; z < N; z++) vrow[z] = vel[r * PR + z];
      rsd_column(N, delta_T + r * N, xH + r * N, vrow, H_z, x_pos, x_off, cell, sub_w, col);
      for (int z = 0; z < N; z++) delta_T[r * N + z] = col[z];
    }
    free(col);
    free(vrow);
  }
  return MRX_OK;
}

int orc_brightness_temp_vel(const mrx_tb_vel_params* p, const float* xH, const float* deltax, const float* TS_box,
                            float* vel, float* delta_T, double* ave_Tb)
{
  const int N = p->tb.grid_dim;
  if (!p->tb.include_spin_temp || p->pec_vels <= 0)
    return orc_brightness_temp(&p->tb, xH, deltax, TS_box, delta_T, ave_Tb);
  if (N < 2 || (N & 1) || !xH || !deltax || !delta_T || !TS_box || !vel || p->pec_vels > 3) return MRX_EINVAL;
  const int PR = 2 * (N / 2 + 1);
  const long NC = (long)N * N * (N / 2 + 1);
  const double total_n_cells = pow((double)N, 3);
  const float redshift = (float)(float)p->tb.redshift;
  const float Hubble_h = (float)(float)p->tb.hubble_h;
  const float OmegaM = (float)(float)p->tb.omega_m;
  const float OmegaB = (float)(OmegaM * p->tb.baryon_frac);
  const float const_factor = (float)(27.0 * (OmegaB * Hubble_h * Hubble_h / 0.023) *
                                     sqrt((0.15 / OmegaM / Hubble_h / Hubble_h) * (1.0 + redshift) / 10.0));
  const float H_z = (float)(float)orc_hubble_z(p, redshift);
  const double T_rad = ORC_TCMB * (1. + redshift);

  /* the cells before velocities (:86-120) */
  for (long r = 0; r < (long)N * N; r++)
    for (int z = 0; z < N; z++) {
      const long ir = r * N + z, ip = r * PR + z;
      const double pixel_deltax = deltax[ip], pixel_x_HI = xH[ir];
      delta_T[ir] = (float)(const_factor * pixel_x_HI * (1.0 + pixel_deltax));
      if (p->pec_vels > 1)
        delta_T[ir] *= (1. + redshift) / (1000. * TS_box[ir]);
      else
        delta_T[ir] *= (1 - T_rad / TS_box[ir]);
    }

  /* comoving velocity (:166-195) */
  for (long r = 0; r < (long)N * N; r++)
    for (int z = 0; z < N; z++) {
      float* v = vel + r * PR + z;
      if (p->trees_id == 1)
        *v = (float)(sqrt(1. / (1. + redshift)) * *v / 1000.);
      else if (p->trees_id == 0 || p->trees_id == 2)
        *v = (float)((1. + redshift) * *v);
      *v *= 1000. * 100. / ORC_MPC;
      *v = (float)(*v / (1. / (1. + redshift)));
    }

  /* dv/dz (:197-208, velocity_gradient reionization.c:2172-2187) */
  float* herm = (float*)malloc(sizeof(float) * 2 * NC);
  float* grad = (float*)malloc(sizeof(float) * (size_t)N * N * PR);
  if (!herm || !grad) {
    free(herm);
    free(grad);
    return MRX_ENOMEM;
  }
  orc_r2c3d_f32(N, vel, herm);
  const float box_f = (float)p->box_size;
  const float delta_k = (float)(2.0 * M_PI / box_f);
  for (long i = 0; i < NC; i++) {
    float re = (float)((double)herm[2 * i] / total_n_cells), im = (float)((double)herm[2 * i + 1] / total_n_cells);
    re = (float)((double)re * p->tb.hubble_h);
    im = (float)((double)im * p->tb.hubble_h);
    const float k_z = (float)(i % (N / 2 + 1)) * delta_k;
    herm[2 * i] = re * 0.0f - im * k_z;
    herm[2 * i + 1] = re * k_z + im * 0.0f;
  }
  orc_c2r3d_f32(N, herm, grad);
  free(herm);

  if (p->pec_vels == 1) { /* capped gradient (:210-236) */
    const double max_v_deriv = fabs(ORC_MAX_DVDR * H_z);
    for (long r = 0; r < (long)N * N; r++)
      for (int z = 0; z < N; z++) {
        double dvdx = (double)grad[r * PR + z];
        if (fabs(dvdx) > max_v_deriv) dvdx = (dvdx < 0) ? -max_v_deriv : max_v_deriv;
        delta_T[r * N + z] /= (dvdx / H_z + 1.0);
      }
  } else { /* optical depth (:240-287) */
    for (long r = 0; r < (long)N * N; r++)
      for (int z = 0; z < N; z++) {
        const long ir = r * N + z;
        const double gc = fabs((double)grad[r * PR + z] / H_z + 1.0);
        if (gc < ORC_FRACT)
          delta_T[ir] = (float)(1000. * (TS_box[ir] - T_rad) / (1. + redshift));
        else
          delta_T[ir] = (float)((1. - exp(-delta_T[ir] / gc)) * 1000. * (TS_box[ir] - T_rad) / (1. + redshift));
      }
    if (p->pec_vels == 3) orc_rsd_los(p, xH, vel, delta_T); /* line-of-sight re-binning (:312-496) */
  }
  free(grad);
  double Ave_Tb = 0.0;
  for (long i = 0; i < (long)N * N * N; i++) Ave_Tb += (double)delta_T[i];
  Ave_Tb /= total_n_cells;
  if (ave_Tb) *ave_Tb = Ave_Tb;
  return MRX_OK;
}
