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
