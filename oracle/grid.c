/*
 * grid.c - construct_baryon_grids (src/core/reionization.c:1353-1615) restated
 * for one rank holding the whole box.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The reference's loop order is kept: zero the grids (:1376-1398, with the SFR
 * history ring moved by one slot), then per property a float buffer in
 * INDEX_REAL order accumulated galaxy by galaxy with `buffer[ind] += v`
 * (v double, :1436-1515), then the final pass into the padded grids
 * (:1530-1600).  pos_to_ngp is misc_tools.c:163-173.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static int pos_to_ngp(double x, double side, int nx, int* bad)
{
  int ind = (int)nearbyint(x / side * (double)nx);
  if (ind > nx - 1) ind = 0;
  if (!(ind > -1)) *bad = 1; /* the reference asserts here */
  return ind;
}

int orc_construct_baryon_grids(const mrx_grid_params* p, long n_gal, const float* pos, const int* type,
                               const double* fesc, const double* gross, const double* bhm, const double* ebhm,
                               float* stars, float* wsfr, float* sfr, float* hist, long* stats)
{
  const int N = p->grid_dim;
  const long np = 2L * N * N * (N / 2 + 1), nr = (long)N * N * N;
  const int spin = p->include_spin_temp, bh = p->bh_feedback;
  /* virial_properties.c:79-93 */
  const double zplus1 = p->redshift + 1;
  const double hubble_at_snapshot = p->hubble * sqrt(p->omega_m * zplus1 * zplus1 * zplus1 +
                                                     p->omega_k * zplus1 * zplus1 + p->omega_lambda);
  const double sfr_timescale = p->reion_sfr_timescale * (1.0 / hubble_at_snapshot);
  /* validity first: the reference asserts inside its loop */
  for (long g = 0; g < n_gal; g++) {
    if (type[g] > 2) continue;
    int bad = 0;
    for (int d = 0; d < 3; d++) pos_to_ngp(pos[3 * g + d], p->box_size, N, &bad);
    if (bad) return (int)(-100 - (g < 1000000 ? g : 1000000));
  }
  for (long ii = 0; ii < np; ii++) {
    stars[ii] = 0.0;
    wsfr[ii] = 0.0;
  }
  if (spin) {
    for (long ii = 0; ii < np; ii++) {
      sfr[ii] = 0.0;
      for (int snap = p->n_store_sfr - 2; snap >= 0; snap--) hist[(snap + 1) * np + ii] = hist[snap * np + ii];
    }
  }
  float* buffer = (float*)malloc(sizeof(float) * nr);
  if (!buffer) return -2;
  long n_dead = 0, n_below = 0;
  for (int prop = 0; prop < 3; prop++) { /* stellar, weighted_sfr, sfr */
    if (!spin && prop == 2) continue;
    for (long ii = 0; ii < nr; ii++) buffer[ii] = (float)0.;
    for (long g = 0; g < n_gal; g++) {
      if (type[g] > 2) {
        if (prop == 0) n_dead++;
        continue;
      }
      int bad = 0;
      const int ix = pos_to_ngp(pos[3 * g], p->box_size, N, &bad);
      const int iy = pos_to_ngp(pos[3 * g + 1], p->box_size, N, &bad);
      const int iz = pos_to_ngp(pos[3 * g + 2], p->box_size, N, &bad);
      const long ind = iz + (long)N * (iy + (long)N * ix);
      switch (prop) {
        case 0:
          buffer[ind] += fesc[g];
          if (bh) {
            if (bhm[g] >= p->bh_mass_limit_reion)
              buffer[ind] += ebhm[g];
            else
              n_below += 1;
          }
          break;
        case 1:
          buffer[ind] += fesc[g];
          if (bh)
            if (bhm[g] >= p->bh_mass_limit_reion) buffer[ind] += ebhm[g] * p->alpha_uv_bh / p->alpha_uv;
          break;
        default:
          buffer[ind] += gross[g];
          break;
      }
    }
    for (int ix = 0; ix < N; ix++)
      for (int iy = 0; iy < N; iy++)
        for (int iz = 0; iz < N; iz++) {
          const long ir = iz + (long)N * (iy + (long)N * ix);
          const long ip = iz + 2L * (N / 2 + 1) * (iy + (long)N * ix);
          if (prop == 0) {
            float val = buffer[ir];
            if (val < 0) val = 0;
            stars[ip] = val;
          } else {
            double val = (double)buffer[ir];
            val = (val > 0) ? val / sfr_timescale : 0;
            if (prop == 1) {
              wsfr[ip] = (float)val;
            } else {
              sfr[ip] = (float)val;
              hist[ip] = (float)val;
            }
          }
        }
  }
  free(buffer);
  if (stats) {
    stats[0] = n_gal - n_dead;
    stats[1] = n_dead;
    stats[2] = n_below;
  }
  return 0;
}
