/*
 * hii.c - the oracle's restatement of filter() and _find_HII_bubbles.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows, operation by operation where the result depends on it:
 *   filter()             reionization.c:2110-2170
 *   RtoM                 find_HII_bubbles.c:26-45
 *   _find_HII_bubbles    find_HII_bubbles.c:47-495 (recombinations and
 *                        minihalos off, as in every configuration measured)
 * Build with -ffp-contract=off: the reference's default build (RelWithDebInfo,
 * baseline x86-64) never fuses multiply-adds, and neither may the restatement.
 */
#include <math.h>
#include <stdbool.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* constants of src/meraxes.h:21-44 and reionization.h:38 */
#define ORC_PLANCK 6.6262e-27
#define ORC_PROTONMASS 1.6726e-24
#define ORC_SIGMA_HI 6.3e-18 /* SIGMA_HI, reionization.h:11 */
#define ORC_REL_TOL ((float)1e-5)
#define ORC_ABS_TOL ((float)1e-8)
#define ORC_L_FACTOR 0.620350491

/* --------------------------------------------------------------- windows */

/* k components exactly as filter() builds them (reionization.c:2112-2135) */
static inline float orc_kcomp(int n, int N, float delta_k)
{
  int middle = N / 2;
  if (n > middle) return (float)(n - N) * delta_k;
  return (float)n * delta_k;
}

/* returns the multiplier filter() applies; *touched = 0 when it leaves the mode alone */
static inline float orc_window_value(float kx, float ky, float kz, float R, int filter_type, int* touched)
{
  float k_mag = sqrtf(kx * kx + ky * ky + kz * kz);
  float kR = k_mag * R;
  *touched = 1;
  switch (filter_type) {
    case 0:
      if (kR > 1e-4) return (float)(3.0 * (sinf(kR) / powf(kR, 3) - cosf(kR) / powf(kR, 2)));
      *touched = 0;
      return 1.0f;
    case 1:
      kR *= 0.413566994;
      if (kR > 1) return 0.0f;
      *touched = 0;
      return 1.0f;
    case 2:
      kR *= 0.643;
      return powf((float)M_E, (float)(-kR * kR / 2.0));
    default:
      *touched = 0;
      return 1.0f;
  }
}

float orc_window_f32(int nx, int ny, int nz, int N, float R, int filter_type, double box_size)
{
  float bs = (float)box_size;
  float delta_k = (float)(2.0 * M_PI / bs);
  int t;
  return orc_window_value(orc_kcomp(nx, N, delta_k), orc_kcomp(ny, N, delta_k), (float)nz * delta_k, R,
                          filter_type, &t);
}

int orc_filter_f32(float* box, int local_ix_start, int slab_nx, int N, float R, int filter_type, double box_size)
{
  if (filter_type < 0 || filter_type > 2) return MRX_EFILTER;
  const int middle = N / 2;
  const float bs = (float)box_size;
  const float delta_k = (float)(2.0 * M_PI / bs);
#pragma omp parallel for schedule(static)
  for (int n_x = 0; n_x < slab_nx; n_x++) {
    float k_x = orc_kcomp(n_x + local_ix_start, N, delta_k);
    for (int n_y = 0; n_y < N; n_y++) {
      float k_y = orc_kcomp(n_y, N, delta_k);
      for (int n_z = 0; n_z <= middle; n_z++) {
        float k_z = (float)n_z * delta_k;
        int touched;
        float w = orc_window_value(k_x, k_y, k_z, R, filter_type, &touched);
        if (!touched) continue;
        long i = 2 * ((long)n_z + (long)(middle + 1) * ((long)n_y + (long)N * n_x));
        box[i] *= w;
        box[i + 1] *= w;
      }
    }
  }
  return MRX_OK;
}

/* FP64 window.  The sharp-k mode set is the reference's own float decision, so
 * the two precisions filter the same modes; the window values are double. */
int orc_filter_f64(double* box, int N, double R, int filter_type, double box_size)
{
  if (filter_type < 0 || filter_type > 2) return MRX_EFILTER;
  const int middle = N / 2;
  const double delta_k = 2.0 * M_PI / box_size;
  const float bs = (float)box_size;
  const float delta_kf = (float)(2.0 * M_PI / bs);
#pragma omp parallel for schedule(static)
  for (int n_x = 0; n_x < N; n_x++) {
    double k_x = (n_x > middle ? n_x - N : n_x) * delta_k;
    for (int n_y = 0; n_y < N; n_y++) {
      double k_y = (n_y > middle ? n_y - N : n_y) * delta_k;
      for (int n_z = 0; n_z <= middle; n_z++) {
        double k_z = n_z * delta_k;
        double kR = sqrt(k_x * k_x + k_y * k_y + k_z * k_z) * R;
        double w = 1.0;
        if (filter_type == 0) {
          if (kR > 1e-4) w = 3.0 * (sin(kR) - kR * cos(kR)) / (kR * kR * kR);
        } else if (filter_type == 1) {
          int touched;
          orc_window_value(orc_kcomp(n_x, N, delta_kf), orc_kcomp(n_y, N, delta_kf), (float)n_z * delta_kf,
                           (float)R, 1, &touched);
          if (touched) w = 0.0;
        } else {
          double x = 0.643 * kR;
          w = exp(-x * x / 2.0);
        }
        long i = 2 * ((long)n_z + (long)(middle + 1) * ((long)n_y + (long)N * n_x));
        box[i] *= w;
        box[i + 1] *= w;
      }
    }
  }
  return MRX_OK;
}

/* ---------------------------------------------------------------- ladder */

int orc_radius_ladder(const mrx_hii_params* p, double* R_out, int max_n)
{
  const double box_size = p->box_size;
  const int N = p->grid_dim;
  double clf = ORC_L_FACTOR;
  if ((box_size / (double)N) < 1.0) clf = 1.0;
  double R = fmin(p->r_bubble_max, ORC_L_FACTOR * box_size);
  int n = 0;
  bool last = false;
  while (!last) {
    if (((R / p->delta_r_factor) <= (clf * box_size / (double)N)) || ((R / p->delta_r_factor) <= p->r_bubble_min)) {
      last = true;
      R = clf * box_size / (double)N;
    }
    if (n < max_n && R_out) R_out[n] = R;
    n++;
    if (n > 100000) return MRX_EINVAL;
    R /= p->delta_r_factor;
  }
  return n;
}

/* ---------------------------------------------------------- clamp kinds */

static inline float orc_clamp_f32(float v, int kind)
{
  switch (kind) {
    case 0:
      return fmaxf(v, -1 + ORC_REL_TOL);
    case 1:
      v = fmaxf(v, 0.0);
      if (v < ORC_ABS_TOL) v = 0;
      return v;
    case 2:
      v = fmaxf(v, 0.0);
      if (v < ORC_ABS_TOL) v = 0;
      return fminf(v, 0.999);
    default:
      return v;
  }
}

static inline double orc_clamp_f64(double v, int kind)
{
  switch (kind) {
    case 0:
      return fmax(v, -1.0 + 1e-5);
    case 1:
      v = fmax(v, 0.0);
      if (v < 1e-8) v = 0;
      return v;
    case 2:
      v = fmax(v, 0.0);
      if (v < 1e-8) v = 0;
      return fmin(v, 0.999);
    default:
      return v;
  }
}

/* --------------------------------------------------- k-space preparation */

static size_t orc_nh(int N) { return (size_t)N * N * (N / 2 + 1); }

/* R2C then X /= N^3 the way find_HII_bubbles.c:168-182 does it: a float
 * complex divided by a double, i.e. in double, rounded back to float. */
static float* orc_forward_f32(int N, const float* field_padded)
{
  size_t nh = orc_nh(N);
  float* u = (float*)malloc(sizeof(float) * 2 * nh);
  orc_r2c3d_f32(N, field_padded, u);
  const double total_n_cells = pow((double)N, 3);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)(2 * nh); i++) u[i] = (float)((double)u[i] / total_n_cells);
  return u;
}

static double* orc_forward_f64(int N, const float* field_padded)
{
  size_t nh = orc_nh(N);
  double* in = (double*)malloc(sizeof(double) * 2 * nh);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)(2 * nh); i++) in[i] = (double)field_padded[i];
  double* u = (double*)malloc(sizeof(double) * 2 * nh);
  orc_r2c3d_f64(N, in, u);
  free(in);
  const double total_n_cells = pow((double)N, 3);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)(2 * nh); i++) u[i] /= total_n_cells;
  return u;
}

int orc_filtered_field_f32(int N, double box_size, int filter_type, const float* field_padded, double R,
                           int apply_filter, int clamp_kind, float* out_padded)
{
  if (N < 2 || (N & 1)) return MRX_EINVAL;
  if (filter_type < 0 || filter_type > 2) return MRX_EFILTER;
  float* u = orc_forward_f32(N, field_padded);
  if (apply_filter) orc_filter_f32(u, 0, N, N, (float)R, filter_type, box_size);
  orc_c2r3d_f32(N, u, out_padded);
  free(u);
  const int PR = 2 * (N / 2 + 1);
#pragma omp parallel for schedule(static)
  for (long r = 0; r < (long)N * N; r++)
    for (int z = 0; z < N; z++) out_padded[r * PR + z] = orc_clamp_f32(out_padded[r * PR + z], clamp_kind);
  return MRX_OK;
}

int orc_filtered_field_f64(int N, double box_size, int filter_type, const float* field_padded, double R,
                           int apply_filter, int clamp_kind, double* out_padded)
{
  if (N < 2 || (N & 1)) return MRX_EINVAL;
  if (filter_type < 0 || filter_type > 2) return MRX_EFILTER;
  double* u = orc_forward_f64(N, field_padded);
  if (apply_filter) orc_filter_f64(u, N, R, filter_type, box_size);
  orc_c2r3d_f64(N, u, out_padded);
  free(u);
  const int PR = 2 * (N / 2 + 1);
#pragma omp parallel for schedule(static)
  for (long r = 0; r < (long)N * N; r++)
    for (int z = 0; z < N; z++) out_padded[r * PR + z] = orc_clamp_f64(out_padded[r * PR + z], clamp_kind);
  return MRX_OK;
}

/* -------------------------------------------------------- bubble finder */

static double orc_RtoM(const mrx_hii_params* p, double R)
{
  switch (p->rtom_filter_type) {
    case 0:
      return (4.0 / 3.0) * M_PI * pow(R, 3) * (p->omega_m * p->rho_crit);
    case 1:
      return pow(2 * M_PI, 1.5) * p->omega_m * p->rho_crit * pow(R, 3);
    default:
      return -1;
  }
}

int orc_find_hii_bubbles(int precision, const mrx_hii_params* p, const float* deltax, const float* stars,
                         const float* weighted_sfr, const float* x_e, float* xH, float* r_bubble, float* J_21,
                         float* z_at_ionization, float* J_21_at_ionization, mrx_hii_globals* out, float* max_ratio)
{
  if (p->include_recombinations) return MRX_EINVAL;
  return orc_find_hii_bubbles_rec(precision, p, deltax, stars, weighted_sfr, x_e, NULL, xH, r_bubble, J_21,
                                  z_at_ionization, J_21_at_ionization, NULL, out, max_ratio);
}

int orc_find_hii_bubbles_rec(int precision, const mrx_hii_params* p, const float* deltax, const float* stars,
                             const float* weighted_sfr, const float* x_e, float* N_rec, float* xH, float* r_bubble,
                             float* J_21, float* z_at_ionization, float* J_21_at_ionization, float* Gamma12,
                             mrx_hii_globals* out, float* max_ratio)
{
  const int N = p->grid_dim;
  const int recomb = p->include_recombinations;
  if (recomb && (!N_rec || !Gamma12)) return MRX_EINVAL;
  if (N < 2 || (N & 1) || !deltax || !stars || !weighted_sfr || !xH || !r_bubble || !z_at_ionization)
    return MRX_EINVAL;
  if (p->filter_type < 0 || p->filter_type > 2) return MRX_EFILTER;
  if (p->rtom_filter_type < 0 || p->rtom_filter_type > 1) return MRX_EINVAL;
  if (p->uvb_flag && (!J_21 || !J_21_at_ionization)) return MRX_EINVAL;
  if (p->include_spin_temp && !x_e) return MRX_EINVAL;

  const double box_size = p->box_size;
  const double pixel_volume = pow(box_size / (double)N, 3);
  const double total_n_cells = pow((double)N, 3);
  const int uvb = p->uvb_flag;
  const double eff = p->efficiency;
  const double redshift = p->redshift;
  const long nreal = (long)N * N * N;
  const int PR = 2 * (N / 2 + 1);
  const size_t npad = (size_t)N * N * PR;

  if (uvb)
    for (long i = 0; i < nreal; i++) J_21[i] = 0.0f;
  for (long i = 0; i < nreal; i++) {
    xH[i] = 1.0f;
    r_bubble[i] = 0.0f;
  }
  if (max_ratio)
    for (long i = 0; i < nreal; i++) max_ratio[i] = 0.0f;

  /* fields: deltax, stars, weighted_sfr, then x_e (spin temperature), then N_rec (:155-163) */
  const int nf = 3 + (p->include_spin_temp ? 1 : 0) + (recomb ? 1 : 0);
  const int fx = 3, fr = p->include_spin_temp ? 4 : 3;
  const float* src[5] = {deltax, stars, weighted_sfr, p->include_spin_temp ? x_e : N_rec, N_rec};
  const int clamp_kind[5] = {0, 1, 1, p->include_spin_temp ? 2 : 1, 1};
  void* unf[5] = {0};
  void* fil[5] = {0};
  for (int f = 0; f < nf; f++) {
    if (precision == 64) {
      unf[f] = orc_forward_f64(N, src[f]);
      fil[f] = malloc(sizeof(double) * npad);
    } else {
      unf[f] = orc_forward_f32(N, src[f]);
      fil[f] = malloc(sizeof(float) * npad);
    }
  }

  double R_ladder[4096];
  int S = orc_radius_ladder(p, R_ladder, 4096);
  if (S < 1 || S > 4096) return MRX_EINVAL;

  for (int step = 0; step < S; step++) {
    const double R = R_ladder[step];
    const bool last = (step == S - 1);
    for (int f = 0; f < nf; f++) {
      if (precision == 64) {
        double* w = (double*)fil[f];
        memcpy(w, unf[f], sizeof(double) * npad);
        if (!last) orc_filter_f64(w, N, R, p->filter_type, box_size);
        orc_c2r3d_f64(N, w, w);
#pragma omp parallel for schedule(static)
        for (long r = 0; r < (long)N * N; r++)
          for (int z = 0; z < N; z++) w[r * PR + z] = orc_clamp_f64(w[r * PR + z], clamp_kind[f]);
      } else {
        float* w = (float*)fil[f];
        memcpy(w, unf[f], sizeof(float) * npad);
        if (!last) orc_filter_f32(w, 0, N, N, (float)R, p->filter_type, box_size);
        orc_c2r3d_f32(N, w, w);
#pragma omp parallel for schedule(static)
        for (long r = 0; r < (long)N * N; r++)
          for (int z = 0; z < N; z++) w[r * PR + z] = orc_clamp_f32(w[r * PR + z], clamp_kind[f]);
      }
    }

    /* find_HII_bubbles.c:317-320 */
    const double J_21_aux_constant = (1.0 + redshift) * (1.0 + redshift) / (4.0 * M_PI) * p->alpha_uv * ORC_PLANCK *
                                     1e21 * R * p->unit_length_cm * p->nion_phot_per_bary / ORC_PROTONMASS *
                                     p->unit_mass_g / pow(p->unit_length_cm, 3) / p->unit_time_s;
    const double M_mean = orc_RtoM(p, R);
    const double R_cubed = R * R * R;
    /* :325-334 */
    double Gamma_R_prefactor = 0.0;
    if (recomb) {
      Gamma_R_prefactor = (1.0 + redshift) * (1.0 + redshift) * R * (p->unit_length_cm / p->hubble_h) * ORC_SIGMA_HI *
                          p->alpha_uv / (p->alpha_uv + 2.75) / 1.0e-12;
      Gamma_R_prefactor *= (p->unit_mass_g / p->unit_time_s) * pow(p->unit_length_cm / p->hubble_h, -3.) *
                           p->nion_phot_per_bary / ORC_PROTONMASS;
    }

#pragma omp parallel for schedule(static)
    for (long r = 0; r < (long)N * N; r++) {
      for (int z = 0; z < N; z++) {
        const long i_real = r * N + z;
        const long i_padded = r * PR + z;
        double dv, sv, wv, xv = 0.0, nv = 0.0;
        if (precision == 64) {
          dv = ((double*)fil[0])[i_padded];
          sv = ((double*)fil[1])[i_padded];
          wv = ((double*)fil[2])[i_padded];
          if (p->include_spin_temp) xv = ((double*)fil[fx])[i_padded];
          if (recomb) nv = ((double*)fil[fr])[i_padded];
        } else {
          dv = (double)((float*)fil[0])[i_padded];
          sv = (double)((float*)fil[1])[i_padded];
          wv = (double)((float*)fil[2])[i_padded];
          if (p->include_spin_temp) xv = (double)((float*)fil[fx])[i_padded];
          if (recomb) nv = (double)((float*)fil[fr])[i_padded];
        }
        const double density_over_mean = 1.0 + dv;
        const double rec = recomb ? nv / density_over_mean : 0.0; /* :360-363 */
        const double f_coll_stars = sv / (M_mean * density_over_mean) * (4.0 / 3.0) * M_PI * R_cubed / pixel_volume;
        const double weighted_sfr_density = wv / pixel_volume;
        const double neutral_fraction = p->include_spin_temp ? 1.0 - xv : 1.0;
        float J_21_aux = 0;
        if (uvb) J_21_aux = (float)(weighted_sfr_density * J_21_aux_constant);

        if (max_ratio) {
          double ratio = f_coll_stars * eff / (neutral_fraction * (1. + rec));
          if ((float)ratio > max_ratio[i_real]) max_ratio[i_real] = (float)ratio;
        }

        if (f_coll_stars * eff > neutral_fraction * (1. + rec)) {
          if (xH[i_real] > ORC_REL_TOL) {
            if (uvb) J_21[i_real] = J_21_aux;
            if (recomb) Gamma12[i_real] = (float)(Gamma_R_prefactor * weighted_sfr_density);
          }
          xH[i_real] = 0;
          r_bubble[i_real] = (float)R;
        } else if (last && (xH[i_real] > ORC_REL_TOL)) {
          xH[i_real] = (float)(neutral_fraction - f_coll_stars * eff);
          if (xH[i_real] < 0.)
            xH[i_real] = (float)0.;
          else if (xH[i_real] > 1.0)
            xH[i_real] = (float)1.;
        }

        if ((xH[i_real] < ORC_REL_TOL) && (z_at_ionization[i_real] < 0)) {
          z_at_ionization[i_real] = (float)redshift;
          if (uvb) J_21_at_ionization[i_real] = J_21_aux * (float)p->gamma_halo_bias;
        }
      }
    }
  }

  /* find_HII_bubbles.c:451-494: sequential double sums, mass weight from the input deltax */
  if (recomb) {
    const int rc = orc_rr_build(p->hubble_h, p->omega_m, p->baryon_frac, p->y_he);
    if (rc != MRX_OK) return rc;
  }
  const float zstep = (float)(p->prev_redshift - redshift);
  const float fabs_dtdz = (float)fabs(orc_dtdz(p->omega_lambda, p->omega_m, p->hubble_h, (float)redshift) /
                                      p->hubble_h);
  double vol_xH = 0.0, mass_xH = 0.0, mass_weight = 0.0, vol_J21 = 0.0;
  for (long r = 0; r < (long)N * N; r++)
    for (int z = 0; z < N; z++) {
      const long i_real = r * N + z;
      const long i_padded = r * PR + z;
      double cell_xH = (double)xH[i_real];
      vol_xH += cell_xH;
      if (uvb) vol_J21 += (double)J_21[i_real];
      double density_over_mean = 1.0 + (double)deltax[i_padded];
      mass_xH += cell_xH * density_over_mean;
      mass_weight += density_over_mean;
      if (recomb) { /* :472-477 */
        const float z_eff = (float)((1. + redshift) * pow(density_over_mean, 1.0 / 3.0) - 1);
        const double dNrec = orc_rr_rate(z_eff, Gamma12[i_real]) * fabs_dtdz * zstep * (1. - cell_xH);
        N_rec[i_padded] += dNrec;
      }
    }
  vol_xH /= total_n_cells;
  if (uvb) vol_J21 /= total_n_cells;
  mass_xH /= mass_weight;
  if (out) {
    out->volume_weighted_global_xH = vol_xH;
    out->mass_weighted_global_xH = mass_xH;
    out->volume_weighted_global_J_21 = vol_J21;
    out->n_radii = S;
  }
  for (int f = 0; f < nf; f++) {
    free(unf[f]);
    free(fil[f]);
  }
  return MRX_OK;
}
