/*
 * oracle.h - CPU restatement of the reference's reionization hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing under oracle/ is part of the product: only
 * the tests, the smoke check and the CPU-baseline leg of the benchmark load it,
 * and only as the checker.  The HIP library never links or calls it.
 *
 * Two precisions, as the reference's own mix dictates:
 *   ref32  float grids, float FFT, float window (libm sinf/cosf/powf exactly as
 *          filter() calls them), double per-cell arithmetic - the reference's
 *          FFTW-float CPU path (find_HII_bubbles.c:47-495, reionization.c:2110-2170).
 *   ref64  the same algorithm with every grid, FFT and window in double: the
 *          FP64 oracle that bounds the error of both ref32 and the GPU.
 *
 * FFTW and GSL are not available here; the FFT is this file set's own
 * mixed-radix transform with FFTW's conventions (unnormalised, forward sign -1,
 * r2c output of N*N*(N/2+1) complex in INDEX_COMPLEX_HERM order).
 */
#ifndef MRX_ORACLE_H
#define MRX_ORACLE_H

#include "../include/mrx_reion.h"

#ifdef __cplusplus
extern "C" {
#endif

/* 1-D complex transform of length n (any n >= 1), interleaved re/im, out of
 * place, sign -1 forward / +1 inverse, unnormalised. */
void orc_fft1d_f32(const float* in, float* out, int n, int sign);
void orc_fft1d_f64(const double* in, double* out, int n, int sign);

/* 3-D r2c: padded real (N,N,2(N/2+1)) -> half spectrum (N,N,N/2+1) complex. */
void orc_r2c3d_f32(int N, const float* in_padded, float* out_herm);
void orc_r2c3d_f64(int N, const double* in_padded, double* out_herm);
/* 3-D c2r: half spectrum -> padded real; in and out may alias. */
void orc_c2r3d_f32(int N, const float* in_herm, float* out_padded);
void orc_c2r3d_f64(int N, const double* in_herm, double* out_padded);

/* filter() restated (reionization.c:2110-2170); float window, float spectrum. */
int orc_filter_f32(float* box_herm, int local_ix_start, int slab_nx, int N, float R, int filter_type,
                   double box_size);
/* FP64 window on a double spectrum. */
int orc_filter_f64(double* box_herm, int N, double R, int filter_type, double box_size);
/* The float window value filter() multiplies mode (nx, ny, nz) by (1 when untouched). */
float orc_window_f32(int nx, int ny, int nz, int N, float R, int filter_type, double box_size);

/* Radius ladder (find_HII_bubbles.c:184-207, :445). */
int orc_radius_ladder(const mrx_hii_params* p, double* R_out, int max_n);

/* One field through R2C, /N^3, filter, C2R and the aliasing clamp. */
int orc_filtered_field_f32(int N, double box_size, int filter_type, const float* field_padded, double R,
                           int apply_filter, int clamp_kind, float* out_padded);
int orc_filtered_field_f64(int N, double box_size, int filter_type, const float* field_padded, double R,
                           int apply_filter, int clamp_kind, double* out_padded);

/*
 * _find_HII_bubbles restated.  precision 32 = ref32, 64 = ref64.  Same pointer
 * contract as mrx_find_hii_bubbles (host memory).  If max_ratio is not NULL it
 * receives, per cell (unpadded), max over the ladder of
 * f_coll*eff / (n_f*(1+rec)): the cell is ionised iff this exceeds 1, so it
 * is the margin that any flip between two implementations must sit on.
 */
int orc_find_hii_bubbles(int precision, const mrx_hii_params* p, const float* deltax, const float* stars,
                         const float* weighted_sfr, const float* x_e, float* xH, float* r_bubble, float* J_21,
                         float* z_at_ionization, float* J_21_at_ionization, mrx_hii_globals* out, float* max_ratio);

/* _ComputeTs restated (ts.c).  The snapshot's host tables (shell plan,
 * frequency integrals, Ly-n sums, prefactors, spline coefficients) come from
 * the oracle's own numpy restatement (oracle/ts_host.py) in an orc_ts_tab;
 * everything per cell is restated in ts.c.  The kappa splines are natural
 * cubic splines given by their knots and second-derivative coefficients c. */
#define ORC_TS_MAX_STEPS 64
#define ORC_TS_MAX_OPS 4096
#define ORC_XHII 14
typedef struct orc_spline {
  int n;
  double x[32], y[32], c[32];
} orc_spline;
typedef struct orc_ts_tab {
  int nsteps, homogeneous, no_light, n_ops;
  double zp, dzp, dt_dzp, const_zp_prefactor, growth_factor_zp, dgrowth_factor_dzp;
  double No, N_b0, f_H, f_He;
  double xe_recfast, T_recfast; /* xion_RECFAST((float)zp), T_RECFAST((float)zp): homogeneous branch */
  double R[ORC_TS_MAX_STEPS], dzpp[ORC_TS_MAX_STEPS], dt_dzpp[ORC_TS_MAX_STEPS], zpp[ORC_TS_MAX_STEPS];
  double zpp_pow[ORC_TS_MAX_STEPS], sum_lyn[ORC_TS_MAX_STEPS];
  double freq_heat[ORC_XHII][ORC_TS_MAX_STEPS], freq_ion[ORC_XHII][ORC_TS_MAX_STEPS];
  double freq_lya[ORC_XHII][ORC_TS_MAX_STEPS];
  float xHII[ORC_XHII];
  orc_spline kappa_HH, kappa_eH, kappa_pH;
  /* SFR-history operations in plan order: 0 zero the first N*N*(N/2+1) floats,
   * 1 scale them by weight, 2 load slot * (float)weight, 3 add it */
  int op_shell[ORC_TS_MAX_OPS], op_kind[ORC_TS_MAX_OPS], op_slot[ORC_TS_MAX_OPS];
  double op_weight[ORC_TS_MAX_OPS];
} orc_ts_tab;

/* The first shell's grid means (ComputeTs.c:225-257, :413-475): the mean
 * collapse fraction at radius R (Mpc) and <x_e_box_prev>, summed in the
 * reference's cell order. */
int orc_ts_first_shell_means(const mrx_ts_params* p, double R, const float* deltax, const float* stars,
                             const float* x_e_box_prev, double* collapse_fraction, double* x_e_ave);
/* Same pointer contract as mrx_compute_ts (host memory); shells_out, if not
 * NULL, receives the num_filter_steps smoothed SFR shells (float,
 * shell-major, unpadded) of the inhomogeneous branch.  The globals reported
 * as collapse_fraction / x_e_ave are the arguments (what the tables used). */
int orc_compute_ts(const mrx_ts_params* p, const orc_ts_tab* t, double collapse_fraction, double x_e_ave,
                   const float* deltax, const float* stars, const float* sfr, const float* sfr_histories,
                   float* x_e_box, float* x_e_box_prev, float* Tk_box, float* TS_box, mrx_ts_globals* out,
                   float* shells_out);

/* ComputeBrightnessTemperatureBox restated (no peculiar velocities). */
int orc_brightness_temp(const mrx_tb_params* p, const float* xH, const float* deltax, const float* TS_box,
                        float* delta_T, double* ave_Tb);

/* Threads the oracle uses (OpenMP); <= 0 leaves the runtime default. */
void orc_set_threads(int n);
int orc_get_threads(void);

/* construct_baryon_grids restated (reionization.c:1353-1615), one rank, host
 * arrays; stats = {n_gridded, n_dead, n_bh_below_limit}.  Returns 0, or
 * -100 - g for the first galaxy g (capped at 1e6) with a negative grid index. */
int orc_construct_baryon_grids(const mrx_grid_params* p, long n_gal, const float* pos, const int* type,
                               const double* fesc, const double* gross, const double* bhm, const double* ebhm,
                               float* stars, float* wsfr, float* sfr, float* hist, long* stats);

/* ComputeBrightnessTemperatureBox with peculiar velocities
 * (BrightnessTemperature.c:126-510); vel (padded) is converted in place. */
int orc_brightness_temp_vel(const mrx_tb_vel_params* p, const float* xH, const float* deltax, const float* TS_box,
                            float* vel, float* delta_T, double* ave_Tb);
/* Its line-of-sight re-binning alone (:312-496): delta_T in place, vel converted. */
int orc_rsd_los(const mrx_tb_vel_params* p, const float* xH, const float* vel, float* delta_T);

/* _find_HII_bubbles with inhomogeneous recombinations (find_HII_bubbles.c:155-163,
 * :246-302, :325-404, :472-477): N_rec (padded) and Gamma12 (unpadded) in/out;
 * the recombination-rate table is recomb.c's (built on first use per cosmology). */
int orc_find_hii_bubbles_rec(int precision, const mrx_hii_params* p, const float* deltax, const float* stars,
                             const float* weighted_sfr, const float* x_e, float* N_rec, float* xH, float* r_bubble,
                             float* J_21, float* z_at_ionization, float* J_21_at_ionization, float* Gamma12,
                             mrx_hii_globals* out, float* max_ratio);

/* recomb.c: recombinations.c restated (MHR00 PDF, self-shielding, rate table)
 * with the oracle's own QAG-61 and natural spline.  orc_set_gk61 installs the
 * 61-point rule (31 nonnegative nodes descending, their Kronrod weights, the
 * 15 Gauss weights of the odd nodes) before any integral is taken. */
void orc_set_gk61(const double* xgk, const double* wgk, const double* wg);
double orc_qag61(double (*f)(double, void*), void* par, double a, double b, double epsrel, int limit,
                 double* abserr);
int orc_rr_build(double hubble_h, double omega_m, double baryon_frac, double y_he);
double orc_rr_rate(double z_eff, double gamma12_bg);
double orc_rr_A(double z);
double orc_rr_C(double z);
double orc_rr_beta(double z);
double orc_rr_Gamma_SS(double Gamma_bg, double Delta, double T_4, double z);
double orc_rr_neutral_fraction(double y_he, double density, double T4, double gamma12);
double orc_rr_recombination_rate(double z_eff, double gamma12_bg, double T4);
double orc_rr_A_aux_integral(double z);
double orc_rr_table(int z_ct, int g);
double orc_rr_lnG(int g);
double orc_dtdz(double omega_l, double omega_m, double hubble_h, float z);

#ifdef __cplusplus
}
#endif

#endif
