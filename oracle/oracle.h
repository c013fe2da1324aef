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

/* _ComputeTs restated (ts.c).  Same pointer contract as mrx_compute_ts (host
 * memory); shells_out, if not NULL, receives the num_filter_steps smoothed SFR
 * shells (float, shell-major, unpadded) of the inhomogeneous branch. */
int orc_ts_export(const mrx_ts_params* p, double* out, int cap);
int orc_compute_ts(const mrx_ts_params* p, const float* deltax, const float* stars, const float* sfr,
                   const float* sfr_histories, float* x_e_box, float* x_e_box_prev, float* Tk_box, float* TS_box,
                   mrx_ts_globals* out, float* shells_out);

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
 * the recombination-rate table is recomb_tables.c (shared with the library). */
int orc_find_hii_bubbles_rec(int precision, const mrx_hii_params* p, const float* deltax, const float* stars,
                             const float* weighted_sfr, const float* x_e, float* N_rec, float* xH, float* r_bubble,
                             float* J_21, float* z_at_ionization, float* J_21_at_ionization, float* Gamma12,
                             mrx_hii_globals* out, float* max_ratio);

#ifdef __cplusplus
}
#endif

#endif
