/*
 * mrx_reion.h - C ABI of the MI355X-native excursion-set reionization path.
 *
 * Plain C, no globals, no torch or HIP types in any signature.  Every grid is
 * float32 in the reference's own layouts (src/core/misc_tools.c:200-220):
 *
 *   padded real  (INDEX_PADDED)        k + 2(N/2+1)(j + N i)   2*N*N*(N/2+1) floats
 *   unpadded real (INDEX_REAL)         k + N(j + N i)          N*N*N floats
 *   half spectrum (INDEX_COMPLEX_HERM) k + (N/2+1)(j + N i)    N*N*(N/2+1) complex
 *                                                              (interleaved re,im floats)
 *
 * Pointers are host or device memory as selected per call; device pointers are
 * used in place, host pointers are staged through the context's device buffers.
 * Every entry point returns MRX_OK (0) or a negative MRX_E* code; the message of
 * the last failure is available from mrx_last_error().  Nothing aborts or throws
 * across this boundary.  A context is bound to one device and must not be used
 * from two host threads at once.
 */
#ifndef MRX_REION_H
#define MRX_REION_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MRX_OK 0
#define MRX_EINVAL (-1)       /* bad argument (size, null pointer, flag) */
#define MRX_ENOMEM (-2)       /* device or host allocation failed */
#define MRX_EDEVICE (-3)      /* HIP runtime error */
#define MRX_EUNSUPPORTED (-4) /* grid size or option outside what this build implements */
#define MRX_EFILTER (-5)      /* invalid filter type (reionization.c:2160-2165 aborts there) */
#define MRX_EIO (-6)          /* a grid file could not be created, opened or written (mrx_io.h) */

/* Filter windows of filter() (reionization.c:2110-2170). */
#define MRX_FILTER_TOPHAT 0 /* real-space top-hat 3(sin x - x cos x)/x^3, x = kR, applied when x > 1e-4 */
#define MRX_FILTER_SHARPK 1 /* zero the mode when 0.413566994 kR > 1 */
#define MRX_FILTER_GAUSS 2  /* exp(-(0.643 kR)^2 / 2) */

typedef struct mrx_ctx mrx_ctx;

/*
 * Knobs of _find_HII_bubbles (find_HII_bubbles.c:47-495).  In the reference they
 * are read from run_globals.params / .physics / .units; here they are explicit.
 */
typedef struct mrx_hii_params {
  int grid_dim;          /* ReionGridDim (N); the whole box lives on this one GPU.  N is  */
                         /* one of 32 48 64 96 128 160 192 256 320 384 512 640 1024;      */
                         /* others return MRX_EUNSUPPORTED (every entry point taking N)  */
  int filter_type;       /* ReionFilterType, MRX_FILTER_*                                   */
  int rtom_filter_type;  /* ReionRtoMFilterType: 0 top-hat, 1 Gaussian (RtoM, :26-45)     */
  int uvb_flag;          /* ReionUVBFlag: write J_21 and J_21_at_ionization                 */
  int include_spin_temp; /* Flag_IncludeSpinTemp: filter x_e as a 4th field (:144-152)      */
  int pointers_on_device;/* 1: every grid pointer of the call is device memory             */
  double box_size;       /* BoxSize [Mpc/h]                                                 */
  double redshift;       /* ZZ[snapshot]                                                    */
  double r_bubble_max;   /* ReionRBubbleMax [Mpc/h]                                         */
  double r_bubble_min;   /* ReionRBubbleMin [Mpc/h]                                         */
  double delta_r_factor; /* ReionDeltaRFactor                                               */
  double efficiency;     /* physics.ReionEfficiency                                         */
  double omega_m;        /* OmegaM                                                          */
  double rho_crit;       /* run_globals.RhoCrit [internal units] (init.c:204)               */
  double alpha_uv;       /* physics.ReionAlphaUV                                            */
  double nion_phot_per_bary; /* physics.ReionNionPhotPerBary                                */
  double gamma_halo_bias;    /* physics.ReionGammaHaloBias                                  */
  double unit_length_cm;     /* units.UnitLength_in_cm                                      */
  double unit_mass_g;        /* units.UnitMass_in_g                                         */
  double unit_time_s;        /* units.UnitTime_in_s                                         */
  /* inhomogeneous recombinations (mrx_find_hii_bubbles_rec only; r_bubble_max is
   * then physics.ReionRBubbleMaxRecomb, find_HII_bubbles.c:186-190) */
  int include_recombinations;/* Flag_IncludeRecombinations                                   */
  int pad_;                  /* unused, 0                                                    */
  double prev_redshift;      /* ZZ[snapshot-1] (ZZ[0] at snapshot 0): zstep (:83-90)         */
  double hubble_h;           /* Hubble_h                                                     */
  double omega_lambda;       /* OmegaLambda (dtdz, XRayHeatingFunctions.c:377-392)           */
  double baryon_frac;        /* BaryonFrac } the hydrogen density No (reionization.h:17-27)  */
  double y_he;               /* physics.Y_He }                                               */
} mrx_hii_params;

/* Scalar results the reference stores in run_globals.reion_grids (:487-494). */
typedef struct mrx_hii_globals {
  double volume_weighted_global_xH;
  double mass_weighted_global_xH;
  double volume_weighted_global_J_21;
  int n_radii; /* radius steps taken, including the last unfiltered one */
} mrx_hii_globals;

/* Knobs of ComputeBrightnessTemperatureBox, spin-temperature branch with
 * Flag_IncludePecVelsFor21cm = 0 (BrightnessTemperature.c:24-124, :512-545). */
typedef struct mrx_tb_params {
  int grid_dim;
  int include_spin_temp; /* Flag_IncludeSpinTemp: multiply by (1 - T_CMB(1+z)/T_S)          */
  int pointers_on_device;
  double redshift;
  double hubble_h;
  double omega_m;
  double baryon_frac;
} mrx_tb_params;

/*
 * Knobs of _ComputeTs (src/core/ComputeTs.c:78-1020) and the X-ray / Ly-alpha
 * physics it calls (src/core/XRayHeatingFunctions.c); minihalos, Lyman-Werner
 * and peculiar velocities off.
 */
typedef struct mrx_ts_params {
  int grid_dim;
  int pointers_on_device;
  int num_filter_steps;    /* TsNumFilterSteps (<= 64)                                   */
  int heating_filter_type; /* TsHeatingFilterType, MRX_FILTER_*                          */
  int rtom_filter_type;    /* ReionRtoMFilterType (collapse fraction of the first shell) */
  int snapshot;            /* index of this snapshot in redshifts[]                      */
  int n_history;           /* SFR history grids held (run_globals.NstoreSnapshots_SFR)   */
  int reproduce_quirks;    /* 1: volume_ave_xe as the reference computes it (ComputeTs.c:957) */
  double box_size;         /* BoxSize [Mpc/h]                                            */
  double hubble_h, omega_m, omega_lambda, omega_r, omega_k, w_lambda, baryon_frac, y_he;
  double l_xray_gal;            /* LXrayGal [erg/s per Msun/yr]                          */
  double nu_xray_gal_threshold; /* NuXrayGalThreshold [eV]                               */
  double spec_index_xray_gal;   /* SpecIndexXrayGal                                      */
  double nu_xray_soft_cut;      /* NuXraySoftCut [eV]                                    */
  double nu_xray_max;           /* NuXrayMax [eV]                                        */
  double reion_efficiency;      /* ReionEfficiency                                       */
  double nion_phot_per_bary;    /* ReionNionPhotPerBary (Pop II Ly-series normalisation) */
  double max_heating_redshift;  /* ReionMaxHeatingRedshift                               */
  double rho_crit;              /* run_globals.RhoCrit [internal units]                  */
  double unit_length_cm, unit_mass_g, unit_time_s;
  const double* redshifts; /* ZZ[0..snapshot]                                            */
  double* stored_fcoll;    /* collapse-fraction history, entry [snapshot] is written      */
  const char* tables_dir;  /* directory of the 21cmFAST tables (TablesForXHeatingDir)    */
  /* 1: the shell tables use fixed_collapse_fraction and fixed_x_e_ave instead of
   * the two grid means the reference measures (ComputeTs.c:413-475).  The
   * differentiable form holds them fixed (a stop-gradient on the tables). */
  int fixed_globals;
  double fixed_collapse_fraction;
  double fixed_x_e_ave;
} mrx_ts_params;

/* Scalar results _ComputeTs leaves in run_globals.reion_grids (:940-985). */
typedef struct mrx_ts_globals {
  double volume_ave_J_alpha, volume_ave_xalpha, volume_ave_Xheat, volume_ave_Xion;
  double volume_ave_TS, volume_ave_TK, volume_ave_xe;
  double collapse_fraction; /* mean collapse fraction of the first shell           */
  double x_e_ave;           /* <x_e> entering the shell tables                      */
  int no_light;             /* 1 when collapse_fraction == 0 (no sources yet)        */
  int homogeneous;          /* 1 when zp >= ReionMaxHeatingRedshift (RECFAST branch) */
} mrx_ts_globals;

/* ---------------------------------------------------------------------------
 * Context.  Replaces init_CUDA() (meraxes_gpu.cu:127-194, declared at
 * src/meraxes_gpu.h:49-57), which selected the device and kept no buffers; the
 * context also owns the device work buffers, sized for the largest grid seen,
 * so repeated snapshots do not reallocate (the CUDA path mallocs per call,
 * find_HII_bubbles_gpu.cu:151-193).
 */
int mrx_init(int device, mrx_ctx** ctx);
/* Number of visible devices (0 when none or on error): init_CUDA picks
 * rank % count (meraxes_gpu.cu:142). */
int mrx_device_count(void);
int mrx_free(mrx_ctx* ctx);
const char* mrx_last_error(const mrx_ctx* ctx);
/* Synchronise the context's stream (all entry points are synchronous on return
 * unless stated, so this is only needed after mrx_set_stream with async use). */
int mrx_synchronize(mrx_ctx* ctx);
/* Run subsequent work on a caller hipStream_t (NULL: the context's own stream). */
int mrx_set_stream(mrx_ctx* ctx, void* hip_stream);

/* ---------------------------------------------------------------------------
 * Radius ladder of the bubble finder (find_HII_bubbles.c:184-207, :445): fills
 * R_out[0..n) with the radii in Mpc/h, largest first, the last one being the
 * unfiltered cell-size step, and returns n (or a negative error).  Host only.
 */
int mrx_radius_ladder(const mrx_hii_params* p, double* R_out, int max_n);

/* ---------------------------------------------------------------------------
 * find_HII_bubbles.  Drop-in for _find_HII_bubbles(snapshot)
 * (find_HII_bubbles.c:47-495) and _find_HII_bubbles_gpu(snapshot, flag)
 * (find_HII_bubbles_gpu.cu:68-644, declared src/meraxes_gpu.h:53).
 *
 * in : deltax, stars, weighted_sfr (padded real); x_e (padded real, only read
 *      when p->include_spin_temp, may be NULL otherwise)
 * out: xH, r_bubble (unpadded real, overwritten);
 *      J_21 (unpadded real, written when p->uvb_flag, may be NULL otherwise)
 * in/out (state carried across snapshots): z_at_ionization, and
 *      J_21_at_ionization (only with p->uvb_flag)
 * globals: volume/mass weighted xH and volume weighted J_21, as the reference.
 * p->include_recombinations must be 0 here (see mrx_find_hii_bubbles_rec);
 * minihalos are not implemented.
 */
int mrx_find_hii_bubbles(mrx_ctx* ctx, const mrx_hii_params* p, const float* deltax, const float* stars,
                         const float* weighted_sfr, const float* x_e, float* xH, float* r_bubble, float* J_21,
                         float* z_at_ionization, float* J_21_at_ionization, mrx_hii_globals* out);

/* The same with inhomogeneous recombinations (Flag_IncludeRecombinations,
 * find_HII_bubbles.c:155-163, :246-302, :325-404, :472-477): N_rec (padded,
 * in/out) is filtered as one more field, a cell ionises when
 * f_coll*eps > n_f (1 + N_rec_R / (1 + delta_R)), Gamma12 (unpadded, in/out)
 * takes Gamma_R_prefactor(R) * sfr_R / V at a cell's first crossing, and after
 * the ladder N_rec += RR(z_eff, Gamma12) |dt/dz| zstep (1 - xH) with the
 * splined MHR00 recombination rate (recombinations.c).  With
 * p->include_recombinations = 0, N_rec and Gamma12 may be NULL and the call is
 * mrx_find_hii_bubbles. */
int mrx_find_hii_bubbles_rec(mrx_ctx* ctx, const mrx_hii_params* p, const float* deltax, const float* stars,
                             const float* weighted_sfr, const float* x_e, float* N_rec, float* xH, float* r_bubble,
                             float* J_21, float* z_at_ionization, float* J_21_at_ionization, float* Gamma12,
                             mrx_hii_globals* out);

/* ---------------------------------------------------------------------------
 * filter().  Drop-in for filter(box, local_ix_start, slab_nx, grid_dim, R,
 * filter_type) (reionization.c:2110-2170, declared reionization.h:72): multiply
 * the half spectrum in place by W(|k| R), |k| built with the global x index and
 * delta_k = 2 pi / box_size.  `box` is INDEX_COMPLEX_HERM, slab_nx planes.
 */
int mrx_filter(mrx_ctx* ctx, float* box, int local_ix_start, int slab_nx, int grid_dim, float R, int filter_type,
               double box_size, int pointers_on_device);

/* ---------------------------------------------------------------------------
 * 3-D real transforms with FFTW's conventions (unnormalised, forward sign -1),
 * the operations the reference runs through fftwf_execute on its r2c / c2r
 * plans (reionization.c:685-736).  in_padded -> out_herm; in_herm -> out_padded
 * (may alias, as the reference's in-place c2r plans do).
 */
int mrx_fft_r2c_3d(mrx_ctx* ctx, int grid_dim, const float* in_padded, float* out_herm, int pointers_on_device);
int mrx_fft_c2r_3d(mrx_ctx* ctx, int grid_dim, const float* in_herm, float* out_padded, int pointers_on_device);

/* ---------------------------------------------------------------------------
 * One field through the bubble finder's smoothing chain at radius R: R2C,
 * divide by N^3 (find_HII_bubbles.c:168-182), filter (skipped when
 * apply_filter = 0, the last ladder step), C2R, and the aliasing clamp of
 * :272-311 for `clamp_kind`: 0 density (max(v, -1 + 1e-5)), 1 source grid
 * (max(v, 0), zero below 1e-8), 2 x_e (as 1, then min(v, 0.999)), -1 none.
 * Output is padded real.  This exposes the filtered fields the threshold sees.
 */
int mrx_filtered_field(mrx_ctx* ctx, int grid_dim, double box_size, int filter_type, const float* field_padded,
                       double R, int apply_filter, int clamp_kind, float* out_padded, int pointers_on_device);

/* ---------------------------------------------------------------------------
 * Spin temperature.  Drop-in for _ComputeTs(snapshot) (ComputeTs.c:78-1020),
 * dispatched where ComputeTs() (:1024-1045) calls it.
 *
 * in : deltax, stars, sfr (padded real, this snapshot); sfr_histories:
 *      p->n_history grids of 2*N*N*(N/2+1) floats each, slot k = k snapshots
 *      back, stored and read exactly as the reference's sfr_histories buffer
 *      (written with INDEX_PADDED, read with INDEX_REAL, reionization.c:1586,
 *      :1767).  Only read when the shells reach back in time.
 * in/out: x_e_box_prev (padded), Tk_box (unpadded) - the carried state.
 * out: x_e_box (padded, = the new x_e_box_prev), TS_box (unpadded).
 */
int mrx_compute_ts(mrx_ctx* ctx, const mrx_ts_params* p, const float* deltax, const float* stars, const float* sfr,
                   const float* sfr_histories, float* x_e_box, float* x_e_box_prev, float* Tk_box, float* TS_box,
                   mrx_ts_globals* out);

/* ---------------------------------------------------------------------------
 * The host tables of one spin-temperature snapshot: what _ComputeTs forms on
 * the host before its per-cell loop (ComputeTs.c:276-377 shell radii,
 * light-cone edges and SFR-history operations; :517-716 frequency integrals
 * at the 14 x_e knots from the lower limit nu(tau_X = 1)
 * (XRayHeatingFunctions.c:228-375, :685-803), Ly-n sums with the kink fix,
 * growth factor, X-ray prefactor; evolveInt's per-shell factors
 * XRayHeatingFunctions.c:1289-1311).  mrx_compute_ts uses exactly these.  Host
 * only, no device work: for callers that log or check the tables.
 * collapse_fraction and x_e_ave are the two grid means the tables depend on
 * (ComputeTs.c:413-475); like the reference, p->stored_fcoll[p->snapshot]
 * receives collapse_fraction (inhomogeneous snapshots only).  Operations are
 * listed in plan order: op_kind 0 zeroes the first N*N*(N/2+1) floats of the
 * working grid, 1 scales them by op_weight, 2 loads history slot op_slot
 * times (float)op_weight, 3 adds it.  More than MRX_TS_TABLES_MAX_OPS
 * operations in all, or a shell needing more than 256, is MRX_EUNSUPPORTED.
 */
#define MRX_TS_TABLES_MAX_STEPS 64
#define MRX_TS_TABLES_MAX_OPS 1024
typedef struct mrx_ts_tables {
  int nsteps, homogeneous, no_light, n_ops;
  double zp, dzp, dt_dzp, const_zp_prefactor, growth_factor_zp, dgrowth_factor_dzp;
  double R[MRX_TS_TABLES_MAX_STEPS];        /* shell radii [Mpc]                               */
  double zpp_edge[MRX_TS_TABLES_MAX_STEPS]; /* light-cone redshift edges                       */
  double zpp[MRX_TS_TABLES_MAX_STEPS], dzpp[MRX_TS_TABLES_MAX_STEPS], dt_dzpp[MRX_TS_TABLES_MAX_STEPS];
  double zpp_pow[MRX_TS_TABLES_MAX_STEPS];  /* (1 + z'')^-SpecIndexXrayGal                     */
  double sum_lyn[MRX_TS_TABLES_MAX_STEPS];
  double nu_lower[MRX_TS_TABLES_MAX_STEPS]; /* max(nu(tau_X = 1), NuXrayGalThreshold) [Hz]     */
  double freq_heat[14][MRX_TS_TABLES_MAX_STEPS], freq_ion[14][MRX_TS_TABLES_MAX_STEPS];
  double freq_lya[14][MRX_TS_TABLES_MAX_STEPS];
  int op_shell[MRX_TS_TABLES_MAX_OPS], op_kind[MRX_TS_TABLES_MAX_OPS], op_slot[MRX_TS_TABLES_MAX_OPS];
  double op_weight[MRX_TS_TABLES_MAX_OPS];
} mrx_ts_tables;

int mrx_ts_snapshot_tables(const mrx_ts_params* p, double collapse_fraction, double x_e_ave, mrx_ts_tables* out);

/* The smoothed SFR shells of the context's last mrx_compute_ts call on the
 * inhomogeneous branch: per shell max(filtered sfr, 0) in float
 * (ComputeTs.c:379-503), before the unit factor the reference folds into its
 * double SMOOTHED_SFR_GAL (:419-421, :487-489).  Shell-major, unpadded:
 * out receives n_shells * grid_dim^3 floats (device memory when
 * pointers_on_device).  MRX_EINVAL when the last call left no shells (the
 * homogeneous branch) or the sizes differ.  The reference keeps these only
 * inside run_globals; this is for callers that inspect them. */
int mrx_ts_last_shells(mrx_ctx* ctx, int grid_dim, int n_shells, float* out, int pointers_on_device);

/* ---------------------------------------------------------------------------
 * Brightness temperature.  Drop-in for ComputeBrightnessTemperatureBox(snapshot)
 * (BrightnessTemperature.c:24-124, :512-545; declared BrightnessTemperature.h:9)
 * for Flag_IncludePecVelsFor21cm = 0.  TS_box is read only with
 * p->include_spin_temp.  *ave_Tb receives volume_ave_Tb.
 */
int mrx_brightness_temp(mrx_ctx* ctx, const mrx_tb_params* p, const float* xH, const float* deltax,
                        const float* TS_box, float* delta_T, double* ave_Tb);

/* ---------------------------------------------------------------------------
 * No counterpart in the reference.  mrx_brightness_temp for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): it allocates nothing, host or
 * device.  Its partial sums live in a fixed device array of 2056 doubles
 * (16 KiB + 64 B) that each device holds until the process ends; that size
 * does not depend on the grid.  Device pointers only (p->pointers_on_device
 * = 1, else MRX_EINVAL); grid_dim even, 2 <= grid_dim <= 2048 (above:
 * MRX_EUNSUPPORTED).  Results equal mrx_brightness_temp's bit for bit.  All
 * calls on one device share that array: calls that overlap on one device
 * (other contexts or streams) race, so serialise them.
 */
int mrx_brightness_temp_noalloc(mrx_ctx* ctx, const mrx_tb_params* p, const float* xH, const float* deltax,
                                const float* TS_box, float* delta_T, double* ave_Tb);

/* ---------------------------------------------------------------------------
 * Kernel timing, the device-side counterpart of the reference's timer_info
 * (src/core/utils.c:205-220) around find_HII_bubbles / ComputeTs
 * (find_HII_bubbles.c:498-531).  While enabled, every kernel launch is
 * bracketed by HIP events on the stream it runs on.  mrx_profile_read()
 * synchronises, writes up to max_kinds kernel names (each name_len bytes into
 * `names`), their summed milliseconds and launch counts, clears the record and
 * returns the number of distinct kernels.
 */
int mrx_profile(mrx_ctx* ctx, int enable);
int mrx_profile_read(mrx_ctx* ctx, int max_kinds, char* names, int name_len, double* ms_total, long* counts);

/* ---------------------------------------------------------------------------
 * 21-cm power spectrum.  Drop-in for Compute_PS (ComputePowerSpectrum.c:44-255,
 * bins from Initialise_PowerSpectrum :18-41): the field is made
 * ((delta_T/<delta_T> - 1) V/N^3) <delta_T>, transformed, and every mode of the
 * half spectrum adds k^3 |F|^2 / (2 pi^2 V) to its bin; bins [0, dk), [dk, f dk),
 * [f dk, f^2 dk) ... while the ceiling stays below N dk, dk = 2 pi / box_size_mpc
 * (box_size_mpc = BoxSize / Hubble_h), f = k_factor (1.35 when <= 1, the
 * reference's).  delta_T is an unpadded device grid.  Writes up to max_bins
 * entries of the bin-mean k, the power (PS_data), its error (PS_error) and
 * the mode counts (each array nullable) and returns the number of bins.
 *
 * mrx_power_spectrum_adjoint: no counterpart in the reference.  For callers
 * that differentiate a loss of the binned power: given dL/dP_b, writes
 * dL/d(delta_T) (unpadded device grid).
 */
int mrx_power_spectrum(mrx_ctx* ctx, int grid_dim, double box_size_mpc, float k_factor, const float* delta_T,
                       int max_bins, double* k, double* power, double* error, long* counts);
int mrx_power_spectrum_adjoint(mrx_ctx* ctx, int grid_dim, double box_size_mpc, float k_factor, const float* delta_T,
                               const double* grad_power, float* grad_delta_T);

/* ---------------------------------------------------------------------------
 * Adjoint of one snapshot: _ComputeTs -> _find_HII_bubbles ->
 * ComputeBrightnessTemperatureBox (ComputeTs.c:78-1020, find_HII_bubbles.c:47-495,
 * BrightnessTemperature.c:24-124).  No counterpart in the reference; for
 * callers that differentiate a loss of delta_T.  Given the forward's inputs
 * (x_e_box_prev and Tk_box as they were BEFORE mrx_compute_ts updated them),
 * its outputs (xH, TS_box, x_e_box) and dL/d(delta_T), writes dL/d(deltax),
 * dL/d(stars) (padded), dL/d(sfr), dL/d(sfr_histories) (padded, per slot),
 * dL/d(x_e_box_prev) (padded) and dL/d(Tk_box) (each of the last four
 * nullable).  The ionisation mask is a stop-gradient (0 almost everywhere);
 * cells partly ionised on the unfiltered last step carry the derivative of
 * xH = n_f - eff f_coll.  `ts` NULL: no spin temperature (tb and hii must not
 * use it).  With ts, the gradient reaches through the shell tables' two grid
 * means (the collapse fraction and <x_e>, ComputeTs.c:413-475): their
 * derivatives are taken by central differences of the host table build and
 * spread back over the cells the means are taken from.  ts->fixed_globals
 * holds them at ts->fixed_collapse_fraction and ts->fixed_x_e_ave instead,
 * forward and adjoint alike (a stop-gradient on the means).
 * Per-radius filtered fields are recomputed, not stored.  Device pointers only.
 */
int mrx_snapshot_adjoint(mrx_ctx* ctx, const mrx_ts_params* ts, const mrx_hii_params* hii, const mrx_tb_params* tb,
                         const float* deltax, const float* stars, const float* sfr, const float* sfr_histories,
                         const float* x_e_box_prev, const float* Tk_box, const float* xH, const float* TS_box,
                         const float* x_e_box, const float* grad_delta_T, float* grad_deltax, float* grad_stars,
                         float* grad_sfr, float* grad_sfr_histories, float* grad_x_e_box_prev, float* grad_Tk_box);

/* ---------------------------------------------------------------------------
 * Batches of independent boxes.  No counterpart in the reference, which runs
 * one box per call (find_HII_bubbles.c:498-532); each box's results equal
 * those of mrx_find_hii_bubbles on that box alone.  For callers that push many
 * small or uneven boxes (parameter sweeps, ensembles) through the bubble
 * finder: boxes sharing grid_dim, box_size, the ladder (r_bubble_max/min,
 * delta_r_factor), the filter, the flags and omega_m * rho_crit form a cohort
 * whose kernels cover all its boxes at once; cohorts run side by side.
 * efficiency, redshift, gamma_halo_bias, alpha_uv, nion_phot_per_bary and the
 * units may differ between the boxes of a cohort.
 *
 * mrx_hii_batch_create validates the boxes (device pointers only, 16-byte
 * aligned, the layouts of mrx_find_hii_bubbles), allocates every work buffer,
 * times the tile-width / radix-schedule variants of each cohort's strided
 * passes and keeps the fastest (unless MRX_BATCH_NO_TUNE), and captures the
 * whole computation - every radius of every box - into one HIP graph (unless
 * MRX_BATCH_NO_GRAPH).  The grids are bound to the plan: mrx_hii_batch_run
 * reads the current contents of the same device buffers on each call and
 * replays the graph as one launch on the context stream, then, if `out` is
 * not NULL, waits and writes one mrx_hii_globals per box (in box order).
 * mrx_hii_batch_describe writes a JSON summary (cohorts, radii, chosen
 * variants, measured variant times) and returns its length.
 */
typedef struct mrx_hii_batch_box {
  mrx_hii_params params; /* params.pointers_on_device must be 1 */
  const float* deltax;
  const float* stars;
  const float* weighted_sfr;
  const float* x_e; /* read when params.include_spin_temp */
  float* xH;
  float* r_bubble;
  float* J_21; /* written when params.uvb_flag */
  float* z_at_ionization;
  float* J_21_at_ionization; /* with params.uvb_flag */
} mrx_hii_batch_box;

typedef struct mrx_hii_batch mrx_hii_batch;

#define MRX_BATCH_NO_GRAPH 1 /* enqueue the kernels on each run instead of replaying a graph */
#define MRX_BATCH_NO_TUNE 2  /* default tile and radix schedules, nothing measured */

int mrx_hii_batch_create(mrx_ctx* ctx, const mrx_hii_batch_box* boxes, int n_boxes, int flags,
                         mrx_hii_batch** out);
int mrx_hii_batch_run(mrx_hii_batch* batch, mrx_hii_globals* out);
int mrx_hii_batch_describe(const mrx_hii_batch* batch, char* buf, int len);
int mrx_hii_batch_free(mrx_hii_batch* batch);

/*
 * Peculiar-velocity branch of ComputeBrightnessTemperatureBox
 * (BrightnessTemperature.c:126-510), applied (as in the reference) only with
 * include_spin_temp and pec_vels > 0; otherwise the call is mrx_brightness_temp.
 * vel (padded) is converted in place to comoving velocity (:166-195, by
 * trees_id); its z-gradient is formed by FFT (/N^3, *Hubble_h, *i k_z,
 * velocity_gradient reionization.c:2172-2187); then
 *   pec_vels 1: delta_T /= (dv/dz capped at |MAX_DVDR H(z)|) / H(z) + 1        (:210-236)
 *   pec_vels 2: optical depth tau = delta_T (1+z)/(1000 T_S) / |dv/dz / H(z) + 1|,
 *               delta_T = (1 - e^-tau) 1000 (T_S - T_CMB(1+z)) / (1+z)       (:240-287)
 *   pec_vels 3: 2, then line-of-sight redshift-space distortions: every cell
 *               split into N_RSD_STEPS = 50 sub-cells moved by the interpolated
 *               displacement v / H(z) and re-binned along z (:312-496).
 */
typedef struct mrx_tb_vel_params {
  mrx_tb_params tb;      /* as for mrx_brightness_temp                                         */
  int pec_vels;          /* Flag_IncludePecVelsFor21cm: 0..3                                   */
  int trees_id;          /* TreesID: 0 VELOCIRAPTOR, 1 GBPTREES, 2 VELOCIRAPTOR_AUG (:171-182)  */
  double box_size;       /* BoxSize [Mpc/h]                                                    */
  double omega_lambda;   /* OmegaLambda  } hubble(z), XRayHeatingFunctions.c:394-404           */
  double omega_r;        /* OmegaR       }                                                     */
} mrx_tb_vel_params;

/* Replaces ComputeBrightnessTemperatureBox (BrightnessTemperature.c:24) with
 * Flag_IncludePecVelsFor21cm > 0.  vel: padded velocity grid, converted in
 * place as the reference does. */
int mrx_brightness_temp_vel(mrx_ctx* ctx, const mrx_tb_vel_params* p, const float* xH, const float* deltax,
                            const float* TS_box, float* vel, float* delta_T, double* ave_Tb);

/*
 * Galaxy gridding: construct_baryon_grids (reionization.c:1353-1615) for one
 * process holding the whole box.  Each galaxy that is not dead (type <= 2) is
 * deposited on its nearest grid point, pos_to_ngp (misc_tools.c:163-173):
 * i = nearbyint(pos / box_size * N), wrapped to 0 when > N-1; a negative index
 * is an error (the reference asserts).  Per cell, the contributions are summed
 * in galaxy order into a float exactly as the reference's `buffer[ind] += v`
 * (v double, the sum rounded to float after each addition):
 *   stars        += fesc_weighted_gsm  (+ effective_bhm when bh_feedback and
 *                                        black_hole_mass >= bh_mass_limit_reion)
 *   weighted_sfr += fesc_weighted_gsm  (+ effective_bhm * alpha_uv_bh / alpha_uv, same condition)
 *   sfr          += gross_stellar_mass (include_spin_temp only)
 * then written to the padded grids (zero elsewhere): stars clamped at 0;
 * weighted_sfr and sfr as (float)(v / sfr_timescale) where v > 0, else 0,
 * with sfr_timescale = reion_sfr_timescale * hubble_time(snapshot)
 * (virial_properties.c:79-93).  With include_spin_temp the SFR history ring
 * (n_store_sfr padded grids) is shifted by one slot first and slot 0 receives
 * the new sfr grid at every real cell (its padding is left as it was, as in the
 * reference).
 */
typedef struct mrx_grid_params {
  int grid_dim;             /* ReionGridDim (N)                                              */
  int include_spin_temp;    /* Flag_IncludeSpinTemp: also the sfr grid and the history ring  */
  int bh_feedback;          /* physics.Flag_BHFeedback                                       */
  int n_store_sfr;          /* NstoreSnapshots_SFR: slots of sfr_histories                   */
  int pointers_on_device;   /* 1: the galaxy arrays and grids are device memory             */
  int pad_;                 /* unused, 0                                                     */
  double box_size;          /* BoxSize [Mpc/h]                                               */
  double reion_sfr_timescale; /* ReionSfrTimescale                                           */
  double hubble;            /* run_globals.Hubble [internal units]                           */
  double omega_m;           /* OmegaM                                                        */
  double omega_k;           /* OmegaK                                                        */
  double omega_lambda;      /* OmegaLambda                                                   */
  double redshift;          /* ZZ[snapshot]                                                  */
  double bh_mass_limit_reion; /* physics.BlackHoleMassLimitReion                             */
  double alpha_uv_bh;       /* physics.ReionAlphaUVBH                                        */
  double alpha_uv;          /* physics.ReionAlphaUV                                          */
} mrx_grid_params;

typedef struct mrx_grid_stats {
  long n_gridded;           /* galaxies deposited (type <= 2)                                */
  long n_dead;              /* galaxies skipped (type > 2)                                   */
  long n_bh_below_limit;    /* N_BlackHoleMassLimitReion (bh_feedback, :1473-1478)           */
  double sfr_timescale;     /* the divisor used                                              */
} mrx_grid_stats;

/* Replaces construct_baryon_grids(snapshot, local_ngals) (reionization.c:1353),
 * with the galaxy fields as arrays (galaxy_t, meraxes.h:547-635: Pos, Type,
 * FescWeightedGSM, GrossStellarMass, BlackHoleMass, EffectiveBHM) in the
 * order of galaxy_to_slab_map.  pos is [n_gal][3] floats in Mpc/h;
 * gross_stellar_mass is needed with include_spin_temp, black_hole_mass and
 * effective_bhm with bh_feedback (otherwise nullable); sfr and sfr_histories
 * with include_spin_temp.  n_gal < 2^31. */
int mrx_construct_baryon_grids(mrx_ctx* ctx, const mrx_grid_params* p, long n_gal, const float* pos,
                               const int* type, const double* fesc_weighted_gsm, const double* gross_stellar_mass,
                               const double* black_hole_mass, const double* effective_bhm, float* stars,
                               float* weighted_sfr, float* sfr, float* sfr_histories, mrx_grid_stats* stats);

/* ---------------------------------------------------------------------------
 * Device copy rate.  No counterpart in the reference: copies `bytes` (a
 * multiple of 16, 16-byte aligned device buffers) from src to dst reps times
 * on the context stream with a streaming kernel (one 16-byte load and one
 * non-temporal store per thread, a grid covering the buffer) after one
 * warm-up copy, and writes
 * the average milliseconds per copy measured with HIP events on that stream.
 * For callers (the benchmark) that set achieved bandwidth beside the device's
 * practical copy rate. */
int mrx_stream_copy(mrx_ctx* ctx, void* dst, const void* src, size_t bytes, int reps, double* ms_per_copy);

/* sizeof of the ABI structs (0 mrx_hii_params, 1 mrx_hii_globals, 2 mrx_tb_params,
 * 3 mrx_ts_params, 4 mrx_ts_globals, 5 mrx_hii_batch_box, 6 mrx_grid_params,
 * 7 mrx_grid_stats, 8 mrx_tb_vel_params, 9 mrx_ts_tables), so a binding can check that its
 * mirror of the layout matches this build. */
size_t mrx_abi_sizeof(int which);

#ifdef __cplusplus
}
#endif

#endif /* MRX_REION_H */
