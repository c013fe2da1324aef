/*
 * ts_tables.h - host-side scalar physics of the spin-temperature step.
 *
 * Everything ComputeTs needs that is not per cell: the 21cmFAST lookup tables
 * and their splines, the shell radii and light-cone redshift edges, the
 * bookkeeping of which past SFR grids make up each shell, and the per-snapshot
 * frequency-integral tables (src/core/ComputeTs.c:32-76, :270-716 and
 * src/core/XRayHeatingFunctions.c).  Plain C99, FP64, run once per snapshot on
 * the host; both the HIP library and the CPU oracle use it, and its numerics
 * are pinned against scipy in the tests.
 */
#ifndef MRX_TS_TABLES_H
#define MRX_TS_TABLES_H

#include "../../include/mrx_reion.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MRX_TS_MAX_STEPS 64
#define MRX_TS_MAX_OPS 24        /* SFR-history operations per k_sfr_ops launch (a by-value kernel argument) */
#define MRX_TS_MAX_SHELL_OPS 256 /* operations one shell may need; a plan needing more is MRX_EUNSUPPORTED */
#define MRX_X_INT_NXHII 14
#define MRX_X_INT_NENERGY 258
#define MRX_NSPEC_MAX 23
#define MRX_RECFAST_NPTS 501

/* natural cubic spline (the interpolation gsl_interp_cspline builds) */
typedef struct mrx_spline {
  int n;
  double x[MRX_RECFAST_NPTS], y[MRX_RECFAST_NPTS];
  double b[MRX_RECFAST_NPTS], c[MRX_RECFAST_NPTS], d[MRX_RECFAST_NPTS];
} mrx_spline;

int mrx_spline_init(mrx_spline* s, const double* x, const double* y, int n);
double mrx_spline_eval(const mrx_spline* s, double x);

/* Tables read from the 21cmFAST files plus cosmology-derived constants. */
typedef struct mrx_ts_data {
  mrx_spline recfast_T, recfast_xe;
  double recfast_zmax;
  mrx_spline kappa_HH, kappa_eH, kappa_pH; /* log T -> log kappa */
  double kHH_t0, kHH_tn, kHH_kn;            /* end points for the clamps/extrapolation */
  int spec_n[MRX_NSPEC_MAX];
  float spec_nu[MRX_NSPEC_MAX], spec_N0_2[MRX_NSPEC_MAX], spec_alpha_2[MRX_NSPEC_MAX];
  float spec_N0_3[MRX_NSPEC_MAX], spec_alpha_3[MRX_NSPEC_MAX];
  float xi_XHII[MRX_X_INT_NXHII];
  float xi_E[MRX_X_INT_NENERGY];
  float xi_fheat[MRX_X_INT_NXHII][MRX_X_INT_NENERGY];
  float xi_nLya[MRX_X_INT_NXHII][MRX_X_INT_NENERGY];
  float xi_nHI[MRX_X_INT_NXHII][MRX_X_INT_NENERGY];
  float xi_nHeI[MRX_X_INT_NXHII][MRX_X_INT_NENERGY];
  float xi_nHeII[MRX_X_INT_NXHII][MRX_X_INT_NENERGY];
  /* cosmology */
  double hubble_h, omega_m, omega_l, omega_r, omega_k, w_lambda;
  double No, He_No, f_H, f_He, N_b0;
  double reion_efficiency;
} mrx_ts_data;

/* One operation on the working SFR grid before a shell's transform
 * (ComputeTs.c:296-376, load_reion_sfr_grids reionization.c:1756-1784). */
enum {
  MRX_SFR_ZERO_HALF = 0, /* sfr[i] = 0 for the first N*N*(N/2+1) floats       */
  MRX_SFR_SCALE_HALF = 1,/* sfr[i] *= weight (double) for the same floats       */
  MRX_SFR_LOAD = 2,      /* sfr[padded] = hist[slot][real] * (float)weight       */
  MRX_SFR_ADD = 3        /* sfr[padded] += hist[slot][real] * (float)weight      */
};
typedef struct mrx_sfr_op {
  int kind;
  int slot;
  double weight;
} mrx_sfr_op;

/* Per-snapshot plan and tables. */
typedef struct mrx_ts_snap {
  int nsteps;
  int homogeneous; /* zp >= ReionMaxHeatingRedshift: RECFAST branch (:189-262) */
  double zp, dzp, dt_dzp;
  double R[MRX_TS_MAX_STEPS];        /* Mpc */
  double zpp_edge[MRX_TS_MAX_STEPS];
  int n_ops[MRX_TS_MAX_STEPS];
  mrx_sfr_op ops[MRX_TS_MAX_STEPS][MRX_TS_MAX_SHELL_OPS];
  int ops_overflow; /* a shell needed more than MRX_TS_MAX_SHELL_OPS operations */
  int max_slot; /* largest history slot read */
  double final_weight; /* `weight` after the shell loop, reused by the Ly-n kink fix */
  /* after the shell pass */
  double collapse_fraction, x_e_ave;
  int no_light;
  double freq_heat[MRX_X_INT_NXHII][MRX_TS_MAX_STEPS];
  double freq_ion[MRX_X_INT_NXHII][MRX_TS_MAX_STEPS];
  double freq_lya[MRX_X_INT_NXHII][MRX_TS_MAX_STEPS];
  double sum_lyn[MRX_TS_MAX_STEPS];
  double nu_lower[MRX_TS_MAX_STEPS]; /* lower frequency limit: max(nu(tau_X = 1), NuXrayGalThreshold) */
  double const_zp_prefactor, growth_factor_zp, dgrowth_factor_dzp;
  /* per-shell factors of evolveInt's z'' sum (XRayHeatingFunctions.c:1289-1311) */
  double zpp[MRX_TS_MAX_STEPS], dzpp[MRX_TS_MAX_STEPS], dt_dzpp[MRX_TS_MAX_STEPS];
  double zpp_pow[MRX_TS_MAX_STEPS]; /* (1+z'')^(-SpecIndexXrayGal) */
} mrx_ts_snap;

/* Loads the tables from `dir` and sets the cosmology; 0 or a negative code. */
int mrx_ts_data_init(mrx_ts_data* d, const mrx_ts_params* p);

/* Shell radii, light-cone edges and SFR-history operations (needs no grids);
 * MRX_EUNSUPPORTED when a shell needs more than MRX_TS_MAX_SHELL_OPS operations. */
int mrx_ts_plan(mrx_ts_snap* s, const mrx_ts_data* d, const mrx_ts_params* p);

/* Frequency-integral tables, sum_lyn and prefactors, once collapse_fraction and
 * x_e_ave of the first shell are known; writes stored_fcoll[snapshot]. */
int mrx_ts_build_tables(mrx_ts_snap* s, const mrx_ts_data* d, const mrx_ts_params* p);

/* dt/dz of the flat LCDM cosmology (XRayHeatingFunctions.c:377-392) */
double mrx_cosmo_dtdz(double omega_l, double omega_m, double hubble_h, float z);

/* scalar physics, exposed for the tests */
double mrx_ts_dtdz(const mrx_ts_data* d, float z);
double mrx_ts_drdz(const mrx_ts_data* d, float z);
double mrx_ts_hubble(const mrx_ts_data* d, float z);
double mrx_ts_dicke(const mrx_ts_data* d, double z);
double mrx_ts_T_RECFAST(const mrx_ts_data* d, float z);
double mrx_ts_xion_RECFAST(const mrx_ts_data* d, float z);
double mrx_ts_kappa_10(const mrx_ts_data* d, double TK);
double mrx_ts_kappa_10_elec(const mrx_ts_data* d, double T);
double mrx_ts_kappa_10_pH(const mrx_ts_data* d, double T);
double mrx_ts_qag61(double (*f)(double, void*), void* par, double a, double b, double epsrel, int limit,
                    double* abserr);
double mrx_ts_brent(double (*f)(double, void*), void* par, double lo, double hi, double relerr, int max_iter);
double mrx_ts_integrate_over_nu(const mrx_ts_data* d, const mrx_ts_params* p, double zp, double local_x_e,
                                double lower_int_limit, int flag);
double mrx_ts_tauX(const mrx_ts_data* d, const mrx_ts_params* p, double nu, double x_e, double x_e_ave, double zp,
                   double zpp);
double mrx_ts_nu_tau_one(const mrx_ts_data* d, const mrx_ts_params* p, double zp, double zpp, double x_e,
                         double x_e_ave);
double mrx_ts_spectral_emissivity(const mrx_ts_data* d, double nu_norm);
float mrx_ts_interp_xint(const mrx_ts_data* d, const float (*tab)[MRX_X_INT_NENERGY], float En, float xHII,
                         float below);
int mrx_ts_locate_xHII_index(const mrx_ts_data* d, float xHII_call);
double mrx_ts_alpha_A(double T);

#ifdef __cplusplus
}
#endif

#endif
