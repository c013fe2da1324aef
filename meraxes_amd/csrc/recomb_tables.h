/*
 * recomb_tables.h - host tables of the inhomogeneous recombinations
 * (src/core/recombinations.c, after 21cmFAST): the MHR00 density-PDF
 * parameters A, C, beta (:159-300) and the recombination rate per baryon
 * RR(z, ln Gamma12) integrated over that PDF with self-shielding (:115-156,
 * init_MHR :56-82), splined in ln Gamma12 at 300 redshifts.  Plain C, shared
 * by the HIP library and the oracle; compile with -ffp-contract=off.
 */
#ifndef MRX_RECOMB_TABLES_H
#define MRX_RECOMB_TABLES_H

#include "ts_tables.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MRX_RR_Z_NPTS 300     /* RR_Z_NPTS, recombinations.h */
#define MRX_RR_DEL_Z 0.2f     /* RR_DEL_Z (a float) */
#define MRX_RR_LNG_NPTS 150   /* RR_lnGamma_NPTS */
#define MRX_RR_LNG_MIN (-10.0) /* RR_lnGamma_min */
#define MRX_RR_DEL_LNG 0.1f   /* RR_DEL_lnGamma (a float) */

typedef struct mrx_rr_tables {
  double No;   /* present hydrogen number density [cm^-3] (reionization.h:25-27) */
  double y_he; /* Y_He */
  mrx_spline A, C, beta;
  double lnG[MRX_RR_LNG_NPTS]; /* knots: RR_lnGamma_min + k * RR_DEL_lnGamma (the product in float) */
  /* natural cubic spline of each redshift row: value and coefficients at the knots */
  double y[MRX_RR_Z_NPTS][MRX_RR_LNG_NPTS], b[MRX_RR_Z_NPTS][MRX_RR_LNG_NPTS];
  double c[MRX_RR_Z_NPTS][MRX_RR_LNG_NPTS], d[MRX_RR_Z_NPTS][MRX_RR_LNG_NPTS];
} mrx_rr_tables;

/* No of the cosmology (reionization.h:17-27). */
double mrx_rr_No(double hubble_h, double omega_m, double baryon_frac, double y_he);

/* Builds every table (the 45,000 rate integrals on up to 16 host threads;
 * each value is computed as the serial loop would). 0 on success. */
int mrx_rr_init(mrx_rr_tables* t, double No, double y_he);

/* splined_recombination_rate (recombinations.c:30-54). */
double mrx_rr_rate(const mrx_rr_tables* t, float z_eff, double gamma12_bg);

/* pieces, exposed for the tests */
double mrx_rr_A_MHR(const mrx_rr_tables* t, double z);
double mrx_rr_C_MHR(const mrx_rr_tables* t, double z);
double mrx_rr_beta_MHR(const mrx_rr_tables* t, double z);
double mrx_rr_Gamma_SS(double Gamma_bg, double Delta, double T_4, double z);
double mrx_rr_neutral_fraction(double y_he, double density, double T4, double gamma12, int usecaseB);
double mrx_rr_recombination_rate(const mrx_rr_tables* t, double z_eff, double gamma12_bg, double T4, int usecaseB);
double mrx_rr_A_aux_integral(const mrx_rr_tables* t, double z);

#ifdef __cplusplus
}
#endif

#endif
