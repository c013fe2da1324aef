/*
 * A stand-in for Meraxes' src/meraxes.h, for compiling dropin/meraxes_hip_dropin.c
 * outside a Meraxes tree (tests/test_dropin.py).  It declares only what the
 * shim reads, with the names and C types of the reference's declarations
 * (run_params_t, physics_params_t, run_units_t, reion_grids_t, galaxy_t,
 * gal_to_slab_t, run_globals_t; mlog_error, mlog, ABORT).  Member order and
 * everything else differ: inside a Meraxes tree the real header is used.
 */
#ifndef DROPIN_STUB_MERAXES_H
#define DROPIN_STUB_MERAXES_H

#include <stddef.h>
#include <stdlib.h>

#define STRLEN 256
#define MLOG_MESG 0

void mlog(const char* fmt, int type, ...);
void mlog_error(const char* fmt, ...);
void stub_abort(int sigterm);
#define ABORT(sigterm) stub_abort(sigterm)

enum tree_ids { VELOCIRAPTOR_TREES, GBPTREES_TREES, VELOCIRAPTOR_TREES_AUG };

typedef struct physics_params_t {
  double ReionEfficiency, ReionNionPhotPerBary, Y_He, BlackHoleMassLimitReion;
  double LXrayGal, NuXrayGalThreshold, SpecIndexXrayGal, NuXraySoftCut, NuXrayMax;
  double ReionMaxHeatingRedshift, ReionGammaHaloBias, ReionAlphaUV, ReionAlphaUVBH;
  double ReionRBubbleMin, ReionRBubbleMax, ReionRBubbleMaxRecomb;
  int Flag_BHFeedback;
} physics_params_t;

typedef struct run_params_t {
  char TablesForXHeatingDir[STRLEN];
  physics_params_t physics;
  double BoxSize, Hubble_h, BaryonFrac, OmegaM, OmegaK, OmegaR, OmegaLambda, wLambda;
  double ReionDeltaRFactor, ReionSfrTimescale;
  int ReionGridDim, ReionFilterType, TsHeatingFilterType, ReionRtoMFilterType, ReionUVBFlag;
  enum tree_ids TreesID;
  int Flag_IncludeSpinTemp, Flag_IncludeRecombinations, Flag_IncludePecVelsFor21cm, TsNumFilterSteps;
} run_params_t;

typedef struct run_units_t {
  double UnitTime_in_s, UnitLength_in_cm, UnitMass_in_g;
} run_units_t;

typedef struct galaxy_t {
  double GrossStellarMass, FescWeightedGSM, BlackHoleMass, EffectiveBHM;
  int Type;
  float Pos[3];
} galaxy_t;

typedef struct gal_to_slab_t {
  int index;
  galaxy_t* galaxy;
  int slab_ind;
} gal_to_slab_t;

typedef struct reion_grids_t {
  ptrdiff_t* slab_nix;
  float *stars, *deltax, *sfr, *sfr_histories, *weighted_sfr;
  float *xH, *z_at_ionization, *J_21_at_ionization, *J_21, *r_bubble;
  float *x_e_box, *x_e_box_prev, *Tk_box, *TS_box, *N_rec, *Gamma12, *delta_T, *vel;
  struct gal_to_slab_t* galaxy_to_slab_map;
  double volume_weighted_global_xH, volume_weighted_global_J_21, mass_weighted_global_xH;
  double volume_ave_J_alpha, volume_ave_xalpha, volume_ave_Xheat, volume_ave_Xion;
  double volume_ave_TS, volume_ave_TK, volume_ave_xe, volume_ave_Tb;
} reion_grids_t;

typedef struct run_globals_t {
  run_params_t params;
  reion_grids_t reion_grids;
  struct run_units_t units;
  int mpi_rank;
  double* ZZ;
  double Hubble, RhoCrit;
  int NstoreSnapshots_SFR;
} run_globals_t;

extern run_globals_t run_globals;

#endif
