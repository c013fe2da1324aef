/*
 * meraxes_hip_dropin.c - the MI355X library behind Meraxes' own GPU dispatch.
 *
 * Compiled inside a Meraxes tree (next to src/core/find_HII_bubbles.c) with
 * -DUSE_CUDA (or -DUSE_HIP, which defines USE_CUDA for the dispatch sites) and
 * linked with meraxes_amd/libmrx_hip.so.  It defines the two symbols Meraxes'
 * GPU build expects, with their signatures (src/meraxes_gpu.h:53-54):
 *   init_CUDA()                          called once from init_gpu() (init.c:25-43)
 *   _find_HII_bubbles_gpu(snapshot, flag) dispatched at find_HII_bubbles.c:503-513
 * and GPU versions of the other two stages of the snapshot, for the dispatch
 * points a maintainer adds where ComputeTs() (ComputeTs.c:1024-1045) and
 * ComputeBrightnessTemperatureBox (BrightnessTemperature.h:9) are called:
 *   _ComputeTs_gpu(snapshot), ComputeBrightnessTemperatureBox_gpu(snapshot),
 *   construct_baryon_grids_gpu(snapshot, local_ngals) (reionization.c:1353).
 * Each unpacks run_globals into one call of the C ABI (include/mrx_reion.h).
 * The grids stay in Meraxes' host arrays (pointers_on_device = 0): the library
 * stages them through its device buffers.  One rank holds the whole box on its
 * GPU; a slab-decomposed run is refused (the CPU path serves it).  Errors are
 * reported the reference's way: mlog_error, then ABORT(EXIT_FAILURE).
 */
#ifdef USE_HIP
#ifndef USE_CUDA
#define USE_CUDA
#endif
#endif

#include <stdbool.h>
#include <stdlib.h>
#include <string.h>

#include "meraxes.h"
#include "mrx_reion.h"

static mrx_ctx* g_mrx = NULL;
/* The collapse-fraction history the spin-temperature tables interpolate in
 * (the reference's stored_fcoll[], XRayHeatingFunctions.h): kept here, since
 * every spin-temperature step of a GPU build goes through this file. */
static double g_stored_fcoll[1000];

static void mrx_check(int rc, const char* what)
{
  if (rc != MRX_OK) {
    mlog_error("MI355X path, %s: %s (code %d)", what, mrx_last_error(g_mrx), rc);
    ABORT(EXIT_FAILURE);
  }
}

static void require_whole_box(void)
{
  const int N = run_globals.params.ReionGridDim;
  if (run_globals.reion_grids.slab_nix[run_globals.mpi_rank] != N) {
    mlog_error("MI355X path: rank %d holds %ld of %d x-planes; the library needs the whole box on one GPU",
               run_globals.mpi_rank, (long)run_globals.reion_grids.slab_nix[run_globals.mpi_rank], N);
    ABORT(EXIT_FAILURE);
  }
}

void init_CUDA(void)
{
  const int ndev = mrx_device_count();
  if (ndev <= 0) {
    mlog_error("MI355X path: no GPU visible");
    ABORT(EXIT_FAILURE);
  }
  mrx_check(mrx_init(run_globals.mpi_rank % ndev, &g_mrx), "mrx_init");
}

void _find_HII_bubbles_gpu(const int snapshot, const bool flag_write_validation_output)
{
  (void)flag_write_validation_output;
  run_params_t* P = &run_globals.params;
  reion_grids_t* G = &run_globals.reion_grids;
  require_whole_box();
  mrx_hii_params p;
  memset(&p, 0, sizeof p);
  p.grid_dim = P->ReionGridDim;
  p.filter_type = P->ReionFilterType;
  p.rtom_filter_type = P->ReionRtoMFilterType;
  p.uvb_flag = P->ReionUVBFlag;
  p.include_spin_temp = P->Flag_IncludeSpinTemp;
  p.pointers_on_device = 0;
  p.box_size = P->BoxSize;
  p.redshift = run_globals.ZZ[snapshot];
  p.r_bubble_max = P->physics.ReionRBubbleMax;
  p.r_bubble_min = P->physics.ReionRBubbleMin;
  p.delta_r_factor = P->ReionDeltaRFactor;
  p.efficiency = P->physics.ReionEfficiency;
  p.omega_m = P->OmegaM;
  p.rho_crit = run_globals.RhoCrit;
  p.alpha_uv = P->physics.ReionAlphaUV;
  p.nion_phot_per_bary = P->physics.ReionNionPhotPerBary;
  p.gamma_halo_bias = P->physics.ReionGammaHaloBias;
  p.unit_length_cm = run_globals.units.UnitLength_in_cm;
  p.unit_mass_g = run_globals.units.UnitMass_in_g;
  p.unit_time_s = run_globals.units.UnitTime_in_s;
  /* inhomogeneous recombinations (find_HII_bubbles.c:83-90, :186-190) */
  p.include_recombinations = P->Flag_IncludeRecombinations;
  if (p.include_recombinations) p.r_bubble_max = P->physics.ReionRBubbleMaxRecomb;
  p.prev_redshift = run_globals.ZZ[snapshot > 0 ? snapshot - 1 : 0];
  p.hubble_h = P->Hubble_h;
  p.omega_lambda = P->OmegaLambda;
  p.baryon_frac = P->BaryonFrac;
  p.y_he = P->physics.Y_He;
  mrx_hii_globals g;
  mrx_check(mrx_find_hii_bubbles_rec(g_mrx, &p, G->deltax, G->stars, G->weighted_sfr,
                                     P->Flag_IncludeSpinTemp ? G->x_e_box : NULL,
                                     P->Flag_IncludeRecombinations ? G->N_rec : NULL, G->xH, G->r_bubble, G->J_21,
                                     G->z_at_ionization, G->J_21_at_ionization,
                                     P->Flag_IncludeRecombinations ? G->Gamma12 : NULL, &g),
            "find_HII_bubbles");
  G->volume_weighted_global_xH = g.volume_weighted_global_xH;
  G->mass_weighted_global_xH = g.mass_weighted_global_xH;
  G->volume_weighted_global_J_21 = g.volume_weighted_global_J_21;
}

void _ComputeTs_gpu(int snapshot)
{
  run_params_t* P = &run_globals.params;
  reion_grids_t* G = &run_globals.reion_grids;
  require_whole_box();
  mrx_ts_params p;
  memset(&p, 0, sizeof p);
  p.grid_dim = P->ReionGridDim;
  p.pointers_on_device = 0;
  p.num_filter_steps = P->TsNumFilterSteps;
  p.heating_filter_type = P->TsHeatingFilterType;
  p.rtom_filter_type = P->ReionRtoMFilterType;
  p.snapshot = snapshot;
  p.n_history = run_globals.NstoreSnapshots_SFR;
  p.reproduce_quirks = 1; /* volume_ave_xe as the reference computes it (ComputeTs.c:957) */
  p.box_size = P->BoxSize;
  p.hubble_h = P->Hubble_h;
  p.omega_m = P->OmegaM;
  p.omega_lambda = P->OmegaLambda;
  p.omega_r = P->OmegaR;
  p.omega_k = P->OmegaK;
  p.w_lambda = P->wLambda;
  p.baryon_frac = P->BaryonFrac;
  p.y_he = P->physics.Y_He;
  p.l_xray_gal = P->physics.LXrayGal;
  p.nu_xray_gal_threshold = P->physics.NuXrayGalThreshold;
  p.spec_index_xray_gal = P->physics.SpecIndexXrayGal;
  p.nu_xray_soft_cut = P->physics.NuXraySoftCut;
  p.nu_xray_max = P->physics.NuXrayMax;
  p.reion_efficiency = P->physics.ReionEfficiency;
  p.nion_phot_per_bary = P->physics.ReionNionPhotPerBary;
  p.max_heating_redshift = P->physics.ReionMaxHeatingRedshift;
  p.rho_crit = run_globals.RhoCrit;
  p.unit_length_cm = run_globals.units.UnitLength_in_cm;
  p.unit_mass_g = run_globals.units.UnitMass_in_g;
  p.unit_time_s = run_globals.units.UnitTime_in_s;
  p.redshifts = run_globals.ZZ;
  p.stored_fcoll = g_stored_fcoll;
  p.tables_dir = P->TablesForXHeatingDir;
  mrx_ts_globals g;
  mrx_check(mrx_compute_ts(g_mrx, &p, G->deltax, G->stars, G->sfr, G->sfr_histories, G->x_e_box, G->x_e_box_prev,
                           G->Tk_box, G->TS_box, &g),
            "ComputeTs");
  G->volume_ave_J_alpha = g.volume_ave_J_alpha;
  G->volume_ave_xalpha = g.volume_ave_xalpha;
  G->volume_ave_Xheat = g.volume_ave_Xheat;
  G->volume_ave_Xion = g.volume_ave_Xion;
  G->volume_ave_TS = g.volume_ave_TS;
  G->volume_ave_TK = g.volume_ave_TK;
  G->volume_ave_xe = g.volume_ave_xe;
}

/* ComputeBrightnessTemperatureBox (BrightnessTemperature.c:24), both branches
 * of Flag_IncludePecVelsFor21cm. */
void ComputeBrightnessTemperatureBox_gpu(int snapshot)
{
  run_params_t* P = &run_globals.params;
  reion_grids_t* G = &run_globals.reion_grids;
  require_whole_box();
  mrx_tb_params tb;
  memset(&tb, 0, sizeof tb);
  tb.grid_dim = P->ReionGridDim;
  tb.include_spin_temp = P->Flag_IncludeSpinTemp;
  tb.pointers_on_device = 0;
  tb.redshift = run_globals.ZZ[snapshot];
  tb.hubble_h = P->Hubble_h;
  tb.omega_m = P->OmegaM;
  tb.baryon_frac = P->BaryonFrac;
  double ave = 0.0;
  if (P->Flag_IncludePecVelsFor21cm > 0 && P->Flag_IncludeSpinTemp) {
    mrx_tb_vel_params v;
    memset(&v, 0, sizeof v);
    v.tb = tb;
    v.pec_vels = P->Flag_IncludePecVelsFor21cm;
    v.trees_id = (int)P->TreesID;
    v.box_size = P->BoxSize;
    v.omega_lambda = P->OmegaLambda;
    v.omega_r = P->OmegaR;
    mrx_check(mrx_brightness_temp_vel(g_mrx, &v, G->xH, G->deltax, G->TS_box, G->vel, G->delta_T, &ave),
              "brightness temperature (velocities)");
  } else {
    mrx_check(mrx_brightness_temp(g_mrx, &tb, G->xH, G->deltax, G->TS_box, G->delta_T, &ave),
              "brightness temperature");
  }
  G->volume_ave_Tb = ave;
}

/* construct_baryon_grids (reionization.c:1353): the fields of the galaxies in
 * galaxy_to_slab_map, gathered into arrays in the order the reference visits them. */
void construct_baryon_grids_gpu(int snapshot, int local_ngals)
{
  run_params_t* P = &run_globals.params;
  reion_grids_t* G = &run_globals.reion_grids;
  require_whole_box();
  gal_to_slab_t* map = G->galaxy_to_slab_map;
  long n = 0;
  for (long live = 0; live < local_ngals; n++)
    if (map[n].galaxy->Type <= 2) live++;
  float* pos = (float*)malloc(sizeof(float) * 3 * (size_t)(n + 1));
  int* type = (int*)malloc(sizeof(int) * (size_t)(n + 1));
  double* f = (double*)malloc(sizeof(double) * 4 * (size_t)(n + 1));
  if (!pos || !type || !f) {
    mlog_error("MI355X path: out of host memory gathering %ld galaxies", n);
    ABORT(EXIT_FAILURE);
  }
  for (long i = 0; i < n; i++) {
    const galaxy_t* g = map[i].galaxy;
    memcpy(pos + 3 * i, g->Pos, sizeof(float) * 3);
    type[i] = g->Type;
    f[i] = g->FescWeightedGSM;
    f[n + i] = g->GrossStellarMass;
    f[2 * n + i] = g->BlackHoleMass;
    f[3 * n + i] = g->EffectiveBHM;
  }
  mrx_grid_params p;
  memset(&p, 0, sizeof p);
  p.grid_dim = P->ReionGridDim;
  p.include_spin_temp = P->Flag_IncludeSpinTemp;
  p.bh_feedback = P->physics.Flag_BHFeedback;
  p.n_store_sfr = run_globals.NstoreSnapshots_SFR;
  p.pointers_on_device = 0;
  p.box_size = P->BoxSize;
  p.reion_sfr_timescale = P->ReionSfrTimescale;
  p.hubble = run_globals.Hubble;
  p.omega_m = P->OmegaM;
  p.omega_k = P->OmegaK;
  p.omega_lambda = P->OmegaLambda;
  p.redshift = run_globals.ZZ[snapshot];
  p.bh_mass_limit_reion = P->physics.BlackHoleMassLimitReion;
  p.alpha_uv_bh = P->physics.ReionAlphaUVBH;
  p.alpha_uv = P->physics.ReionAlphaUV;
  mrx_grid_stats s;
  mrx_check(mrx_construct_baryon_grids(g_mrx, &p, n, pos, type, f, f + n, f + 2 * n, f + 3 * n, G->stars,
                                       G->weighted_sfr, G->sfr, G->sfr_histories, &s),
            "construct_baryon_grids");
  free(pos);
  free(type);
  free(f);
}
