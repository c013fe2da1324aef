/*
 * mrx_io.h - C ABI of the reionization grid files (libmrx_io.so).
 *
 * The reference writes one HDF5 file per snapshot, "<OutputDir>/<FileNameGalaxies>_grids_<snapshot>.hdf5"
 * (src/core/reionization.c:1640-1646): save_reion_input_grids (:1648-1754)
 * creates it with the input grids, save_reion_output_grids (:1786-2050) opens it
 * again and adds the output grids and their volume averages as attributes.
 * Every grid is an N x N x N float32 dataset (H5T_NATIVE_FLOAT) chunked
 * {1, N, N} along x (:1682, :1827), unpadded whatever the memory layout
 * (padded grids are de-padded, :1688-1726, :1864-1870); every average is
 * a one-element double attribute (H5LTset_attribute_double).
 *
 * These entry points write the same files from grids held in HBM (or on the
 * host, pointers_on_device = 0): a kernel de-pads and converts units on the
 * device, the unpadded slabs are copied to pinned host memory in x-slabs of a
 * few MiB, and the host writes slab s while slab s+1 is copied.  Datasets,
 * shapes, chunking, types, values and attributes equal the reference's; the
 * file is written by one process with the default (serial) file driver where
 * the reference's ranks write collectively through MPI-IO (the bytes of the
 * datasets do not depend on the driver).  Not written: the light cone
 * (Flag_ConstructLightcone, :1880-1925) and the mini-halo grids (USE_MINI_HALOS).
 *
 * Return codes are those of mrx_reion.h; MRX_EIO when HDF5 fails (file
 * missing, directory not writable, dataset already present), with the
 * operation in mrx_last_error().  HDF5's own error printing is suppressed for
 * the duration of each call and restored afterwards.
 */
#ifndef MRX_IO_H
#define MRX_IO_H

#include <stddef.h>

#include "mrx_reion.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Run flags and units the two writers read (run_globals.params, units). */
typedef struct mrx_save_params {
  int grid_dim;                 /* ReionGridDim                                                    */
  int uvb_flag;                 /* ReionUVBFlag: J_21 (+ its average), J_21_at_ionization, Mvir_crit */
  int include_spin_temp;        /* Flag_IncludeSpinTemp: sfr (input file); TS_box, Tk_box, x_e_box   */
  int compute_21cm_bright_temp; /* Flag_Compute21cmBrightTemp: delta_T                               */
  int compute_ps;               /* Flag_ComputePS: k_bins, PS_data, PS_error (:1966-2040)           */
  int ps_length;                /* PS_Length                                                       */
  int pointers_on_device;       /* grids in device memory (1) or host memory (0); PS arrays are host */
  int pad_;                     /* unused, 0                                                       */
  double unit_mass_g;           /* units.UnitMass_in_g  } sfr and weighted_sfr are written in       */
  double unit_time_s;           /* units.UnitTime_in_s  } M_sun/yr (:1691-1693, :1723-1725)          */
} mrx_save_params;

/* The output grids and averages of one snapshot (reion_grids_t, src/meraxes.h:366-523).
 * Unpadded grids (INDEX_REAL) except x_e_box_prev (INDEX_PADDED).  A grid whose
 * flag is off is not read and may be NULL. */
typedef struct mrx_output_grids {
  const float* xH;
  const float* z_at_ionization;
  const float* r_bubble;
  const float* J_21;               /* uvb_flag */
  const float* J_21_at_ionization; /* uvb_flag */
  const float* Mvir_crit;          /* uvb_flag */
  const float* TS_box;             /* include_spin_temp */
  const float* Tk_box;             /* include_spin_temp */
  const float* x_e_box_prev;       /* include_spin_temp, padded: written as "x_e_box" (:1864-1870) */
  const float* delta_T;            /* compute_21cm_bright_temp */
  const float* PS_k;               /* compute_ps, ps_length floats, host memory */
  const float* PS_data;
  const float* PS_error;
  double volume_weighted_global_xH;   /* attributes of "xH" (:1927-1928) */
  double mass_weighted_global_xH;
  double volume_weighted_global_J_21; /* of "J_21" (:1836) */
  double volume_ave_TS;               /* of "TS_box" (:1930-1945) */
  double volume_ave_TK;               /* of "Tk_box" */
  double volume_ave_xe;               /* of "x_e_box" */
  double volume_ave_J_alpha;          /* of "TS_box" */
  double volume_ave_xalpha;
  double volume_ave_Xheat;
  double volume_ave_Xion;
  double volume_ave_Tb;               /* of "delta_T" (:1959) */
} mrx_output_grids;

/* gen_grids_fname (reionization.c:1640-1646): "<output_dir>/<file_name_galaxies>_grids_<snapshot>.hdf5",
 * or without the directory when relative != 0.  Returns the length written
 * (excluding the terminator) or MRX_EINVAL when it does not fit in len bytes. */
int mrx_grids_fname(char* buf, size_t len, const char* output_dir, const char* file_name_galaxies, int snapshot,
                    int relative);

/* save_reion_input_grids (reionization.c:1648-1754): creates (truncates) the file
 * at path and writes "sfr" (include_spin_temp only, in M_sun/yr), "deltax",
 * "stars" and "weighted_sfr" (in M_sun/yr), all from padded grids. */
int mrx_save_reion_input_grids(mrx_ctx* ctx, const mrx_save_params* p, const char* path, const float* deltax,
                               const float* stars, const float* weighted_sfr, const float* sfr);

/* save_reion_output_grids (reionization.c:1786-2050): opens the existing file
 * at path for writing and adds, in the reference's order, "xH",
 * "z_at_ionization", "r_bubble"; with uvb_flag "J_21" (+ volume_weighted_global_J_21),
 * "J_21_at_ionization", "Mvir_crit"; with include_spin_temp "TS_box", "Tk_box",
 * "x_e_box"; with compute_21cm_bright_temp "delta_T"; the averages as
 * attributes; with compute_ps "k_bins", "PS_data", "PS_error" ({ps_length}, contiguous). */
int mrx_save_reion_output_grids(mrx_ctx* ctx, const mrx_save_params* p, const char* path,
                                const mrx_output_grids* g);

/* sizeof of this header's structs as compiled into the library: 0 mrx_save_params,
 * 1 mrx_output_grids (0 for any other value). */
size_t mrx_io_abi_sizeof(int which);

#ifdef __cplusplus
}
#endif
#endif /* MRX_IO_H */
