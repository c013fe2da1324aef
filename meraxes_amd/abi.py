"""ctypes mirror of include/mrx_reion.h (structs and constants only).

The struct layouts here must match the header field for field; the C-ABI
tests compare ``ctypes.sizeof`` against ``mrx_abi_sizeof`` exported by the
library, so a drift fails loudly.
"""
import ctypes as C

MRX_OK = 0
MRX_EINVAL = -1
MRX_ENOMEM = -2
MRX_EDEVICE = -3
MRX_EUNSUPPORTED = -4
MRX_EFILTER = -5
MRX_EIO = -6

ERRORS = {
    MRX_EINVAL: "MRX_EINVAL",
    MRX_ENOMEM: "MRX_ENOMEM",
    MRX_EDEVICE: "MRX_EDEVICE",
    MRX_EUNSUPPORTED: "MRX_EUNSUPPORTED",
    MRX_EFILTER: "MRX_EFILTER",
    MRX_EIO: "MRX_EIO",
}

FILTER_TOPHAT = 0
FILTER_SHARPK = 1
FILTER_GAUSS = 2


class HiiParams(C.Structure):
    """mrx_hii_params: knobs of _find_HII_bubbles (find_HII_bubbles.c:47-495)."""

    _fields_ = [
        ("grid_dim", C.c_int),
        ("filter_type", C.c_int),
        ("rtom_filter_type", C.c_int),
        ("uvb_flag", C.c_int),
        ("include_spin_temp", C.c_int),
        ("pointers_on_device", C.c_int),
        ("box_size", C.c_double),
        ("redshift", C.c_double),
        ("r_bubble_max", C.c_double),
        ("r_bubble_min", C.c_double),
        ("delta_r_factor", C.c_double),
        ("efficiency", C.c_double),
        ("omega_m", C.c_double),
        ("rho_crit", C.c_double),
        ("alpha_uv", C.c_double),
        ("nion_phot_per_bary", C.c_double),
        ("gamma_halo_bias", C.c_double),
        ("unit_length_cm", C.c_double),
        ("unit_mass_g", C.c_double),
        ("unit_time_s", C.c_double),
        ("include_recombinations", C.c_int),
        ("pad_", C.c_int),
        ("prev_redshift", C.c_double),
        ("hubble_h", C.c_double),
        ("omega_lambda", C.c_double),
        ("baryon_frac", C.c_double),
        ("y_he", C.c_double),
    ]


class HiiGlobals(C.Structure):
    """mrx_hii_globals: the scalars _find_HII_bubbles leaves in reion_grids."""

    _fields_ = [
        ("volume_weighted_global_xH", C.c_double),
        ("mass_weighted_global_xH", C.c_double),
        ("volume_weighted_global_J_21", C.c_double),
        ("n_radii", C.c_int),
    ]


class TbParams(C.Structure):
    """mrx_tb_params: knobs of ComputeBrightnessTemperatureBox."""

    _fields_ = [
        ("grid_dim", C.c_int),
        ("include_spin_temp", C.c_int),
        ("pointers_on_device", C.c_int),
        ("redshift", C.c_double),
        ("hubble_h", C.c_double),
        ("omega_m", C.c_double),
        ("baryon_frac", C.c_double),
    ]


class TsParams(C.Structure):
    """mrx_ts_params: knobs of _ComputeTs (ComputeTs.c:78-1020)."""

    _fields_ = [
        ("grid_dim", C.c_int),
        ("pointers_on_device", C.c_int),
        ("num_filter_steps", C.c_int),
        ("heating_filter_type", C.c_int),
        ("rtom_filter_type", C.c_int),
        ("snapshot", C.c_int),
        ("n_history", C.c_int),
        ("reproduce_quirks", C.c_int),
        ("box_size", C.c_double),
        ("hubble_h", C.c_double),
        ("omega_m", C.c_double),
        ("omega_lambda", C.c_double),
        ("omega_r", C.c_double),
        ("omega_k", C.c_double),
        ("w_lambda", C.c_double),
        ("baryon_frac", C.c_double),
        ("y_he", C.c_double),
        ("l_xray_gal", C.c_double),
        ("nu_xray_gal_threshold", C.c_double),
        ("spec_index_xray_gal", C.c_double),
        ("nu_xray_soft_cut", C.c_double),
        ("nu_xray_max", C.c_double),
        ("reion_efficiency", C.c_double),
        ("nion_phot_per_bary", C.c_double),
        ("max_heating_redshift", C.c_double),
        ("rho_crit", C.c_double),
        ("unit_length_cm", C.c_double),
        ("unit_mass_g", C.c_double),
        ("unit_time_s", C.c_double),
        ("redshifts", C.POINTER(C.c_double)),
        ("stored_fcoll", C.POINTER(C.c_double)),
        ("tables_dir", C.c_char_p),
        ("fixed_globals", C.c_int),
        ("fixed_collapse_fraction", C.c_double),
        ("fixed_x_e_ave", C.c_double),
    ]


class TsGlobals(C.Structure):
    """mrx_ts_globals: the scalars _ComputeTs leaves in reion_grids."""

    _fields_ = [
        ("volume_ave_J_alpha", C.c_double),
        ("volume_ave_xalpha", C.c_double),
        ("volume_ave_Xheat", C.c_double),
        ("volume_ave_Xion", C.c_double),
        ("volume_ave_TS", C.c_double),
        ("volume_ave_TK", C.c_double),
        ("volume_ave_xe", C.c_double),
        ("collapse_fraction", C.c_double),
        ("x_e_ave", C.c_double),
        ("no_light", C.c_int),
        ("homogeneous", C.c_int),
    ]


def fptr(a):
    """float* of a numpy array or an int device address (None -> NULL)."""
    if a is None:
        return None
    if isinstance(a, int):
        return C.c_void_p(a)
    return a.ctypes.data_as(C.c_void_p)


MRX_BATCH_NO_GRAPH = 1
MRX_BATCH_NO_TUNE = 2


class HiiBatchBox(C.Structure):
    """mrx_hii_batch_box: one box of a batch (device grids only)."""

    _fields_ = [
        ("params", HiiParams),
        ("deltax", C.c_void_p),
        ("stars", C.c_void_p),
        ("weighted_sfr", C.c_void_p),
        ("x_e", C.c_void_p),
        ("xH", C.c_void_p),
        ("r_bubble", C.c_void_p),
        ("J_21", C.c_void_p),
        ("z_at_ionization", C.c_void_p),
        ("J_21_at_ionization", C.c_void_p),
    ]


class GridParams(C.Structure):
    """mrx_grid_params: knobs of construct_baryon_grids (reionization.c:1353-1615)."""

    _fields_ = [
        ("grid_dim", C.c_int),
        ("include_spin_temp", C.c_int),
        ("bh_feedback", C.c_int),
        ("n_store_sfr", C.c_int),
        ("pointers_on_device", C.c_int),
        ("pad_", C.c_int),
        ("box_size", C.c_double),
        ("reion_sfr_timescale", C.c_double),
        ("hubble", C.c_double),
        ("omega_m", C.c_double),
        ("omega_k", C.c_double),
        ("omega_lambda", C.c_double),
        ("redshift", C.c_double),
        ("bh_mass_limit_reion", C.c_double),
        ("alpha_uv_bh", C.c_double),
        ("alpha_uv", C.c_double),
    ]


class GridStats(C.Structure):
    """mrx_grid_stats: galaxy counts of one gridding call and the SFR timescale used."""

    _fields_ = [
        ("n_gridded", C.c_long),
        ("n_dead", C.c_long),
        ("n_bh_below_limit", C.c_long),
        ("sfr_timescale", C.c_double),
    ]


class TbVelParams(C.Structure):
    """mrx_tb_vel_params: ComputeBrightnessTemperatureBox with peculiar velocities."""

    _fields_ = [
        ("tb", TbParams),
        ("pec_vels", C.c_int),
        ("trees_id", C.c_int),
        ("box_size", C.c_double),
        ("omega_lambda", C.c_double),
        ("omega_r", C.c_double),
    ]


_TSM, _TSO = 64, 1024  # MRX_TS_TABLES_MAX_STEPS, MRX_TS_TABLES_MAX_OPS


class TsTables(C.Structure):
    """mrx_ts_tables: the host tables of one spin-temperature snapshot."""

    _fields_ = [
        ("nsteps", C.c_int), ("homogeneous", C.c_int), ("no_light", C.c_int), ("n_ops", C.c_int),
        ("zp", C.c_double), ("dzp", C.c_double), ("dt_dzp", C.c_double), ("const_zp_prefactor", C.c_double),
        ("growth_factor_zp", C.c_double), ("dgrowth_factor_dzp", C.c_double),
        ("R", C.c_double * _TSM), ("zpp_edge", C.c_double * _TSM), ("zpp", C.c_double * _TSM),
        ("dzpp", C.c_double * _TSM), ("dt_dzpp", C.c_double * _TSM), ("zpp_pow", C.c_double * _TSM),
        ("sum_lyn", C.c_double * _TSM), ("nu_lower", C.c_double * _TSM),
        ("freq_heat", (C.c_double * _TSM) * 14), ("freq_ion", (C.c_double * _TSM) * 14),
        ("freq_lya", (C.c_double * _TSM) * 14),
        ("op_shell", C.c_int * _TSO), ("op_kind", C.c_int * _TSO), ("op_slot", C.c_int * _TSO),
        ("op_weight", C.c_double * _TSO),
    ]


class SaveParams(C.Structure):
    """mrx_save_params (include/mrx_io.h): flags and units of the grid files."""

    _fields_ = [
        ("grid_dim", C.c_int),
        ("uvb_flag", C.c_int),
        ("include_spin_temp", C.c_int),
        ("compute_21cm_bright_temp", C.c_int),
        ("compute_ps", C.c_int),
        ("ps_length", C.c_int),
        ("pointers_on_device", C.c_int),
        ("pad_", C.c_int),
        ("unit_mass_g", C.c_double),
        ("unit_time_s", C.c_double),
    ]


OUTPUT_GRID_NAMES = ("xH", "z_at_ionization", "r_bubble", "J_21", "J_21_at_ionization", "Mvir_crit", "TS_box",
                     "Tk_box", "x_e_box_prev", "delta_T", "PS_k", "PS_data", "PS_error")
OUTPUT_AVERAGES = ("volume_weighted_global_xH", "mass_weighted_global_xH", "volume_weighted_global_J_21",
                   "volume_ave_TS", "volume_ave_TK", "volume_ave_xe", "volume_ave_J_alpha", "volume_ave_xalpha",
                   "volume_ave_Xheat", "volume_ave_Xion", "volume_ave_Tb")


class OutputGrids(C.Structure):
    """mrx_output_grids (include/mrx_io.h): the output grids and averages of one snapshot."""

    _fields_ = [(n, C.c_void_p) for n in OUTPUT_GRID_NAMES] + [(n, C.c_double) for n in OUTPUT_AVERAGES]
