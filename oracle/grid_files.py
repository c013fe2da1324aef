"""Oracle of the reionization grid files: what each file must hold.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker; the product
never imports it.

A restatement in numpy of save_reion_input_grids and save_reion_output_grids
(src/core/reionization.c:1648-1754, :1786-2050) for one process holding the
whole box: each grid becomes an N^3 float32 dataset chunked (1, N, N)
(:1680-1683, :1825-1828), padded grids de-padded (grid_index INDEX_PADDED ->
INDEX_REAL, :1688-1726, :1864-1870), sfr and weighted_sfr converted to M_sun/yr
as (float)(v * UnitMass_in_g / UnitTime_in_s * SEC_PER_YEAR / SOLAR_MASS)
(:1691-1693, :1723-1725), and each average a one-element double attribute
(H5LTset_attribute_double, :1836, :1927-1959).  The power spectrum vectors
are contiguous {PS_Length} float datasets (:1966-2040).
"""
import numpy as np

SEC_PER_YEAR = 3.155e7  # src/meraxes.h:33
SOLAR_MASS = 1.989e33  # src/meraxes.h:22


def unpad(a, N):
    """INDEX_PADDED -> INDEX_REAL (misc_tools.c:200-220): drop the two padding floats of each row."""
    return np.ascontiguousarray(np.asarray(a, np.float32).reshape(N, N, N + 2)[:, :, :N])


def msun_per_yr(a, N, unit_mass_g, unit_time_s):
    """(float)(v * UnitMass_in_g / UnitTime_in_s * SEC_PER_YEAR / SOLAR_MASS), left to right in double."""
    v = unpad(a, N).astype(np.float64)
    return (((v * unit_mass_g) / unit_time_s) * SEC_PER_YEAR / SOLAR_MASS).astype(np.float32)


def input_file(p, deltax, stars, weighted_sfr, sfr=None):
    """{name: float32 (N, N, N)} of save_reion_input_grids (:1648-1754), in the order written."""
    N = p.grid_dim
    out = {}
    if p.include_spin_temp:
        out["sfr"] = msun_per_yr(sfr, N, p.unit_mass_g, p.unit_time_s)
    out["deltax"] = unpad(deltax, N)
    out["stars"] = unpad(stars, N)
    out["weighted_sfr"] = msun_per_yr(weighted_sfr, N, p.unit_mass_g, p.unit_time_s)
    return out


def output_file(p, grids, averages):
    """(datasets, attributes) of save_reion_output_grids (:1786-2050).

    datasets: {name: float32 array} in the order written; attributes:
    {dataset name: {attribute name: value}}.
    """
    N = p.grid_dim
    ds, at = {}, {}

    def real(name):
        return np.ascontiguousarray(np.asarray(grids[name], np.float32).reshape(N, N, N))

    ds["xH"] = real("xH")
    ds["z_at_ionization"] = real("z_at_ionization")
    ds["r_bubble"] = real("r_bubble")
    if p.uvb_flag:  # :1834-1845
        ds["J_21"] = real("J_21")
        at.setdefault("J_21", {})["volume_weighted_global_J_21"] = averages["volume_weighted_global_J_21"]
        ds["J_21_at_ionization"] = real("J_21_at_ionization")
        ds["Mvir_crit"] = real("Mvir_crit")
    if p.include_spin_temp:  # :1856-1871
        ds["TS_box"] = real("TS_box")
        ds["Tk_box"] = real("Tk_box")
        ds["x_e_box"] = unpad(grids["x_e_box_prev"], N)
    if p.compute_21cm_bright_temp:  # :1873-1878
        ds["delta_T"] = real("delta_T")
    at.setdefault("xH", {})["volume_weighted_global_xH"] = averages["volume_weighted_global_xH"]
    at["xH"]["mass_weighted_global_xH"] = averages["mass_weighted_global_xH"]
    if p.include_spin_temp:  # :1930-1950
        at.setdefault("TS_box", {})["volume_ave_TS"] = averages["volume_ave_TS"]
        at.setdefault("Tk_box", {})["volume_ave_TK"] = averages["volume_ave_TK"]
        at.setdefault("x_e_box", {})["volume_ave_xe"] = averages["volume_ave_xe"]
        for k in ("volume_ave_J_alpha", "volume_ave_xalpha", "volume_ave_Xheat", "volume_ave_Xion"):
            at["TS_box"][k] = averages[k]
    if p.compute_21cm_bright_temp:
        at.setdefault("delta_T", {})["volume_ave_Tb"] = averages["volume_ave_Tb"]
    if p.compute_ps:  # :1966-2040
        L = p.ps_length
        ds["k_bins"] = np.asarray(grids["PS_k"], np.float32)[:L]
        ds["PS_data"] = np.asarray(grids["PS_data"], np.float32)[:L]
        ds["PS_error"] = np.asarray(grids["PS_error"], np.float32)[:L]
    return ds, at
