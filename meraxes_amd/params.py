"""Physical parameters and units of the measured configurations.

Values follow the reference's own parameter files: cosmology and units of
input/params/simulations/Tiamat.par:8-28, reionization knobs of
input/params/defaults.par:196-254, and the unit derivations of set_units
(src/core/init.c:184-205).
"""
import math

import ctypes as C
import os

import numpy as np

from .abi import GridParams, HiiParams, SaveParams, TbParams, TbVelParams, TsParams

# src/meraxes.h:21-44
GRAVITY = 6.672e-8
SOLAR_MASS = 1.989e33
SEC_PER_YEAR = 3.155e7
HUBBLE = 3.2407789e-18
MPC = 3.086e24
TCMB = 2.728

# Tiamat.par
HUBBLE_H = 0.678
OMEGA_M = 0.308
OMEGA_L = 0.692
BARYON_FRAC = 0.15714
Y_HE = 0.24
UNIT_LENGTH_CM = 3.08568e24  # Mpc/h
UNIT_MASS_G = 1.989e43  # 1e10 Msun/h
UNIT_VELOCITY_CM_S = 1e5  # km/s


def units():
    """UnitTime_in_s, G and RhoCrit as set_units computes them (init.c:184-205)."""
    unit_time_s = UNIT_LENGTH_CM / UNIT_VELOCITY_CM_S
    G = GRAVITY / UNIT_LENGTH_CM**3 * UNIT_MASS_G * unit_time_s**2
    hubble = HUBBLE * unit_time_s
    rho_crit = 3 * hubble * hubble / (8 * math.pi * G)
    return {"unit_time_s": unit_time_s, "G": G, "rho_crit": rho_crit}


def hii_params(grid_dim, box_size, redshift, filter_type=0, efficiency=20.0, uvb_flag=0,
               include_spin_temp=0, r_bubble_max=50.0, r_bubble_min=0.4068, delta_r_factor=1.1,
               rtom_filter_type=0, on_device=0, include_recombinations=0, prev_redshift=None):
    """mrx_hii_params for one snapshot (defaults.par:221-251 for the unset knobs)."""
    u = units()
    p = HiiParams()
    p.grid_dim = grid_dim
    p.filter_type = filter_type
    p.rtom_filter_type = rtom_filter_type
    p.uvb_flag = uvb_flag
    p.include_spin_temp = include_spin_temp
    p.pointers_on_device = on_device
    p.box_size = box_size
    p.redshift = redshift
    p.r_bubble_max = r_bubble_max
    p.r_bubble_min = r_bubble_min
    p.delta_r_factor = delta_r_factor
    p.efficiency = efficiency
    p.omega_m = OMEGA_M
    p.rho_crit = u["rho_crit"]
    p.alpha_uv = 5.0
    p.nion_phot_per_bary = 4000.0
    p.gamma_halo_bias = 2.0
    p.unit_length_cm = UNIT_LENGTH_CM
    p.unit_mass_g = UNIT_MASS_G
    p.unit_time_s = u["unit_time_s"]
    # recombinations (defaults.par:253 ReionRBubbleMaxRecomb = 33.9 when on)
    p.include_recombinations = include_recombinations
    p.prev_redshift = redshift if prev_redshift is None else prev_redshift
    p.hubble_h = HUBBLE_H
    p.omega_lambda = OMEGA_L
    p.baryon_frac = BARYON_FRAC
    p.y_he = Y_HE
    return p


def grid_params(grid_dim, box_size, redshift, include_spin_temp=0, n_store_sfr=8, bh_feedback=1,
                reion_sfr_timescale=0.5, bh_mass_limit_reion=-1.0, alpha_uv_bh=1.57, alpha_uv=5.0, on_device=0):
    """mrx_grid_params (defaults.par:183-191, :227-228, :248 for the knobs; Tiamat.par cosmology)."""
    u = units()
    p = GridParams()
    p.grid_dim = grid_dim
    p.include_spin_temp = include_spin_temp
    p.bh_feedback = bh_feedback
    p.n_store_sfr = n_store_sfr
    p.pointers_on_device = on_device
    p.box_size = box_size
    p.reion_sfr_timescale = reion_sfr_timescale
    p.hubble = HUBBLE * u["unit_time_s"]  # run_globals.Hubble (init.c:201)
    p.omega_m = OMEGA_M
    p.omega_k = 0.0
    p.omega_lambda = OMEGA_L
    p.redshift = redshift
    p.bh_mass_limit_reion = bh_mass_limit_reion
    p.alpha_uv_bh = alpha_uv_bh
    p.alpha_uv = alpha_uv
    return p


def tb_params(grid_dim, redshift, include_spin_temp=0, on_device=0):
    p = TbParams()
    p.grid_dim = grid_dim
    p.include_spin_temp = include_spin_temp
    p.pointers_on_device = on_device
    p.redshift = redshift
    p.hubble_h = HUBBLE_H
    p.omega_m = OMEGA_M
    p.baryon_frac = BARYON_FRAC
    return p


def tb_vel_params(grid_dim, box_size, redshift, pec_vels=3, trees_id=1, on_device=0):
    """mrx_tb_vel_params: the spin-temperature delta_T with peculiar velocities (Tiamat cosmology)."""
    p = TbVelParams()
    p.tb = tb_params(grid_dim, redshift, include_spin_temp=1, on_device=on_device)
    p.pec_vels = pec_vels
    p.trees_id = trees_id
    p.box_size = box_size
    p.omega_lambda = OMEGA_L
    p.omega_r = 0.0
    return p


def save_params(grid_dim, uvb_flag=1, include_spin_temp=1, compute_21cm_bright_temp=1, compute_ps=0, ps_length=0,
                on_device=0):
    """mrx_save_params: flags of the grid files, Tiamat units (UnitTime = UnitLength / UnitVelocity)."""
    p = SaveParams()
    p.grid_dim = grid_dim
    p.uvb_flag = uvb_flag
    p.include_spin_temp = include_spin_temp
    p.compute_21cm_bright_temp = compute_21cm_bright_temp
    p.compute_ps = compute_ps
    p.ps_length = ps_length
    p.pointers_on_device = on_device
    p.unit_mass_g = UNIT_MASS_G
    p.unit_time_s = units()["unit_time_s"]
    return p


def radius_ladder(p):
    """Radii of the bubble finder, largest first (find_HII_bubbles.c:184-207, :445)."""
    N = p.grid_dim
    clf = 0.620350491
    if p.box_size / N < 1.0:
        clf = 1.0
    R = min(p.r_bubble_max, 0.620350491 * p.box_size)
    out = []
    while True:
        last = (R / p.delta_r_factor <= clf * p.box_size / N) or (R / p.delta_r_factor <= p.r_bubble_min)
        if last:
            R = clf * p.box_size / N
        out.append(R)
        if last:
            return out
        R /= p.delta_r_factor


TABLES_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "21cmfast_tables")


def ts_params(grid_dim, box_size, snapshot, redshifts, stored_fcoll, n_history, efficiency=20.0,
              max_heating_redshift=20.0, num_filter_steps=40, heating_filter_type=1, on_device=0,
              reproduce_quirks=0):
    """mrx_ts_params with the defaults.par X-ray knobs (defaults.par:48-55, :247-254).

    ``redshifts`` and ``stored_fcoll`` are float64 numpy arrays the struct points
    into; keep them alive while the struct is used.
    """
    u = units()
    p = TsParams()
    p.grid_dim = grid_dim
    p.pointers_on_device = on_device
    p.num_filter_steps = num_filter_steps
    p.heating_filter_type = heating_filter_type
    p.rtom_filter_type = 0
    p.snapshot = snapshot
    p.n_history = n_history
    p.reproduce_quirks = reproduce_quirks
    p.box_size = box_size
    p.hubble_h = HUBBLE_H
    p.omega_m = OMEGA_M
    p.omega_lambda = OMEGA_L
    p.omega_r = 0.0
    p.omega_k = 0.0
    p.w_lambda = -1.0
    p.baryon_frac = BARYON_FRAC
    p.y_he = Y_HE
    p.l_xray_gal = 3.16e40
    p.nu_xray_gal_threshold = 500.0
    p.spec_index_xray_gal = 1.0
    p.nu_xray_soft_cut = 2000.0
    p.nu_xray_max = 10000.0
    p.reion_efficiency = efficiency
    p.nion_phot_per_bary = 4000.0
    p.max_heating_redshift = max_heating_redshift
    p.rho_crit = u["rho_crit"]
    p.unit_length_cm = UNIT_LENGTH_CM
    p.unit_mass_g = UNIT_MASS_G
    p.unit_time_s = u["unit_time_s"]
    assert redshifts.dtype == np.float64 and stored_fcoll.dtype == np.float64
    p.redshifts = redshifts.ctypes.data_as(C.POINTER(C.c_double))
    p.stored_fcoll = stored_fcoll.ctypes.data_as(C.POINTER(C.c_double))
    p.tables_dir = TABLES_DIR.encode()
    return p


def snapshot_redshifts(n=30, z_hi=20.0, z_lo=6.0):
    """The 30 snapshot redshifts of the full-snapshot configuration, z = 20 .. 6."""
    return np.linspace(z_hi, z_lo, n).astype(np.float64)
