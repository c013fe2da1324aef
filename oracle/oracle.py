"""ctypes front end of the C oracle (liboracle.so, built by oracle/Makefile).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
benchmark's CPU-baseline leg, always as the checker.  The product never
imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from meraxes_amd.abi import GridParams, HiiGlobals, HiiParams, TbParams, TbVelParams, TsGlobals, TsParams, fptr

from . import ts_host

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        vp, i, d, f = C.c_void_p, C.c_int, C.c_double, C.c_float
        L.orc_find_hii_bubbles.argtypes = [i, C.POINTER(HiiParams)] + [vp] * 9 + [C.POINTER(HiiGlobals), vp]
        L.orc_filtered_field_f32.argtypes = [i, d, i, vp, d, i, i, vp]
        L.orc_filtered_field_f64.argtypes = [i, d, i, vp, d, i, i, vp]
        L.orc_r2c3d_f32.argtypes = [i, vp, vp]
        L.orc_c2r3d_f32.argtypes = [i, vp, vp]
        L.orc_r2c3d_f64.argtypes = [i, vp, vp]
        L.orc_c2r3d_f64.argtypes = [i, vp, vp]
        L.orc_fft1d_f32.argtypes = [vp, vp, i, i]
        L.orc_fft1d_f64.argtypes = [vp, vp, i, i]
        L.orc_filter_f32.argtypes = [vp, i, i, i, f, i, d]
        L.orc_filter_f64.argtypes = [vp, i, d, i, d]
        L.orc_window_f32.argtypes = [i, i, i, i, f, i, d]
        L.orc_window_f32.restype = f
        L.orc_radius_ladder.argtypes = [C.POINTER(HiiParams), vp, i]
        L.orc_brightness_temp.argtypes = [C.POINTER(TbParams), vp, vp, vp, vp, C.POINTER(C.c_double)]
        L.orc_compute_ts.argtypes = [C.POINTER(TsParams), vp, d, d] + [vp] * 8 + [C.POINTER(TsGlobals), vp]
        L.orc_ts_first_shell_means.argtypes = [C.POINTER(TsParams), d, vp, vp, vp, vp, vp]
        L.orc_construct_baryon_grids.argtypes = [C.POINTER(GridParams), C.c_long] + [vp] * 11
        L.orc_brightness_temp_vel.argtypes = [C.POINTER(TbVelParams)] + [vp] * 5 + [C.POINTER(C.c_double)]
        L.orc_rsd_los.argtypes = [C.POINTER(TbVelParams), vp, vp, vp]
        L.orc_find_hii_bubbles_rec.argtypes = [i, C.POINTER(HiiParams)] + [vp] * 11 + [C.POINTER(HiiGlobals), vp]
        L.orc_set_threads.argtypes = [i]
        L.orc_get_threads.restype = i
        L.orc_set_gk61.argtypes = [vp, vp, vp]
        L.orc_qag61.restype = d
        L.orc_qag61.argtypes = [vp, vp, d, d, d, i, vp]
        # the 61-point Gauss-Kronrod rule of oracle/ts_host.py, for recomb.c's QAG
        _rule = [np.ascontiguousarray(a, np.float64) for a in (ts_host.XGK, ts_host.WGK, ts_host.WG)]
        L.orc_set_gk61(*[a.ctypes.data_as(vp) for a in _rule])
        _lib = L
    return _lib


def _chk(rc, what):
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed with {rc}")


def set_threads(n):
    lib().orc_set_threads(int(n))


def get_threads():
    return lib().orc_get_threads()


def fft1d(x, sign=-1):
    x = np.ascontiguousarray(x)
    if x.dtype == np.complex64:
        out = np.empty_like(x)
        lib().orc_fft1d_f32(fptr(x), fptr(out), x.size, sign)
    else:
        x = x.astype(np.complex128)
        out = np.empty_like(x)
        lib().orc_fft1d_f64(fptr(x), fptr(out), x.size, sign)
    return out


def r2c3d(padded):
    """padded real (N,N,2(N/2+1)) -> complex (N,N,N/2+1); float32 or float64."""
    N = padded.shape[0]
    if padded.dtype == np.float32:
        out = np.empty((N, N, N // 2 + 1), np.complex64)
        lib().orc_r2c3d_f32(N, fptr(np.ascontiguousarray(padded)), fptr(out))
    else:
        padded = np.ascontiguousarray(padded, dtype=np.float64)
        out = np.empty((N, N, N // 2 + 1), np.complex128)
        lib().orc_r2c3d_f64(N, fptr(padded), fptr(out))
    return out


def c2r3d(herm):
    N = herm.shape[0]
    if herm.dtype == np.complex64:
        out = np.empty((N, N, 2 * (N // 2 + 1)), np.float32)
        lib().orc_c2r3d_f32(N, fptr(np.ascontiguousarray(herm)), fptr(out))
    else:
        herm = np.ascontiguousarray(herm, dtype=np.complex128)
        out = np.empty((N, N, 2 * (N // 2 + 1)), np.float64)
        lib().orc_c2r3d_f64(N, fptr(herm), fptr(out))
    return out


def filter_f32(herm, R, filter_type, box_size, local_ix_start=0):
    herm = np.ascontiguousarray(herm, dtype=np.complex64).copy()
    N = herm.shape[1]
    _chk(lib().orc_filter_f32(fptr(herm), local_ix_start, herm.shape[0], N, R, filter_type, box_size), "filter")
    return herm


def filter_f64(herm, R, filter_type, box_size):
    herm = np.ascontiguousarray(herm, dtype=np.complex128).copy()
    _chk(lib().orc_filter_f64(fptr(herm), herm.shape[0], R, filter_type, box_size), "filter64")
    return herm


def window_f32(nx, ny, nz, N, R, filter_type, box_size):
    return lib().orc_window_f32(nx, ny, nz, N, R, filter_type, box_size)


def radius_ladder(p):
    buf = np.zeros(4096, np.float64)
    n = lib().orc_radius_ladder(C.byref(p), fptr(buf), 4096)
    _chk(0 if n > 0 else n, "ladder")
    return buf[:n].copy()


def filtered_field(N, box_size, filter_type, field_padded, R, apply_filter=1, clamp_kind=-1, precision=32):
    field_padded = np.ascontiguousarray(field_padded, dtype=np.float32)
    if precision == 64:
        out = np.empty((N, N, 2 * (N // 2 + 1)), np.float64)
        rc = lib().orc_filtered_field_f64(N, box_size, filter_type, fptr(field_padded), R, apply_filter,
                                          clamp_kind, fptr(out))
    else:
        out = np.empty((N, N, 2 * (N // 2 + 1)), np.float32)
        rc = lib().orc_filtered_field_f32(N, box_size, filter_type, fptr(field_padded), R, apply_filter,
                                          clamp_kind, fptr(out))
    _chk(rc, "filtered_field")
    return out


def find_hii_bubbles(p, deltax, stars, wsfr, x_e=None, z_at_ion=None, J21_at_ion=None, precision=32,
                     want_ratio=False, N_rec=None, Gamma12=None):
    """Runs the restated _find_HII_bubbles; returns a dict of unpadded outputs.

    With p.include_recombinations, N_rec (padded) and Gamma12 (unpadded) are
    copied, updated and returned under the same names.
    """
    N = p.grid_dim
    if p.include_recombinations:
        return _find_hii_bubbles_rec(p, deltax, stars, wsfr, x_e, z_at_ion, J21_at_ion, precision, want_ratio,
                                     N_rec, Gamma12)
    shp = (N, N, N)
    xH = np.empty(shp, np.float32)
    rb = np.empty(shp, np.float32)
    J21 = np.empty(shp, np.float32) if p.uvb_flag else None
    z_in = (np.full(shp, -1.0, np.float32) if z_at_ion is None else np.array(z_at_ion, np.float32, copy=True))
    J21a = None
    if p.uvb_flag:
        J21a = np.zeros(shp, np.float32) if J21_at_ion is None else np.array(J21_at_ion, np.float32, copy=True)
    ratio = np.empty(shp, np.float32) if want_ratio else None
    g = HiiGlobals()
    args = [np.ascontiguousarray(a, dtype=np.float32) if a is not None else None for a in (deltax, stars, wsfr, x_e)]
    rc = lib().orc_find_hii_bubbles(precision, C.byref(p), *[fptr(a) for a in args], fptr(xH), fptr(rb), fptr(J21),
                                    fptr(z_in), fptr(J21a), C.byref(g), fptr(ratio))
    _chk(rc, "find_hii_bubbles")
    return {
        "xH": xH, "r_bubble": rb, "J_21": J21, "z_at_ionization": z_in, "J_21_at_ionization": J21a,
        "max_ratio": ratio, "volume_weighted_global_xH": g.volume_weighted_global_xH,
        "mass_weighted_global_xH": g.mass_weighted_global_xH,
        "volume_weighted_global_J_21": g.volume_weighted_global_J_21, "n_radii": g.n_radii,
    }


def _find_hii_bubbles_rec(p, deltax, stars, wsfr, x_e, z_at_ion, J21_at_ion, precision, want_ratio, N_rec, Gamma12):
    N = p.grid_dim
    shp = (N, N, N)
    xH = np.empty(shp, np.float32)
    rb = np.empty(shp, np.float32)
    J21 = np.empty(shp, np.float32) if p.uvb_flag else None
    z_in = (np.full(shp, -1.0, np.float32) if z_at_ion is None else np.array(z_at_ion, np.float32, copy=True))
    J21a = None
    if p.uvb_flag:
        J21a = np.zeros(shp, np.float32) if J21_at_ion is None else np.array(J21_at_ion, np.float32, copy=True)
    nrec = np.array(N_rec, np.float32, copy=True)
    g12 = np.zeros(shp, np.float32) if Gamma12 is None else np.array(Gamma12, np.float32, copy=True)
    ratio = np.empty(shp, np.float32) if want_ratio else None
    g = HiiGlobals()
    args = [np.ascontiguousarray(a, dtype=np.float32) if a is not None else None for a in (deltax, stars, wsfr, x_e)]
    rc = lib().orc_find_hii_bubbles_rec(precision, C.byref(p), *[fptr(a) for a in args], fptr(nrec), fptr(xH),
                                        fptr(rb), fptr(J21), fptr(z_in), fptr(J21a), fptr(g12), C.byref(g),
                                        fptr(ratio))
    _chk(rc, "find_hii_bubbles_rec")
    return {
        "xH": xH, "r_bubble": rb, "J_21": J21, "z_at_ionization": z_in, "J_21_at_ionization": J21a,
        "N_rec": nrec, "Gamma12": g12, "max_ratio": ratio,
        "volume_weighted_global_xH": g.volume_weighted_global_xH,
        "mass_weighted_global_xH": g.mass_weighted_global_xH,
        "volume_weighted_global_J_21": g.volume_weighted_global_J_21, "n_radii": g.n_radii,
    }


def rr_lib():
    """The oracle's recombination functions (oracle/recomb.c), for the tests."""
    L = lib()
    d = C.c_double
    sig = {"orc_rr_rate": [d, d], "orc_rr_A": [d], "orc_rr_C": [d], "orc_rr_beta": [d],
           "orc_rr_Gamma_SS": [d, d, d, d], "orc_rr_neutral_fraction": [d, d, d, d],
           "orc_rr_recombination_rate": [d, d, d], "orc_rr_A_aux_integral": [d], "orc_rr_table": [C.c_int, C.c_int],
           "orc_rr_lnG": [C.c_int], "orc_dtdz": [d, d, d, C.c_float]}
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = d
    L.orc_rr_build.argtypes = [d, d, d, d]
    L.orc_rr_build.restype = C.c_int
    return L


def brightness_temp(p, xH, deltax, TS_box=None):
    N = p.grid_dim
    out = np.empty((N, N, N), np.float32)
    ave = C.c_double()
    args = [np.ascontiguousarray(a, dtype=np.float32) if a is not None else None for a in (xH, deltax, TS_box)]
    _chk(lib().orc_brightness_temp(C.byref(p), *[fptr(a) for a in args], fptr(out), C.byref(ave)), "tb")
    return out, ave.value


def brightness_temp_vel(p, xH, deltax, TS_box, vel):
    """delta_T with peculiar velocities; returns (delta_T, <delta_T>, converted vel) - vel is copied."""
    N = p.tb.grid_dim
    out = np.empty((N, N, N), np.float32)
    ave = C.c_double()
    v = np.array(vel, np.float32, copy=True)
    args = [np.ascontiguousarray(a, dtype=np.float32) for a in (xH, deltax, TS_box)]
    _chk(lib().orc_brightness_temp_vel(C.byref(p), *[fptr(a) for a in args], fptr(v), fptr(out), C.byref(ave)),
         "tb_vel")
    return out, ave.value, v


def rsd_los(p, xH, vel_converted, delta_T):
    """The sub-cell re-binning alone on a delta_T grid (returns a new grid)."""
    out = np.array(delta_T, np.float32, copy=True)
    _chk(lib().orc_rsd_los(C.byref(p), fptr(np.ascontiguousarray(xH, np.float32)),
                           fptr(np.ascontiguousarray(vel_converted, np.float32)), fptr(out)), "rsd_los")
    return out


class _Spline(C.Structure):
    _fields_ = [("n", C.c_int), ("x", C.c_double * 32), ("y", C.c_double * 32), ("c", C.c_double * 32)]


_MS, _MO, _NX = 64, 4096, 14


class TsTab(C.Structure):
    """ctypes mirror of orc_ts_tab (oracle.h)."""
    _fields_ = [("nsteps", C.c_int), ("homogeneous", C.c_int), ("no_light", C.c_int), ("n_ops", C.c_int),
                ("zp", C.c_double), ("dzp", C.c_double), ("dt_dzp", C.c_double), ("const_zp_prefactor", C.c_double),
                ("growth_factor_zp", C.c_double), ("dgrowth_factor_dzp", C.c_double),
                ("No", C.c_double), ("N_b0", C.c_double), ("f_H", C.c_double), ("f_He", C.c_double),
                ("xe_recfast", C.c_double), ("T_recfast", C.c_double),
                ("R", C.c_double * _MS), ("dzpp", C.c_double * _MS), ("dt_dzpp", C.c_double * _MS),
                ("zpp", C.c_double * _MS), ("zpp_pow", C.c_double * _MS), ("sum_lyn", C.c_double * _MS),
                ("freq_heat", (C.c_double * _MS) * _NX), ("freq_ion", (C.c_double * _MS) * _NX),
                ("freq_lya", (C.c_double * _MS) * _NX), ("xHII", C.c_float * _NX),
                ("kappa_HH", _Spline), ("kappa_eH", _Spline), ("kappa_pH", _Spline),
                ("op_shell", C.c_int * _MO), ("op_kind", C.c_int * _MO), ("op_slot", C.c_int * _MO),
                ("op_weight", C.c_double * _MO)]


def _fill_spline(dst, sp):
    dst.n = len(sp.x)
    for i in range(dst.n):
        dst.x[i], dst.y[i], dst.c[i] = sp.x[i], sp.y[i], sp.c[i]


def ts_tables(p, deltax=None, stars=None, x_e_box_prev=None):
    """The snapshot's host tables from the oracle's restatement (oracle/ts_host.py).

    The two grid means the tables depend on (ComputeTs.c:413-475) are taken
    from the grids (or from p.fixed_* when p.fixed_globals is set); like the
    reference, stored_fcoll[snapshot] receives the collapse fraction.
    Returns (TsSnapshot, orc_ts_tab, collapse_fraction, x_e_ave)."""
    h = ts_host.host(p)
    s = h.plan()
    cf = xe = 0.0
    if deltax is not None:
        R0 = ts_host.L_FACTOR * (p.box_size / p.hubble_h) / float(np.float32(p.grid_dim))
        cfv, xev = C.c_double(), C.c_double()
        args = [np.ascontiguousarray(a, dtype=np.float32) for a in (deltax, stars, x_e_box_prev)]
        _chk(lib().orc_ts_first_shell_means(C.byref(p), R0, *[fptr(a) for a in args], C.byref(cfv), C.byref(xev)),
             "ts_first_shell_means")
        cf, xe = cfv.value, (0.0 if s.homogeneous else xev.value)
    if p.fixed_globals and not s.homogeneous:
        cf, xe = p.fixed_collapse_fraction, p.fixed_x_e_ave
    h.tables(s, cf, xe, p.stored_fcoll)
    if not s.homogeneous and p.stored_fcoll:
        p.stored_fcoll[p.snapshot] = cf
    t = TsTab()
    t.nsteps, t.homogeneous = s.nsteps, int(s.homogeneous)
    t.zp, t.dzp, t.dt_dzp = s.zp, s.dzp, s.dt_dzp
    t.No, t.N_b0, t.f_H, t.f_He = h.No, h.N_b0, h.f_H, h.f_He
    for k in range(ts_host.NXHII):
        t.xHII[k] = float(ts_host.XHII[k])
    _fill_spline(t.kappa_HH, h.d.kappa_HH)
    _fill_spline(t.kappa_eH, h.d.kappa_eH)
    _fill_spline(t.kappa_pH, h.d.kappa_pH)
    if s.homogeneous:
        t.xe_recfast, t.T_recfast = s.xe_recfast, s.T_recfast
        return s, t, cf, xe
    if len(s.ops) > _MO:
        raise ValueError("more SFR-history operations than the oracle holds")
    t.no_light = s.no_light
    t.const_zp_prefactor, t.growth_factor_zp, t.dgrowth_factor_dzp = (s.const_zp_prefactor, s.growth_factor_zp,
                                                                      s.dgrowth_factor_dzp)
    for c in range(s.nsteps):
        t.R[c], t.dzpp[c], t.dt_dzpp[c], t.zpp[c] = s.R[c], s.dzpp[c], s.dt_dzpp[c], s.zpp[c]
        t.zpp_pow[c], t.sum_lyn[c] = s.zpp_pow[c], s.sum_lyn[c]
        for k in range(ts_host.NXHII):
            t.freq_heat[k][c], t.freq_ion[k][c], t.freq_lya[k][c] = (s.freq_heat[k, c], s.freq_ion[k, c],
                                                                     s.freq_lya[k, c])
    t.n_ops = len(s.ops)
    for k, (c, kind, slot, w) in enumerate(s.ops):
        t.op_shell[k], t.op_kind[k], t.op_slot[k], t.op_weight[k] = c, kind, slot, w
    return s, t, cf, xe


def compute_ts(p, deltax, stars, sfr, sfr_histories, x_e_box_prev, Tk_box, want_shells=False):
    """Runs the restated _ComputeTs; x_e_box_prev / Tk_box are copied, not modified."""
    N = p.grid_dim
    xe_prev = np.array(x_e_box_prev, np.float32, copy=True)
    Tk = np.array(Tk_box, np.float32, copy=True)
    xe = np.empty_like(xe_prev)
    TS = np.empty((N, N, N), np.float32)
    shells = np.empty((p.num_filter_steps, N, N, N), np.float32) if want_shells else None
    _, t, cf, xav = ts_tables(p, deltax, stars, xe_prev)
    g = TsGlobals()
    args = [np.ascontiguousarray(a, dtype=np.float32) if a is not None else None
            for a in (deltax, stars, sfr, sfr_histories)]
    _chk(lib().orc_compute_ts(C.byref(p), C.byref(t), cf, xav, *[fptr(a) for a in args], fptr(xe), fptr(xe_prev),
                              fptr(Tk), fptr(TS), C.byref(g), fptr(shells)), "compute_ts")
    out = {k: getattr(g, k) for k, _ in TsGlobals._fields_}
    out.update({"x_e_box": xe, "x_e_box_prev": xe_prev, "Tk_box": Tk, "TS_box": TS, "shells": shells})
    return out


def _spline_dict(sp):
    """Knots and per-interval coefficients (y + dx (b + dx (c + dx d))) of a natural spline."""
    x, y, c = np.asarray(sp.x), np.asarray(sp.y), np.asarray(sp.c)
    h = np.diff(x)
    b = np.zeros_like(x)
    d = np.zeros_like(x)
    b[:-1] = (y[1:] - y[:-1]) / h - h * (c[1:] + 2.0 * c[:-1]) / 3.0
    d[:-1] = (c[1:] - c[:-1]) / (3.0 * h)
    return {"x": x.copy(), "y": y.copy(), "b": b, "c": c.copy(), "d": d}


def ts_export(p):
    """Constants of the inhomogeneous spin-temperature step (p.fixed_globals must be set)."""
    if not p.fixed_globals:
        raise ValueError("ts_export needs p.fixed_globals")
    s, _, _, _ = ts_tables(p)
    if s.homogeneous:
        raise ValueError("ts_export: homogeneous snapshot")
    h = ts_host.host(p)
    out = {"nsteps": s.nsteps, "no_light": s.no_light, "zp": s.zp, "dzp": s.dzp, "dt_dzp": s.dt_dzp,
           "const_zp_prefactor": s.const_zp_prefactor, "growth": s.growth_factor_zp,
           "dgrowth": s.dgrowth_factor_dzp, "N_b0": h.N_b0, "No": h.No, "f_H": h.f_H, "f_He": h.f_He,
           "H": h.hubble(s.zp), "hubble_h": p.hubble_h}
    for name in ("R", "dt_dzpp", "dzpp", "zpp", "zpp_pow", "sum_lyn"):
        out[name] = np.array(getattr(s, name), np.float64)
    for name in ("freq_heat", "freq_ion", "freq_lya"):
        out[name] = np.array(getattr(s, name), np.float64)
    out["XH"] = ts_host.XHII.copy()
    out["kHH"], out["keH"], out["kpH"] = (_spline_dict(h.d.kappa_HH), _spline_dict(h.d.kappa_eH),
                                          _spline_dict(h.d.kappa_pH))
    out["ops"] = list(s.ops)
    return out


def construct_baryon_grids(p, pos, type_, fesc, gross=None, bhm=None, ebhm=None, hist=None):
    """construct_baryon_grids restated (reionization.c:1353-1615) on host arrays.

    hist (spin temperature: [n_store_sfr] padded grids) is copied, shifted and
    returned.  Returns stars, weighted_sfr, sfr, hist and the counts.
    """
    N = p.grid_dim
    shape = (N, N, 2 * (N // 2 + 1))
    pos = np.ascontiguousarray(pos, np.float32)
    type_ = np.ascontiguousarray(type_, np.int32)
    fesc = np.ascontiguousarray(fesc, np.float64)
    conv = (lambda a: None if a is None else np.ascontiguousarray(a, np.float64))
    gross, bhm, ebhm = conv(gross), conv(bhm), conv(ebhm)
    stars = np.empty(shape, np.float32)
    wsfr = np.empty(shape, np.float32)
    sfr = np.empty(shape, np.float32) if p.include_spin_temp else None
    hist = None if hist is None else np.array(hist, np.float32, copy=True)
    stats = np.zeros(3, np.int64)
    rc = lib().orc_construct_baryon_grids(C.byref(p), len(pos), fptr(pos), fptr(type_), fptr(fesc), fptr(gross),
                                          fptr(bhm), fptr(ebhm), fptr(stars), fptr(wsfr), fptr(sfr), fptr(hist),
                                          fptr(stats))
    if rc <= -100:
        raise ValueError(f"galaxy {-100 - rc} has a negative grid index")
    _chk(rc, "construct_baryon_grids")
    return {"stars": stars, "weighted_sfr": wsfr, "sfr": sfr, "sfr_histories": hist, "n_gridded": int(stats[0]),
            "n_dead": int(stats[1]), "n_bh_below_limit": int(stats[2])}
