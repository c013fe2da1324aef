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
        L.orc_compute_ts.argtypes = [C.POINTER(TsParams)] + [vp] * 8 + [C.POINTER(TsGlobals), vp]
        L.orc_ts_export.argtypes = [C.POINTER(TsParams), vp, i]
        L.orc_construct_baryon_grids.argtypes = [C.POINTER(GridParams), C.c_long] + [vp] * 11
        L.orc_brightness_temp_vel.argtypes = [C.POINTER(TbVelParams)] + [vp] * 5 + [C.POINTER(C.c_double)]
        L.orc_rsd_los.argtypes = [C.POINTER(TbVelParams), vp, vp, vp]
        L.orc_find_hii_bubbles_rec.argtypes = [i, C.POINTER(HiiParams)] + [vp] * 11 + [C.POINTER(HiiGlobals), vp]
        L.orc_set_threads.argtypes = [i]
        L.orc_get_threads.restype = i
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


def rr_tables_lib():
    """The oracle library's recombination-table functions (recomb_tables.c), for the tests."""
    L = lib()
    d = C.c_double
    for name, args in (("mrx_rr_No", [d, d, d, d]), ("mrx_rr_neutral_fraction", [d, d, d, d, C.c_int]),
                       ("mrx_rr_Gamma_SS", [d, d, d, d]), ("mrx_cosmo_dtdz", [d, d, d, C.c_float])):
        f = getattr(L, name)
        f.argtypes = args
        f.restype = d
    for name, args in (("mrx_rr_A_MHR", [C.c_void_p, d]), ("mrx_rr_C_MHR", [C.c_void_p, d]),
                       ("mrx_rr_beta_MHR", [C.c_void_p, d]), ("mrx_rr_rate", [C.c_void_p, C.c_float, d]),
                       ("mrx_rr_recombination_rate", [C.c_void_p, d, d, d, C.c_int])):
        f = getattr(L, name)
        f.argtypes = args
        f.restype = d
    L.mrx_rr_init.argtypes = [C.c_void_p, d, d]
    L.mrx_rr_init.restype = C.c_int
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


def compute_ts(p, deltax, stars, sfr, sfr_histories, x_e_box_prev, Tk_box, want_shells=False):
    """Runs the restated _ComputeTs; x_e_box_prev / Tk_box are copied, not modified."""
    N = p.grid_dim
    xe_prev = np.array(x_e_box_prev, np.float32, copy=True)
    Tk = np.array(Tk_box, np.float32, copy=True)
    xe = np.empty_like(xe_prev)
    TS = np.empty((N, N, N), np.float32)
    shells = np.empty((p.num_filter_steps, N, N, N), np.float32) if want_shells else None
    g = TsGlobals()
    args = [np.ascontiguousarray(a, dtype=np.float32) if a is not None else None
            for a in (deltax, stars, sfr, sfr_histories)]
    _chk(lib().orc_compute_ts(C.byref(p), *[fptr(a) for a in args], fptr(xe), fptr(xe_prev), fptr(Tk), fptr(TS),
                              C.byref(g), fptr(shells)), "compute_ts")
    out = {k: getattr(g, k) for k, _ in TsGlobals._fields_}
    out.update({"x_e_box": xe, "x_e_box_prev": xe_prev, "Tk_box": Tk, "TS_box": TS, "shells": shells})
    return out


def ts_export(p):
    """Constants of the inhomogeneous spin-temperature step (p.fixed_globals must be set)."""
    buf = np.zeros(200000, np.float64)
    n = lib().orc_ts_export(C.byref(p), buf.ctypes.data_as(C.c_void_p), len(buf))
    _chk(0 if n > 0 else n, "ts_export")
    o = buf[:n]
    ns = int(o[0])
    out = {"nsteps": ns, "no_light": int(o[1]), "zp": o[2], "dzp": o[3], "dt_dzp": o[4],
           "const_zp_prefactor": o[5], "growth": o[6], "dgrowth": o[7], "N_b0": o[8], "No": o[9], "f_H": o[10],
           "f_He": o[11], "H": o[12], "hubble_h": o[13]}
    nop = int(o[14])
    k = 16
    for name in ("R", "dt_dzpp", "dzpp", "zpp", "zpp_pow", "sum_lyn"):
        out[name] = o[k:k + ns].copy()
        k += ns
    for name in ("freq_heat", "freq_ion", "freq_lya"):
        out[name] = o[k:k + 14 * ns].reshape(14, ns).copy()
        k += 14 * ns
    out["XH"] = o[k:k + 14].astype(np.float32)
    k += 14
    for name in ("kHH", "keH", "kpH"):
        m = int(o[k])
        k += 1
        sp = {}
        for f in ("x", "y", "b", "c", "d"):
            sp[f] = o[k:k + m].copy()
            k += 27
        out[name] = sp
    ops = []
    for _ in range(nop):
        ops.append((int(o[k]), int(o[k + 1]), int(o[k + 2]), float(o[k + 3])))
        k += 4
    out["ops"] = ops
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
