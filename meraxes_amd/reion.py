"""Host-side mirror of the reference's reionization interface over the C ABI.

Same names and argument meaning as the reference functions each call replaces:

* ``find_hii_bubbles``  -> _find_HII_bubbles (src/core/find_HII_bubbles.c:47-495)
* ``filter_``           -> filter (src/core/reionization.c:2110-2170)
* ``fft_r2c_3d`` / ``fft_c2r_3d`` -> the r2c / c2r plans (reionization.c:685-736)
* ``filtered_field``    -> one field of the smoothing chain (find_HII_bubbles.c:214-311)
* ``brightness_temp``   -> ComputeBrightnessTemperatureBox (BrightnessTemperature.c:24-549)
* ``compute_ts``        -> _ComputeTs (ComputeTs.c:78-1020)
* ``power_spectrum``    -> Compute_PS (ComputePowerSpectrum.c:44-255)
* ``construct_baryon_grids`` -> construct_baryon_grids (reionization.c:1353-1615)
* ``brightness_temp_vel`` -> the peculiar-velocity branch (BrightnessTemperature.c:126-510)
* ``gen_grids_fname``, ``save_reion_input_grids``, ``save_reion_output_grids``
  -> the HDF5 grid files (reionization.c:1640-2050), through libmrx_io.so
* ``HiiBatch``          -> many boxes of _find_HII_bubbles in one graph launch
  (no counterpart in the reference, which runs one box per call)
* ``power_spectrum_adjoint``, ``snapshot_adjoint`` -> backward passes of
  Compute_PS and of one snapshot (Ts -> bubbles -> delta_T), no counterpart
  in the reference; meraxes_amd.autograd wraps them as autograd functions

Grids are numpy arrays (host memory, staged by the library) or torch tensors
on the GPU (used in place, on torch's current stream).  Layouts are the
reference's: padded real (N, N, 2(N/2+1)), unpadded real (N, N, N), half
spectrum (N, N, N/2+1) complex64.  Invalid arguments raise MrxError with the
library's message, where the reference aborts (e.g. reionization.c:2160-2165).
"""
import ctypes as C
import os

import numpy as np

from . import _lib, abi
from ._lib import Context, MrxError, load  # noqa: F401
from .abi import GridStats, HiiBatchBox, HiiGlobals, TsGlobals, TsParams, TsTables

try:  # torch is optional for host-memory use
    import torch
except Exception:  # pragma: no cover
    torch = None


def _is_dev(a):
    return torch is not None and isinstance(a, torch.Tensor) and a.is_cuda


def _ptr(a):
    if a is None:
        return None
    if _is_dev(a):
        if not a.is_contiguous():
            raise ValueError("device grids must be contiguous")
        return C.c_void_p(a.data_ptr())
    if not (isinstance(a, np.ndarray) and a.flags.c_contiguous):
        raise ValueError("host grids must be C-contiguous numpy arrays")
    return a.ctypes.data_as(C.c_void_p)


def _bind_stream(ctx, on_dev):
    """Orders the library's work after torch's pending work on the current stream.

    A torch side stream is handed to the library, which then enqueues on it
    without synchronising.  torch's legacy default stream cannot be handed over
    (the library's own stream does not wait on it), so its pending work is
    waited for first; the library then synchronises its own stream before
    returning.
    """
    if on_dev:
        cur = torch.cuda.current_stream()
        if cur.cuda_stream:
            ctx.check(ctx.lib.mrx_set_stream(ctx.ptr, C.c_void_p(cur.cuda_stream)), "set_stream")
            return
        cur.synchronize()
    ctx.check(ctx.lib.mrx_set_stream(ctx.ptr, None), "set_stream")


def _empty_like_real(ref, N, fill=None):
    if _is_dev(ref):
        t = torch.empty((N, N, N), dtype=torch.float32, device=ref.device)
        if fill is not None:
            t.fill_(fill)
        return t
    out = np.empty((N, N, N), np.float32)
    if fill is not None:
        out.fill(fill)
    return out


def radius_ladder(params):
    lib = load()
    buf = np.zeros(4096, np.float64)
    n = lib.mrx_radius_ladder(C.byref(params), buf.ctypes.data_as(C.c_void_p), 4096)
    if n < 0:
        raise MrxError(n, "radius ladder")
    return buf[:n].copy()


def find_hii_bubbles(ctx, params, deltax, stars, weighted_sfr, x_e=None, z_at_ionization=None,
                     J_21_at_ionization=None, out=None, N_rec=None, Gamma12=None):
    """One call of the bubble finder; returns the output grids and globals.

    ``z_at_ionization`` / ``J_21_at_ionization`` are the state carried between
    snapshots (updated in place when given; fresh -1 / 0 grids otherwise).
    ``out`` may hold preallocated ``xH``, ``r_bubble``, ``J_21`` grids.  With
    params.include_recombinations, ``N_rec`` (padded) and ``Gamma12`` are the
    recombination state carried between snapshots, updated in place (N_rec is
    required; Gamma12 defaults to a fresh zero grid).
    """
    N = params.grid_dim
    on_dev = _is_dev(deltax)
    params.pointers_on_device = 1 if on_dev else 0
    out = dict(out or {})
    xH = out.get("xH") if out.get("xH") is not None else _empty_like_real(deltax, N)
    rb = out.get("r_bubble") if out.get("r_bubble") is not None else _empty_like_real(deltax, N)
    J21 = None
    if params.uvb_flag:
        J21 = out.get("J_21") if out.get("J_21") is not None else _empty_like_real(deltax, N)
        if J_21_at_ionization is None:
            J_21_at_ionization = _empty_like_real(deltax, N, 0.0)
    if z_at_ionization is None:
        z_at_ionization = _empty_like_real(deltax, N, -1.0)
    g = HiiGlobals()
    _bind_stream(ctx, on_dev)
    if params.include_recombinations:
        if N_rec is None:
            raise ValueError("include_recombinations needs the N_rec grid")
        if Gamma12 is None:
            Gamma12 = _empty_like_real(deltax, N, 0.0)
        ctx.check(ctx.lib.mrx_find_hii_bubbles_rec(
            ctx.ptr, C.byref(params), _ptr(deltax), _ptr(stars), _ptr(weighted_sfr), _ptr(x_e), _ptr(N_rec), _ptr(xH),
            _ptr(rb), _ptr(J21), _ptr(z_at_ionization), _ptr(J_21_at_ionization), _ptr(Gamma12), C.byref(g)),
            "find_hii_bubbles_rec")
    else:
        ctx.check(ctx.lib.mrx_find_hii_bubbles(ctx.ptr, C.byref(params), _ptr(deltax), _ptr(stars),
                                               _ptr(weighted_sfr), _ptr(x_e), _ptr(xH), _ptr(rb), _ptr(J21),
                                               _ptr(z_at_ionization), _ptr(J_21_at_ionization), C.byref(g)),
                  "find_hii_bubbles")
    return {
        "xH": xH, "r_bubble": rb, "J_21": J21, "z_at_ionization": z_at_ionization,
        "J_21_at_ionization": J_21_at_ionization if params.uvb_flag else None,
        "N_rec": N_rec, "Gamma12": Gamma12,
        "volume_weighted_global_xH": g.volume_weighted_global_xH,
        "mass_weighted_global_xH": g.mass_weighted_global_xH,
        "volume_weighted_global_J_21": g.volume_weighted_global_J_21, "n_radii": g.n_radii,
    }


def filter_(ctx, box, R, filter_type, box_size, local_ix_start=0):
    """In-place filter() of a half-spectrum slab (shape (slab_nx, N, N/2+1) complex64)."""
    on_dev = _is_dev(box)
    slab_nx, N = box.shape[0], box.shape[1]
    _bind_stream(ctx, on_dev)
    ctx.check(ctx.lib.mrx_filter(ctx.ptr, _ptr(box), local_ix_start, slab_nx, N, float(R), filter_type, box_size,
                                 1 if on_dev else 0), "filter")
    return box


def fft_r2c_3d(ctx, padded, out=None):
    N = padded.shape[0]
    on_dev = _is_dev(padded)
    if out is None:
        if on_dev:
            out = torch.empty((N, N, N // 2 + 1), dtype=torch.complex64, device=padded.device)
        else:
            out = np.empty((N, N, N // 2 + 1), np.complex64)
    _bind_stream(ctx, on_dev)
    ctx.check(ctx.lib.mrx_fft_r2c_3d(ctx.ptr, N, _ptr(padded), _ptr(out), 1 if on_dev else 0), "fft_r2c_3d")
    return out


def fft_c2r_3d(ctx, herm, out=None):
    N = herm.shape[0]
    on_dev = _is_dev(herm)
    if out is None:
        if on_dev:
            out = torch.empty((N, N, 2 * (N // 2 + 1)), dtype=torch.float32, device=herm.device)
        else:
            out = np.empty((N, N, 2 * (N // 2 + 1)), np.float32)
    _bind_stream(ctx, on_dev)
    ctx.check(ctx.lib.mrx_fft_c2r_3d(ctx.ptr, N, _ptr(herm), _ptr(out), 1 if on_dev else 0), "fft_c2r_3d")
    return out


def filtered_field(ctx, field_padded, box_size, filter_type, R, apply_filter=1, clamp_kind=-1, out=None):
    N = field_padded.shape[0]
    on_dev = _is_dev(field_padded)
    if out is None:
        if on_dev:
            out = torch.empty_like(field_padded)
        else:
            out = np.empty_like(field_padded)
    _bind_stream(ctx, on_dev)
    ctx.check(ctx.lib.mrx_filtered_field(ctx.ptr, N, box_size, filter_type, _ptr(field_padded), float(R),
                                         int(apply_filter), int(clamp_kind), _ptr(out), 1 if on_dev else 0),
              "filtered_field")
    return out


def brightness_temp(ctx, params, xH, deltax, TS_box=None, out=None):
    N = params.grid_dim
    on_dev = _is_dev(xH)
    params.pointers_on_device = 1 if on_dev else 0
    if out is None:
        out = _empty_like_real(xH, N)
    ave = C.c_double()
    _bind_stream(ctx, on_dev)
    ctx.check(ctx.lib.mrx_brightness_temp(ctx.ptr, C.byref(params), _ptr(xH), _ptr(deltax), _ptr(TS_box), _ptr(out),
                                          C.byref(ave)), "brightness_temp")
    return out, ave.value


def brightness_temp_vel(ctx, params, xH, deltax, TS_box, vel, out=None):
    """delta_T with peculiar velocities (BrightnessTemperature.c:126-510).

    ``params`` is an mrx_tb_vel_params; ``vel`` (padded) is converted in place
    to comoving velocity, as the reference does.  Returns (delta_T, <delta_T>).
    """
    N = params.tb.grid_dim
    on_dev = _is_dev(xH)
    params.tb.pointers_on_device = 1 if on_dev else 0
    if out is None:
        out = _empty_like_real(xH, N)
    ave = C.c_double()
    _bind_stream(ctx, on_dev)
    ctx.check(ctx.lib.mrx_brightness_temp_vel(ctx.ptr, C.byref(params), _ptr(xH), _ptr(deltax), _ptr(TS_box),
                                              _ptr(vel), _ptr(out), C.byref(ave)), "brightness_temp_vel")
    return out, ave.value


def gen_grids_fname(output_dir, file_name_galaxies, snapshot, relative=False):
    """gen_grids_fname (reionization.c:1640-1646): the snapshot's grid file name."""
    L = _lib.load_io()
    buf = C.create_string_buffer(4096)
    n = L.mrx_grids_fname(buf, len(buf), (output_dir or "").encode(), file_name_galaxies.encode(), int(snapshot),
                          1 if relative else 0)
    if n < 0:
        raise _lib.MrxError(n, "grids_fname: name does not fit")
    return buf.value.decode()


def save_reion_input_grids(ctx, params, path, deltax, stars, weighted_sfr, sfr=None):
    """save_reion_input_grids (reionization.c:1648-1754): creates the grid file at ``path``.

    ``params`` is an mrx_save_params; the grids are padded, on the device
    (torch) or the host (numpy).  sfr is needed with include_spin_temp.
    """
    L = _lib.load_io()
    on_dev = _is_dev(deltax)
    params.pointers_on_device = 1 if on_dev else 0
    _bind_stream(ctx, on_dev)
    ctx.check(L.mrx_save_reion_input_grids(ctx.ptr, C.byref(params), os.fsencode(path), _ptr(deltax), _ptr(stars),
                                           _ptr(weighted_sfr), _ptr(sfr)), "save_reion_input_grids")


def save_reion_output_grids(ctx, params, path, grids, averages=None):
    """save_reion_output_grids (reionization.c:1786-2050): adds the output grids to ``path``.

    ``grids`` maps the names of mrx_output_grids (xH, z_at_ionization,
    r_bubble, J_21, J_21_at_ionization, Mvir_crit, TS_box, Tk_box,
    x_e_box_prev (padded), delta_T; PS_k, PS_data, PS_error as host float32
    arrays) to arrays; ``averages`` maps the attribute names
    (volume_weighted_global_xH, ...) to floats.
    """
    L = _lib.load_io()
    g = abi.OutputGrids()
    keep = []
    on_dev = _is_dev(grids["xH"])
    for name in abi.OUTPUT_GRID_NAMES:
        a = grids.get(name)
        if a is None:
            continue
        if name.startswith("PS_"):
            a = np.ascontiguousarray(a, dtype=np.float32)
        elif _is_dev(a) != on_dev:
            raise ValueError(f"{name}: every grid must be on the same side (host or device)")
        keep.append(a)
        setattr(g, name, _ptr(a))
    for name, v in (averages or {}).items():
        if name not in abi.OUTPUT_AVERAGES:
            raise KeyError(name)
        setattr(g, name, float(v))
    params.pointers_on_device = 1 if on_dev else 0
    _bind_stream(ctx, on_dev)
    ctx.check(L.mrx_save_reion_output_grids(ctx.ptr, C.byref(params), os.fsencode(path), C.byref(g)),
              "save_reion_output_grids")


def profile(ctx, enable=True):
    """Start (or stop) per-kernel event timing on the context's stream."""
    ctx.check(ctx.lib.mrx_profile(ctx.ptr, 1 if enable else 0), "profile")


def profile_read(ctx, max_kinds=64):
    """{kernel name: (total ms, launches)} since the last read; clears the record."""
    name_len = 64
    names = C.create_string_buffer(max_kinds * name_len)
    ms = np.zeros(max_kinds, np.float64)
    cnt = np.zeros(max_kinds, np.int64)
    n = ctx.lib.mrx_profile_read(ctx.ptr, max_kinds, names, name_len, ms.ctypes.data_as(C.c_void_p),
                                 cnt.ctypes.data_as(C.c_void_p))
    ctx.check(0 if n >= 0 else n, "profile_read")
    out = {}
    raw = names.raw
    for k in range(min(n, max_kinds)):
        nm = raw[k * name_len:(k + 1) * name_len].split(b"\0", 1)[0].decode()
        out[nm] = (float(ms[k]), int(cnt[k]))
    return out


def compute_ts(ctx, params, deltax, stars, sfr, sfr_histories, x_e_box_prev, Tk_box, x_e_box=None, TS_box=None):
    """One spin-temperature step; x_e_box_prev and Tk_box are updated in place.

    Returns the output grids and the volume averages.  ``sfr_histories`` holds
    params.n_history padded grids, slot k being k snapshots back.
    """
    N = params.grid_dim
    on_dev = _is_dev(deltax)
    params.pointers_on_device = 1 if on_dev else 0
    if x_e_box is None:
        x_e_box = torch.empty_like(x_e_box_prev) if on_dev else np.empty_like(x_e_box_prev)
    if TS_box is None:
        TS_box = _empty_like_real(deltax, N)
    g = TsGlobals()
    _bind_stream(ctx, on_dev)
    ctx.check(ctx.lib.mrx_compute_ts(ctx.ptr, C.byref(params), _ptr(deltax), _ptr(stars), _ptr(sfr),
                                     _ptr(sfr_histories), _ptr(x_e_box), _ptr(x_e_box_prev), _ptr(Tk_box),
                                     _ptr(TS_box), C.byref(g)), "compute_ts")
    out = {k: getattr(g, k) for k, _ in TsGlobals._fields_}
    out.update({"x_e_box": x_e_box, "x_e_box_prev": x_e_box_prev, "Tk_box": Tk_box, "TS_box": TS_box})
    return out


def ts_last_shells(ctx, grid_dim, n_shells):
    """The smoothed SFR shells of the last compute_ts call (max(filtered sfr, 0),
    ComputeTs.c:379-503) as a host array [n_shells, N, N, N]."""
    out = np.empty((n_shells, grid_dim, grid_dim, grid_dim), np.float32)
    ctx.check(ctx.lib.mrx_ts_last_shells(ctx.ptr, grid_dim, n_shells, _ptr(out), 0), "ts_last_shells")
    return out


def ts_snapshot_tables(params, collapse_fraction, x_e_ave):
    """The host tables of one spin-temperature snapshot (mrx_ts_snapshot_tables):
    an mrx_ts_tables.  Host only; params.stored_fcoll[snapshot] receives
    collapse_fraction."""
    t = TsTables()
    rc = load().mrx_ts_snapshot_tables(C.byref(params), float(collapse_fraction), float(x_e_ave), C.byref(t))
    if rc != 0:
        raise MrxError(rc, "ts_snapshot_tables")
    return t


def construct_baryon_grids(ctx, params, pos, type_, fesc_weighted_gsm, gross_stellar_mass=None, black_hole_mass=None,
                           effective_bhm=None, stars=None, weighted_sfr=None, sfr=None, sfr_histories=None):
    """Galaxies to the padded stars / weighted_sfr (/ sfr) grids (reionization.c:1353-1615).

    The galaxy fields are arrays in galaxy_to_slab_map order: ``pos`` [n, 3]
    float32 in Mpc/h, ``type_`` int32, the masses float64 - numpy arrays (host)
    or CUDA tensors (device), all alike.  With params.include_spin_temp,
    ``sfr_histories`` ([n_store_sfr] padded grids) is updated in place: moved by
    one slot, slot 0 = the new sfr grid.  Returns the grids and the counts.
    """
    N = params.grid_dim
    on_dev = _is_dev(pos)
    params.pointers_on_device = 1 if on_dev else 0
    n = int(pos.shape[0]) if pos is not None else 0
    want = {"pos": (pos, "float32"), "type": (type_, "int32"), "fesc_weighted_gsm": (fesc_weighted_gsm, "float64"),
            "gross_stellar_mass": (gross_stellar_mass, "float64"), "black_hole_mass": (black_hole_mass, "float64"),
            "effective_bhm": (effective_bhm, "float64")}
    for k, (a, dt) in want.items():
        if a is None:
            continue
        if _is_dev(a) != on_dev:
            raise ValueError(f"{k}: galaxy arrays must all be host or all be device memory")
        if str(a.dtype).replace("torch.", "") != dt:
            raise ValueError(f"{k} must be {dt}, got {a.dtype}")
        if (a.shape[0] if a.ndim else 0) != n or (k == "pos" and tuple(a.shape[1:]) != (3,)):
            raise ValueError(f"{k} has shape {tuple(a.shape)} for {n} galaxies")
    shape = (N, N, 2 * (N // 2 + 1))

    def grid():
        if on_dev:
            return torch.empty(shape, dtype=torch.float32, device=pos.device)
        return np.empty(shape, np.float32)

    stars = grid() if stars is None else stars
    weighted_sfr = grid() if weighted_sfr is None else weighted_sfr
    if params.include_spin_temp and sfr is None:
        sfr = grid()
    st = GridStats()
    _bind_stream(ctx, on_dev)
    ctx.check(ctx.lib.mrx_construct_baryon_grids(
        ctx.ptr, C.byref(params), n, _ptr(pos), _ptr(type_), _ptr(fesc_weighted_gsm), _ptr(gross_stellar_mass),
        _ptr(black_hole_mass), _ptr(effective_bhm), _ptr(stars), _ptr(weighted_sfr), _ptr(sfr), _ptr(sfr_histories),
        C.byref(st)), "construct_baryon_grids")
    return {"stars": stars, "weighted_sfr": weighted_sfr, "sfr": sfr, "sfr_histories": sfr_histories,
            "n_gridded": st.n_gridded, "n_dead": st.n_dead, "n_bh_below_limit": st.n_bh_below_limit,
            "sfr_timescale": st.sfr_timescale}


def power_spectrum(ctx, delta_T, box_size_mpc, k_factor=1.35, max_bins=64):
    """Binned 21-cm power of an unpadded device grid: dict of k, power, error, counts."""
    N = delta_T.shape[0]
    if not _is_dev(delta_T):
        raise ValueError("power_spectrum takes a device grid")
    k = np.zeros(max_bins)
    pw = np.zeros(max_bins)
    er = np.zeros(max_bins)
    cn = np.zeros(max_bins, np.int64)
    _bind_stream(ctx, True)
    n = ctx.lib.mrx_power_spectrum(ctx.ptr, N, float(box_size_mpc), float(k_factor), _ptr(delta_T.contiguous()),
                                   max_bins, k.ctypes.data_as(C.c_void_p), pw.ctypes.data_as(C.c_void_p),
                                   er.ctypes.data_as(C.c_void_p), cn.ctypes.data_as(C.c_void_p))
    ctx.check(0 if n >= 0 else n, "power_spectrum")
    return {"k": k[:n].copy(), "power": pw[:n].copy(), "error": er[:n].copy(), "counts": cn[:n].copy()}


def power_spectrum_adjoint(ctx, delta_T, box_size_mpc, grad_power, k_factor=1.35, out=None):
    """dL/d(delta_T) from dL/dP_b (a host array over the bins)."""
    N = delta_T.shape[0]
    g = np.ascontiguousarray(np.asarray(grad_power, np.float64))
    if out is None:
        out = torch.empty((N, N, N), dtype=torch.float32, device=delta_T.device)
    _bind_stream(ctx, True)
    ctx.check(ctx.lib.mrx_power_spectrum_adjoint(ctx.ptr, N, float(box_size_mpc), float(k_factor),
                                                 _ptr(delta_T.contiguous()), g.ctypes.data_as(C.c_void_p),
                                                 _ptr(out)), "power_spectrum_adjoint")
    return out


def snapshot_adjoint(ctx, ts_params, hii_params, tb_params, deltax, stars, sfr, sfr_histories, x_e_box_prev, Tk_box,
                     xH, TS_box, x_e_box, grad_delta_T, want=("deltax", "stars", "sfr", "sfr_histories",
                                                             "x_e_box_prev", "Tk_box")):
    """Backward pass of one snapshot; returns {name: gradient} for ``want``.

    x_e_box_prev / Tk_box are the values BEFORE the forward's spin-temperature
    step; xH, TS_box, x_e_box are the forward's outputs.
    """
    for p in (ts_params, hii_params, tb_params):
        if p is not None:
            p.pointers_on_device = 1
    g = {}
    g["deltax"] = torch.empty_like(deltax)
    g["stars"] = torch.empty_like(stars)
    if ts_params is not None:
        if "sfr" in want:
            g["sfr"] = torch.empty_like(sfr)
        if "sfr_histories" in want and sfr_histories is not None:
            g["sfr_histories"] = torch.empty_like(sfr_histories)
        if "x_e_box_prev" in want:
            g["x_e_box_prev"] = torch.empty_like(x_e_box_prev)
        if "Tk_box" in want:
            g["Tk_box"] = torch.empty_like(Tk_box)
    _bind_stream(ctx, True)
    ctx.check(ctx.lib.mrx_snapshot_adjoint(
        ctx.ptr, None if ts_params is None else C.byref(ts_params), C.byref(hii_params), C.byref(tb_params),
        _ptr(deltax), _ptr(stars), _ptr(sfr), _ptr(sfr_histories), _ptr(x_e_box_prev), _ptr(Tk_box), _ptr(xH),
        _ptr(TS_box), _ptr(x_e_box), _ptr(grad_delta_T.contiguous()), _ptr(g["deltax"]), _ptr(g["stars"]),
        _ptr(g.get("sfr")), _ptr(g.get("sfr_histories")), _ptr(g.get("x_e_box_prev")), _ptr(g.get("Tk_box"))),
        "snapshot_adjoint")
    return g


class HiiBatch:
    """Many boxes through the bubble finder as one replayed HIP graph.

    ``boxes`` is a list of dicts with ``params`` (HiiParams) and device grids
    ``deltax``, ``stars``, ``weighted_sfr`` (padded), optional ``x_e``, and
    optional preallocated ``xH``, ``r_bubble``, ``J_21``, ``z_at_ionization``,
    ``J_21_at_ionization`` (allocated here otherwise, z_at_ionization = -1).
    The grids stay bound to the plan: ``run()`` recomputes every box from the
    current contents of its input tensors.
    """

    def __init__(self, ctx, boxes, graph=True, tune=True):
        self.ctx = ctx
        self.boxes = []
        arr = (HiiBatchBox * len(boxes))()
        for i, b in enumerate(boxes):
            p = b["params"]
            p.pointers_on_device = 1
            N = p.grid_dim
            d = b["deltax"]
            if not _is_dev(d):
                raise ValueError("batch grids must be torch tensors on the GPU")
            o = {k: b.get(k) for k in ("xH", "r_bubble", "J_21", "z_at_ionization", "J_21_at_ionization")}
            for k in ("xH", "r_bubble"):
                if o[k] is None:
                    o[k] = _empty_like_real(d, N)
            if o["z_at_ionization"] is None:
                o["z_at_ionization"] = _empty_like_real(d, N, -1.0)
            if p.uvb_flag:
                if o["J_21"] is None:
                    o["J_21"] = _empty_like_real(d, N)
                if o["J_21_at_ionization"] is None:
                    o["J_21_at_ionization"] = _empty_like_real(d, N, 0.0)
            keep = dict(b)
            keep.update(o)
            self.boxes.append(keep)
            arr[i].params = p
            for k in ("deltax", "stars", "weighted_sfr", "x_e", "xH", "r_bubble", "J_21", "z_at_ionization",
                      "J_21_at_ionization"):
                v = keep.get(k)
                setattr(arr[i], k, None if v is None else v.data_ptr())
        self._arr = arr
        flags = (0 if graph else 1) | (0 if tune else 2)
        self.ptr = C.c_void_p()
        _bind_stream(ctx, True)
        ctx.check(ctx.lib.mrx_hii_batch_create(ctx.ptr, arr, len(boxes), flags, C.byref(self.ptr)), "hii_batch_create")

    def run(self, sync=True):
        """Replays the plan; with ``sync`` returns the per-box globals."""
        _bind_stream(self.ctx, True)
        if not sync:
            self.ctx.check(self.ctx.lib.mrx_hii_batch_run(self.ptr, None), "hii_batch_run")
            return None
        g = (HiiGlobals * len(self.boxes))()
        self.ctx.check(self.ctx.lib.mrx_hii_batch_run(self.ptr, g), "hii_batch_run")
        return [{"volume_weighted_global_xH": x.volume_weighted_global_xH,
                 "mass_weighted_global_xH": x.mass_weighted_global_xH,
                 "volume_weighted_global_J_21": x.volume_weighted_global_J_21, "n_radii": x.n_radii} for x in g]

    def describe(self):
        import json

        buf = C.create_string_buffer(1 << 16)
        n = self.ctx.lib.mrx_hii_batch_describe(self.ptr, buf, len(buf))
        if n < 0:
            raise MrxError(n, "hii_batch_describe")
        return json.loads(buf.value.decode())

    def close(self):
        if self.ptr:
            self.ctx.lib.mrx_hii_batch_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
