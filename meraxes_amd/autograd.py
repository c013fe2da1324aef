"""Differentiable form of one snapshot and of the 21-cm power spectrum.

``snapshot(...)`` runs the reference's order - the spin-temperature step
(_ComputeTs, ComputeTs.c:78-1020), the bubble finder with x_e
(_find_HII_bubbles, find_HII_bubbles.c:47-495) and the brightness temperature
(BrightnessTemperature.c:24-124) - on the HIP path and returns delta_T as a
torch tensor whose backward is the library's hand-written adjoint
(mrx_snapshot_adjoint): the ionisation mask is a stop-gradient, per-radius
fields are recomputed in the backward rather than stored, and the shell
tables' two grid means are held fixed.  ``power_spectrum(...)`` is Compute_PS
(ComputePowerSpectrum.c:44-255) with mrx_power_spectrum_adjoint as its
backward.  Gradients flow to deltax, stars, sfr, the SFR history, x_e_prev and
T_K; weighted_sfr only feeds J_21, so its gradient is zero.
"""
import torch

from . import reion


class _Snapshot(torch.autograd.Function):
    @staticmethod
    def forward(fctx, lib, tsp, hp, bp, deltax, stars, weighted_sfr, sfr, hist, xe_prev, Tk):
        N = hp.grid_dim
        xe_in = xe_prev.detach().clone()
        Tk_in = Tk.detach().clone()
        xe_box = torch.empty_like(xe_in)
        TS = torch.empty((N, N, N), dtype=torch.float32, device=deltax.device)
        Tk_out = Tk_in.clone()
        if tsp is not None:
            xe_work = xe_in.clone()
            reion.compute_ts(lib, tsp, deltax, stars, sfr, hist, xe_work, Tk_out, x_e_box=xe_box, TS_box=TS)
        r = reion.find_hii_bubbles(lib, hp, deltax, stars, weighted_sfr,
                                   x_e=xe_box if hp.include_spin_temp else None)
        dT, _ = reion.brightness_temp(lib, bp, r["xH"], deltax, TS if bp.include_spin_temp else None)
        fctx.save_for_backward(deltax, stars, sfr, hist, xe_in, Tk_in, r["xH"], TS, xe_box)
        fctx.lib = lib
        fctx.params = (tsp, hp, bp)
        fctx.wsfr_shape = weighted_sfr.shape
        xH = r["xH"]
        fctx.mark_non_differentiable(xH, TS, xe_box, Tk_out)
        return dT, xH, TS, xe_box, Tk_out

    @staticmethod
    def backward(fctx, g_dT, g_xH, g_TS, g_xe, g_Tk):
        deltax, stars, sfr, hist, xe_in, Tk_in, xH, TS, xe_box = fctx.saved_tensors
        tsp, hp, bp = fctx.params
        if g_dT is None:
            g_dT = torch.zeros_like(xH)
        g = reion.snapshot_adjoint(fctx.lib, tsp, hp, bp, deltax, stars, sfr, hist, xe_in, Tk_in, xH, TS, xe_box,
                                   g_dT.float())
        zeros = torch.zeros(fctx.wsfr_shape, dtype=torch.float32, device=deltax.device)
        return (None, None, None, None, g["deltax"], g["stars"], zeros, g.get("sfr"), g.get("sfr_histories"),
                g.get("x_e_box_prev"), g.get("Tk_box"))


def snapshot(lib, ts_params, hii_params, tb_params, deltax, stars, weighted_sfr, sfr=None, sfr_histories=None,
             x_e_prev=None, Tk=None):
    """delta_T (N, N, N) of one snapshot, differentiable; also returns (xH, T_S, x_e, T_K) detached.

    ``ts_params`` None runs without spin temperature (then hii_params and
    tb_params must not include it).  Grids are device tensors in the
    reference's layouts (padded deltax, stars, weighted_sfr, sfr, x_e_prev;
    sfr_histories (n_history, N, N, N+2); unpadded T_K).
    """
    N = hii_params.grid_dim
    dev = deltax.device
    if sfr is None:
        sfr = torch.zeros_like(deltax)
    if sfr_histories is None:
        sfr_histories = torch.zeros((1,) + tuple(deltax.shape), dtype=torch.float32, device=dev)
    if x_e_prev is None:
        x_e_prev = torch.zeros_like(deltax)
    if Tk is None:
        Tk = torch.zeros((N, N, N), dtype=torch.float32, device=dev)
    for p in (ts_params, hii_params, tb_params):
        if p is not None:
            p.pointers_on_device = 1
    dT, xH, TS, xe, Tk_new = _Snapshot.apply(lib, ts_params, hii_params, tb_params, deltax, stars, weighted_sfr, sfr,
                                             sfr_histories, x_e_prev, Tk)
    return dT, (xH, TS, xe, Tk_new)


class _PowerSpectrum(torch.autograd.Function):
    @staticmethod
    def forward(fctx, lib, delta_T, box_size_mpc, k_factor):
        r = reion.power_spectrum(lib, delta_T.detach(), box_size_mpc, k_factor)
        fctx.save_for_backward(delta_T.detach())
        fctx.lib = lib
        fctx.args = (box_size_mpc, k_factor)
        return torch.as_tensor(r["power"], dtype=torch.float64, device=delta_T.device)

    @staticmethod
    def backward(fctx, g_p):
        (dT,) = fctx.saved_tensors
        box, kf = fctx.args
        g = reion.power_spectrum_adjoint(fctx.lib, dT, box, g_p.detach().double().cpu().numpy(), kf)
        return None, g, None, None


def power_spectrum(lib, delta_T, box_size_mpc, k_factor=1.35):
    """Binned power (float64 tensor over the bins), differentiable in delta_T."""
    return _PowerSpectrum.apply(lib, delta_T, box_size_mpc, k_factor)


def k_factor_for_bins(N, n_bins):
    """A bin growth factor that gives n_bins bins of Compute_PS's scheme on an N^3 grid."""
    import math

    return float(math.exp(math.log(N) / (n_bins - 0.5)))
