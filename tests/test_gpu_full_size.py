"""The benchmark's 512^3 box on the GPU, checked through size-independent properties.

The oracle finishes in seconds only on small boxes; at the full size the
HIP path is checked against what the mathematics fixes, whatever the size:
the FFT round trip is the identity and conserves power (Parseval); the
filter keeps the mean (W(0) = 1) and is linear; the bubble finder's globals
are the means of its own grids, every ionised cell carries a ladder radius and
the snapshot's redshift, raising the efficiency never raises a cell's xH, and
the x pass gives the same bits for one or two radii per launch.
"""
import os

import numpy as np
import pytest

from meraxes_amd import params as P
from meraxes_amd import reion, synth
from meraxes_amd._lib import Context

pytestmark = pytest.mark.gpu

N, L, Z = 512, 400.0, 8.0


@pytest.fixture(scope="module")
def fields(ctx):
    import torch

    print(f"full size: generating {N}^3 inputs", flush=True)
    d, s, w = synth.snapshot_inputs_torch(N, 11, "cuda")
    torch.cuda.synchronize()
    yield d, s, w
    del d, s, w
    torch.cuda.empty_cache()


def test_fft_round_trip_and_parseval_512(ctx, fields):
    import torch

    print("full size: FFT round trip", flush=True)
    d = fields[0]
    H = reion.fft_r2c_3d(ctx, d)
    back = reion.fft_c2r_3d(ctx, H)  # unnormalised, as FFTW: N^3 times the input
    torch.cuda.synchronize()
    x = d[:, :, :N].double()
    y = back[:, :, :N].double() / N**3
    assert (torch.linalg.norm(y - x) / torch.linalg.norm(x)).item() < 1e-6
    # Parseval over the half spectrum: the kz = 1 .. N/2-1 planes stand for two modes each
    p = H.abs().double() ** 2
    total = p[:, :, 0].sum() + p[:, :, N // 2].sum() + 2 * p[:, :, 1:N // 2].sum()
    assert abs(total.item() / (N**3 * (x * x).sum().item()) - 1) < 1e-5


@pytest.mark.parametrize("ftype,R", [(0, 5.0), (2, 2.0), (1, 20.0)])
def test_filter_keeps_the_mean_and_is_linear_512(ctx, fields, ftype, R):
    import torch

    print(f"full size: filter type {ftype} R={R}", flush=True)
    d, s, _ = fields
    s1 = s / s[:, :, :N].std()  # a second field of comparable scale
    fd = reion.filtered_field(ctx, d, L, ftype, R)
    fs = reion.filtered_field(ctx, s1, L, ftype, R)
    mix = (0.3 * d + 1.7 * s1).contiguous()
    fm = reion.filtered_field(ctx, mix, L, ftype, R)
    torch.cuda.synchronize()
    for f, g in ((fd, d), (fs, s1)):
        m_in, m_out = g[:, :, :N].double().mean().item(), f[:, :, :N].double().mean().item()
        assert abs(m_out - m_in) <= 1e-6 * max(abs(m_in), g[:, :, :N].double().std().item())
    lin = 0.3 * fd[:, :, :N].double() + 1.7 * fs[:, :, :N].double()
    err = torch.linalg.norm(fm[:, :, :N].double() - lin) / torch.linalg.norm(lin)
    assert err.item() < 1e-5


def test_bubble_finder_invariants_512(ctx, fields):
    import torch

    d, s, w = fields
    p1 = P.hii_params(N, L, Z, uvb_flag=1, efficiency=20.0)
    p2 = P.hii_params(N, L, Z, uvb_flag=1, efficiency=40.0)
    print("full size: bubble finder, eps 20 and 40", flush=True)
    o1 = reion.find_hii_bubbles(ctx, p1, d, s, w)
    o2 = reion.find_hii_bubbles(ctx, p2, d, s, w)
    torch.cuda.synchronize()
    dl = d[:, :, :N].double()
    radii = torch.tensor(np.asarray(reion.radius_ladder(p1), np.float32), device=d.device)
    for o in (o1, o2):
        xh = o["xH"].double()
        assert abs(o["volume_weighted_global_xH"] - xh.mean().item()) < 1e-9
        mass = ((xh * (1 + dl)).sum() / (1 + dl).sum()).item()
        assert abs(o["mass_weighted_global_xH"] - mass) < 1e-9
        assert abs(o["volume_weighted_global_J_21"] - o["J_21"].double().mean().item()) < 1e-9 * max(
            1.0, abs(o["volume_weighted_global_J_21"]))
        ion = o["xH"] == 0
        assert 0 < ion.double().mean().item() < 1
        assert torch.isin(o["r_bubble"][ion], radii).all()
        # z_at_ionization is set wherever xH ends below 1e-5 (find_HII_bubbles.c:427-434)
        low = o["xH"] < 1e-5
        assert (o["z_at_ionization"][low] == np.float32(Z)).all()
        assert (o["z_at_ionization"][~low] == -1).all()
        assert ((o["xH"] >= 0) & (o["xH"] <= 1)).all()
    # a larger efficiency ionises at least every cell the smaller one does, and
    # leaves no partially ionised cell more neutral
    assert (o2["xH"] <= o1["xH"]).all()
    # one radius per x pass gives the same bits as the default two
    print("full size: one radius per x pass", flush=True)
    old = os.environ.get("MRX_LADDER_RADII")
    os.environ["MRX_LADDER_RADII"] = "1"
    try:
        c1 = Context(0)
    finally:
        if old is None:
            del os.environ["MRX_LADDER_RADII"]
        else:
            os.environ["MRX_LADDER_RADII"] = old
    try:
        o3 = reion.find_hii_bubbles(c1, p1, d, s, w)
        torch.cuda.synchronize()
    finally:
        c1.close()
    for k in ("xH", "r_bubble", "J_21", "z_at_ionization", "J_21_at_ionization"):
        assert torch.equal(o3[k], o1[k]), k
    assert o3["volume_weighted_global_xH"] == o1["volume_weighted_global_xH"]


def test_spin_temperature_and_brightness_temperature_512(ctx, fields):
    """The full-snapshot configuration's Ts and delta_T at 512^3, 400 Mpc/h:
    snapshot 0 (z = 20 = ReionMaxHeatingRedshift) takes the homogeneous branch,
    where x_e and T_K equal the RECFAST splines at z everywhere
    (ComputeTs.c:189-262; values from the oracle's restatement of the table);
    snapshot 24 (z = 8.41, 40 shells, an 8-slot SFR history) reports averages
    that are the means of its own grids (ComputeTs.c:940-985), keeps x_e in
    [0, 1) and T_S, T_K positive; then the bubble finder with x_e as the fourth
    field, and delta_T = C xH (1 + delta) (1 - T_CMB (1+z) / T_S)
    (BrightnessTemperature.c:46-124) recomputed cell by cell in float64 from
    the GPU's own grids, <delta_T> their mean."""
    import torch

    from oracle import ts_host

    d, s, w = fields
    zz = P.snapshot_redshifts()
    nh = 8
    fc = np.zeros(len(zz))
    fc[:24] = np.linspace(1e-4, 5e-3, 24)
    xe_prev = torch.zeros_like(d)
    Tk = torch.zeros((N, N, N), device=d.device)
    # snapshot 0: homogeneous
    print("full size: spin temperature, homogeneous branch", flush=True)
    hist = torch.zeros((nh,) + tuple(w.shape), device=d.device)
    p0 = P.ts_params(N, L, 0, zz, fc, nh, on_device=1)
    g0 = reion.compute_ts(ctx, p0, d, s, w, hist, xe_prev, Tk)
    torch.cuda.synchronize()
    h = ts_host.host(p0)
    assert g0["homogeneous"] == 1
    assert (xe_prev[:, :, :N] == np.float32(h.xion_RECFAST(zz[0]))).all()
    assert (Tk == np.float32(h.T_RECFAST(zz[0]))).all()
    TS0 = g0["TS_box"].double()
    assert torch.isfinite(TS0).all() and (TS0 > 0).all()
    assert abs(g0["volume_ave_TS"] - TS0.mean().item()) <= 1e-9 * TS0.mean().item()
    # snapshot 24: inhomogeneous, 40 shells
    print("full size: spin temperature, 40 shells", flush=True)
    snap = 24
    z = float(zz[snap])
    for k in range(nh):
        hist[k] = w * float(((1 + z) / (1 + zz[max(snap - k, 0)])) ** 3)
    xe_prev.fill_(2e-3)
    Tk.fill_(30.0)
    p1 = P.ts_params(N, L, snap, zz, fc, nh, on_device=1)
    g1 = reion.compute_ts(ctx, p1, d, s, w, hist, xe_prev, Tk)
    torch.cuda.synchronize()
    assert g1["homogeneous"] == 0 and g1["no_light"] == 0
    xe = g1["x_e_box"][:, :, :N].double()
    TK, TS = g1["Tk_box"].double(), g1["TS_box"].double()
    assert torch.equal(g1["x_e_box"], xe_prev)
    assert ((xe >= 0) & (xe < 1)).all() and (TK > 0).all() and (TS > 0).all() and torch.isfinite(TS).all()
    for key, grid in (("volume_ave_TS", TS), ("volume_ave_TK", TK), ("volume_ave_xe", xe)):
        assert abs(g1[key] - grid.mean().item()) <= 1e-9 * abs(grid.mean().item()), key
    assert TK.mean().item() > 30.0  # X-ray heating outruns adiabatic cooling over this step
    # bubbles with x_e, then delta_T from the GPU's own grids
    print("full size: bubble finder with x_e and the brightness temperature", flush=True)
    hp = P.hii_params(N, L, z, efficiency=20.0, uvb_flag=1, include_spin_temp=1)
    o = reion.find_hii_bubbles(ctx, hp, d, s, w, x_e=g1["x_e_box"])
    bp = P.tb_params(N, z, include_spin_temp=1)
    dT, ave = reion.brightness_temp(ctx, bp, o["xH"], d, g1["TS_box"])
    torch.cuda.synchronize()
    # the constant as the reference forms it (BrightnessTemperature.c:46-52): float
    # cosmology, the products of floats in float, the rest in double
    Om, h_ = np.float32(P.OMEGA_M), np.float32(P.HUBBLE_H)
    Ob = np.float32(float(Om) * P.BARYON_FRAC)
    zf = float(np.float32(z))
    obh2 = float(np.float32(np.float32(Ob * h_) * h_))
    C = float(np.float32(27.0 * (obh2 / 0.023) * np.sqrt((0.15 / float(Om) / float(h_) / float(h_)) * (1.0 + zf)
                                                       / 10.0)))
    want = C * o["xH"].double() * (1 + d[:, :, :N].double()) * (1 - 2.728 * (1 + zf) / TS)
    got = dT.double()
    err = (got - want).abs() / want.abs().clamp_min(1e-30)
    big = want.abs() > 1e-3 * want.abs().max()
    assert err[big].max().item() < 1e-6
    assert (got - want).abs().max().item() < 1e-6 * want.abs().max().item()
    assert abs(ave - got.mean().item()) <= 1e-9 * abs(got.mean().item())
    del hist, xe_prev, Tk, o, dT
    torch.cuda.empty_cache()
