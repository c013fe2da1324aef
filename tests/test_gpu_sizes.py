"""Grid sizes with a factor 5 (160, 320, 640) and 1024^3, the size of the
reference's own Genesis_L210_N4320.par (ReionGridDim 1024).  FFTW plans take
any ReionGridDim (reionization.c:685-736); the HIP path covers these with a
radix-5 butterfly in the same Stockham engine.

160^3 and 320^3 are checked against the oracle as the other sizes are
(transforms against FP64, filtered fields against ref32 within rel-L2 1e-5,
the bubble finder's parity contract, spin temperature within 1e-4 per cell);
640^3 and 1024^3, where the oracle would take minutes, through the
size-independent properties of the 512^3 tests: the FFT round trip is the
identity and keeps Parseval's sum, the window keeps the mean and is linear,
the bubble finder's globals are the means of its own grids and every ionised
cell carries a ladder radius.
"""
import numpy as np
import pytest

from meraxes_amd import params as P
from meraxes_amd import reion, synth
from meraxes_amd._lib import Context
from oracle import oracle as O

from test_gpu_hii import check_parity, run_pair

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    dt = np.complex128 if np.iscomplexobj(a) or np.iscomplexobj(b) else np.float64
    a, b = a.astype(dt), b.astype(dt)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("N", [160, 320])
def test_transforms_match_fp64(ctx, N):
    d = synth.to_padded(synth.gaussian_field(N, 13))
    print(f"N={N}: r2c and c2r against FP64", flush=True)
    H = reion.fft_r2c_3d(ctx, d)
    ref = O.r2c3d(d.astype(np.float64))
    e_f = rel_l2(H, ref)
    back = reion.fft_c2r_3d(ctx, ref.astype(np.complex64))
    ref_b = O.c2r3d(ref)
    e_b = rel_l2(back[:, :, :N], ref_b[:, :, :N])
    print(f"N={N}: r2c rel L2 {e_f:.2e}, c2r rel L2 {e_b:.2e}", flush=True)
    assert e_f < 2e-6 and e_b < 2e-6


@pytest.mark.parametrize("ftype", [0, 1, 2])
def test_filtered_fields_160(ctx, ftype):
    N, L = 160, 125.0
    d, s, w = synth.snapshot_inputs(N, 4)
    R_all = P.radius_ladder(P.hii_params(N, L, 8.0, filter_type=ftype))
    picks = [R_all[0], R_all[len(R_all) // 2], R_all[-2], R_all[-1]]
    for field, kind in ((d, 0), (s, 1), (w, 1)):
        for i, R in enumerate(picks):
            apply = 0 if i == len(picks) - 1 else 1
            g = reion.filtered_field(ctx, field, L, ftype, R, apply, kind)
            o32 = O.filtered_field(N, L, ftype, field, R, apply, kind)
            e32 = rel_l2(g[:, :, :N], o32[:, :, :N])
            print(f"N={N} type={ftype} kind={kind} R={R:.3f}: gpu-ref32 {e32:.2e}", flush=True)
            assert e32 <= 1e-5


def partial_cells_vs_fp64(ctx, p, d, s, w, g, o, label, xe=None):
    """The partially neutral cells (n_f - eff f_coll from the float round trip
    of the unfiltered fields) must be as close to the FP64 oracle as the float
    oracle is, as at 256^3."""
    o64 = O.find_hii_bubbles(p, d, s, w, x_e=xe, precision=64)
    neutral = (g["xH"] >= 1e-5) & (o["xH"] >= 1e-5) & (o64["xH"] >= 1e-5)
    e_g = float(np.abs(g["xH"][neutral].astype(np.float64) - o64["xH"][neutral]).max())
    e_32 = float(np.abs(o["xH"][neutral].astype(np.float64) - o64["xH"][neutral]).max())
    print(f"{label} partial xH vs FP64: gpu {e_g:.2e}, float oracle {e_32:.2e}", flush=True)
    assert e_g <= 2 * e_32 + 1e-6


def test_bubble_finder_160_with_x_e(ctx):
    """160^3 on a 125 Mpc/h box (the 512^3, 400 Mpc/h cell size), top-hat, UVB
    on, x_e as the fourth field."""
    N = 160
    print(f"{N}^3 with x_e: oracle and GPU", flush=True)
    d, s, w = synth.snapshot_inputs(N, 21)
    rng = np.random.default_rng(8)
    xe = synth.to_padded((0.02 + 0.05 * rng.random((N, N, N))).astype(np.float32))
    p = P.hii_params(N, 125.0, 8.0, efficiency=30.0, uvb_flag=1, include_spin_temp=1)
    g, o = run_pair(ctx, p, d, s, w, xe=xe)
    check_parity(g, o, N, "160 x_e", eff=30.0, neutral_vs_fp64=True)
    partial_cells_vs_fp64(ctx, p, d, s, w, g, o, "160 x_e", xe=xe)


def test_bubble_finder_320_sharpk(ctx):
    """320^3 on a 250 Mpc/h box, sharp-k with UVB."""
    N = 320
    print(f"{N}^3 sharp-k: oracle and GPU", flush=True)
    d, s, w = synth.snapshot_inputs(N, 22)
    p = P.hii_params(N, 250.0, 8.0, efficiency=20.0, uvb_flag=1, filter_type=1)
    g, o = run_pair(ctx, p, d, s, w)
    check_parity(g, o, N, "320 sharp-k", eff=20.0, neutral_vs_fp64=True)
    partial_cells_vs_fp64(ctx, p, d, s, w, g, o, "320 sharp-k")


def test_spin_temperature_160(ctx):
    """Two snapshots of the spin temperature at 160^3 (125 Mpc/h): x_e, T_K and
    T_S within 1e-4 per cell, each side carrying its own state."""
    N, L = 160, 125.0
    zz = P.snapshot_redshifts()
    nh = 8
    fc = {"g": np.zeros(len(zz)), "o": np.zeros(len(zz))}
    PR = 2 * (N // 2 + 1)
    hist = np.zeros((nh, N, N, PR), np.float32)
    st = {k: {"xe": np.zeros((N, N, PR), np.float32), "Tk": np.zeros((N, N, N), np.float32)} for k in "go"}
    for snap in (0, 12):
        d, s, w = synth.snapshot_inputs(N, 40 + snap)
        hist = np.roll(hist, 1, axis=0)
        hist[0] = w
        print(f"Ts at {N}^3, snapshot {snap}", flush=True)
        tg = reion.compute_ts(ctx, P.ts_params(N, L, snap, zz, fc["g"], nh), d, s, w, hist, st["g"]["xe"],
                              st["g"]["Tk"])
        to = O.compute_ts(P.ts_params(N, L, snap, zz, fc["o"], nh), d, s, w, hist, st["o"]["xe"], st["o"]["Tk"])
        # the GPU call updates its state in place; the oracle returns copies
        st["o"]["xe"], st["o"]["Tk"] = to["x_e_box_prev"], to["Tk_box"]
        for k in ("x_e_box", "Tk_box", "TS_box"):
            a = tg[k][:, :, :N].astype(np.float64)
            b = to[k][:, :, :N].astype(np.float64)
            r = float((np.abs(a - b) / np.maximum(np.abs(b), 1e-30)).max())
            print(f"  {k}: max rel {r:.2e}", flush=True)
            assert r <= 1e-4


@pytest.fixture
def own_ctx():
    """A context of its own, released after the test: a 1024^3 box's buffers
    (spectra, shells) are tens of GiB, which the session's context would keep."""
    import torch

    c = Context(0)
    yield c
    c.close()
    torch.cuda.empty_cache()


@pytest.mark.parametrize("N,L", [(640, 500.0), (1024, 800.0)])
def test_large_grid_properties(own_ctx, N, L):
    import torch

    ctx = own_ctx
    print(f"{N}^3: generating inputs on the GPU", flush=True)
    d, s, w = synth.snapshot_inputs_torch(N, 5, "cuda")
    torch.cuda.synchronize()
    # FFT round trip and Parseval
    H = reion.fft_r2c_3d(ctx, d)
    back = reion.fft_c2r_3d(ctx, H)
    torch.cuda.synchronize()
    x = d[:, :, :N].double()
    y = back[:, :, :N].double() / float(N) ** 3
    e_rt = (torch.linalg.norm(y - x) / torch.linalg.norm(x)).item()
    p = H.abs().double() ** 2
    total = p[:, :, 0].sum() + p[:, :, N // 2].sum() + 2 * p[:, :, 1:N // 2].sum()
    e_ps = abs(total.item() / (float(N) ** 3 * (x * x).sum().item()) - 1)
    del H, back, y, p
    print(f"{N}^3: round trip rel L2 {e_rt:.2e}, Parseval {e_ps:.2e}", flush=True)
    assert e_rt < 1e-6 and e_ps < 1e-5
    # top-hat window: the mean is kept, the filter is linear
    R = 4.0 * L / N
    s1 = s / s[:, :, :N].std()
    fd = reion.filtered_field(ctx, d, L, 0, R)
    fs = reion.filtered_field(ctx, s1, L, 0, R)
    fm = reion.filtered_field(ctx, (0.3 * d + 1.7 * s1).contiguous(), L, 0, R)
    torch.cuda.synchronize()
    for f, g in ((fd, d), (fs, s1)):
        m_in, m_out = g[:, :, :N].double().mean().item(), f[:, :, :N].double().mean().item()
        assert abs(m_out - m_in) <= 1e-6 * max(abs(m_in), g[:, :, :N].double().std().item())
    lin = 0.3 * fd[:, :, :N].double() + 1.7 * fs[:, :, :N].double()
    e_lin = (torch.linalg.norm(fm[:, :, :N].double() - lin) / torch.linalg.norm(lin)).item()
    print(f"{N}^3: filter linearity rel L2 {e_lin:.2e}", flush=True)
    assert e_lin < 1e-5
    del fd, fs, fm, lin, s1, x
    torch.cuda.empty_cache()
    # the bubble finder: globals are the means of its grids, ionised cells carry a ladder radius
    hp = P.hii_params(N, L, 8.0, uvb_flag=1, efficiency=20.0)
    print(f"{N}^3: bubble finder", flush=True)
    o = reion.find_hii_bubbles(ctx, hp, d, s, w)
    torch.cuda.synchronize()
    xH = o["xH"].double()
    dm = 1.0 + d[:, :, :N].double()
    assert abs(o["volume_weighted_global_xH"] - xH.mean().item()) <= 1e-9
    assert abs(o["mass_weighted_global_xH"] - ((xH * dm).sum() / dm.sum()).item()) <= 1e-9
    ion = o["xH"] == 0
    radii = torch.tensor(P.radius_ladder(hp), dtype=torch.float32, device=xH.device)
    rb = o["r_bubble"]
    assert bool((rb[~ion] == 0).all())
    assert bool(torch.isin(rb[ion], radii).all())
    assert bool((o["z_at_ionization"][ion] == np.float32(8.0)).all())
    print(f"{N}^3: <xH>_V {o['volume_weighted_global_xH']:.6f}, ionised {ion.double().mean().item():.4f}",
          flush=True)
    assert 0.0 < o["volume_weighted_global_xH"] < 1.0


def test_spin_temperature_1024(own_ctx):
    """The spin-temperature step at 1024^3 (800 Mpc/h, the 512^3 / 400 Mpc/h
    cell size): 40 shells of 4 GiB each stay resident (160 GiB) beside the
    step's grids on one GPU.  Snapshot 24 with an 8-slot SFR history: the
    averages are the means of the step's own grids (ComputeTs.c:940-985), x_e
    stays in [0, 1), T_S and T_K positive and finite, and X-ray heating raises
    the mean T_K over the step, as at 512^3."""
    import torch

    ctx = own_ctx
    N, L = 1024, 800.0
    torch.cuda.empty_cache()  # earlier tests' cached blocks: the shells need 160 GiB
    print(f"{N}^3: spin temperature, generating inputs on the GPU", flush=True)
    d, s, w = synth.snapshot_inputs_torch(N, 6, "cuda")
    zz = P.snapshot_redshifts()
    nh, snap = 8, 24
    z = float(zz[snap])
    fc = np.zeros(len(zz))
    fc[:snap] = np.linspace(1e-4, 5e-3, snap)
    hist = torch.empty((nh,) + tuple(w.shape), device=w.device)
    for k in range(nh):
        hist[k] = w * float(((1 + z) / (1 + zz[max(snap - k, 0)])) ** 3)
    xe_prev = torch.full_like(d, 2e-3)
    Tk = torch.full((N, N, N), 30.0, device=d.device)
    print(f"{N}^3: spin temperature, 40 shells", flush=True)
    g = reion.compute_ts(ctx, P.ts_params(N, L, snap, zz, fc, nh, on_device=1), d, s, w, hist, xe_prev, Tk)
    torch.cuda.synchronize()
    del d, s, w, hist
    assert g["homogeneous"] == 0 and g["no_light"] == 0
    # checked 64 x-planes at a time: float64 copies of the whole grids would
    # need 24 GiB more beside the resident shells
    grids = {"volume_ave_xe": g["x_e_box"], "volume_ave_TK": g["Tk_box"], "volume_ave_TS": g["TS_box"]}
    sums = dict.fromkeys(grids, 0.0)
    for x0 in range(0, N, 64):
        xe = grids["volume_ave_xe"][x0:x0 + 64, :, :N].double()
        TK = grids["volume_ave_TK"][x0:x0 + 64].double()
        TS = grids["volume_ave_TS"][x0:x0 + 64].double()
        assert ((xe >= 0) & (xe < 1)).all() and (TK > 0).all() and (TS > 0).all() and torch.isfinite(TS).all()
        for key, grid in (("volume_ave_xe", xe), ("volume_ave_TK", TK), ("volume_ave_TS", TS)):
            sums[key] += grid.sum().item()
        del xe, TK, TS
    for key in grids:
        mean = sums[key] / float(N) ** 3
        assert abs(g[key] - mean) <= 1e-9 * abs(mean), key
    print(f"{N}^3: <T_S> {g['volume_ave_TS']:.4f} <T_K> {g['volume_ave_TK']:.4f} <x_e> {g['volume_ave_xe']:.3e}, "
          f"device memory in use {torch.cuda.memory_allocated() / 2**30:.1f} GiB (torch)", flush=True)
    assert sums["volume_ave_TK"] / float(N) ** 3 > 30.0
