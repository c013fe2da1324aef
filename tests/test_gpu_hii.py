"""GPU parity of find_HII_bubbles: the HIP path vs the oracle on identical seeded grids.

Checks (the three of the parity contract):
1. filtered fields: covered in test_gpu_fft_filter.py (rel L2 <= 1e-5);
2. global x_HI within 1e-5 of ref32, and at most 1e-4 of cells flip between
   ionised and neutral, each flip within 1e-4 relative of the threshold (the
   oracle's per-cell max over the ladder of f_coll*eff/n_f is that margin);
3. every result is also compared with the FP64 oracle and the error printed.
"""
import numpy as np
import pytest

from meraxes_amd import params as P
from meraxes_amd import reion, synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def compare(g, o, N, label):
    ion_g = g["xH"] < 1e-5
    ion_o = o["xH"] < 1e-5
    flips = ion_g != ion_o
    nflip = int(flips.sum())
    frac = nflip / N**3
    margin = np.abs(o["max_ratio"].astype(np.float64) - 1.0)
    worst = float(margin[flips].max()) if nflip else 0.0
    dxh = abs(g["volume_weighted_global_xH"] - o["volume_weighted_global_xH"])
    dmass = abs(g["mass_weighted_global_xH"] - o["mass_weighted_global_xH"])
    rb_mis = (g["r_bubble"] != o["r_bubble"]) & ~flips
    rb_margin = float(margin[rb_mis].min()) if rb_mis.any() else 0.0
    print(f"{label}: <xH>_V gpu {g['volume_weighted_global_xH']:.8f} oracle {o['volume_weighted_global_xH']:.8f} "
          f"|d|={dxh:.2e} |d mass|={dmass:.2e} flips {nflip} ({frac:.2e}) worst margin {worst:.2e} "
          f"r_bubble mismatches {int(rb_mis.sum())} partial-xH max|d| "
          f"{np.abs(g['xH'][~ion_g & ~ion_o] - o['xH'][~ion_g & ~ion_o]).max() if (~ion_g & ~ion_o).any() else 0:.2e}",
          flush=True)
    return dxh, dmass, frac, worst, flips


def run_pair(ctx, p, d, s, w, xe=None, z_in=None, J21_in=None):
    o = O.find_hii_bubbles(p, d, s, w, x_e=xe, z_at_ion=z_in, J21_at_ion=J21_in, want_ratio=True)
    g = reion.find_hii_bubbles(ctx, p, d, s, w, x_e=xe,
                               z_at_ionization=None if z_in is None else z_in.copy(),
                               J_21_at_ionization=None if J21_in is None else J21_in.copy())
    return g, o


def check_parity(g, o, N, label, eff=20.0, neutral_vs_fp64=False):
    dxh, dmass, frac, worst, flips = compare(g, o, N, label)
    assert g["n_radii"] == o["n_radii"]
    assert dxh <= 1e-5
    assert dmass <= 1e-5
    assert frac <= 1e-4
    assert worst <= 1e-4
    # cells that agree on ionisation agree on everything the reference derives
    # from it: a crossing sets xH = 0 (find_HII_bubbles.c:317-443); a cell that
    # never crosses takes the last step's n_f - eff f_coll, which may fall below
    # 1e-5 without being 0 and then carries the transforms' float noise
    keep = ~flips
    ion = (o["xH"] < 1e-5) & keep
    crossed = ion & (o["r_bubble"] > 0) & (g["r_bubble"] > 0)
    assert (g["xH"][crossed] == 0).all() and (o["xH"][crossed] == 0).all()
    assert np.abs(g["xH"][ion].astype(np.float64) - o["xH"][ion]).max(initial=0.0) < 1e-5
    neutral = keep & ~ion
    # partial xH = n_f - eff * f_coll on the unfiltered last step: the float
    # transforms' ~1e-7 (of the field's rms) error is multiplied by eff (larger
    # boxes check these cells against the FP64 oracle instead: neutral_vs_fp64)
    if not neutral_vs_fp64:
        assert np.allclose(g["xH"][neutral], o["xH"][neutral], rtol=0, atol=2.5e-6 * eff)
    zi = keep & (o["z_at_ionization"] >= 0)
    assert np.array_equal(g["z_at_ionization"][zi], o["z_at_ionization"][zi])
    return flips


@pytest.mark.parametrize("ftype", [0, 1, 2])
def test_config1_64(ctx, ftype):
    """64^3, L=100 Mpc/h, z=8, eps=20, seed 0 (the reference's CPU-runnable case)."""
    N = 64
    d, s, w = synth.snapshot_inputs(N, 0)
    p = P.hii_params(N, 100.0, 8.0, filter_type=ftype, efficiency=20.0)
    g, o = run_pair(ctx, p, d, s, w)
    check_parity(g, o, N, f"C1 type {ftype}")
    o64 = O.find_hii_bubbles(p, d, s, w, precision=64, want_ratio=True)
    compare(g, o64, N, f"C1 type {ftype} vs ref64")


def test_uvb_and_carried_state(ctx):
    """J_21 / J_21_at_ionization with ReionUVBFlag and a z_at_ionization carried in."""
    N = 64
    d, s, w = synth.snapshot_inputs(N, 4)
    p = P.hii_params(N, 100.0, 9.0, filter_type=1, efficiency=60.0, uvb_flag=1)
    rng = np.random.default_rng(1)
    z_in = np.where(rng.random((N, N, N)) < 0.01, np.float32(10.0), np.float32(-1.0)).astype(np.float32)
    J_in = np.where(z_in > 0, np.float32(0.5), np.float32(0.0)).astype(np.float32)
    g, o = run_pair(ctx, p, d, s, w, z_in=z_in, J21_in=J_in)
    flips = check_parity(g, o, N, "UVB", eff=60.0)
    keep = ~flips
    both = keep & (o["xH"] < 1e-5)
    # J_21 is the filtered SFR at the first-crossing radius: float transform
    # error relative to the field's scale, large relative to small values
    dJ = np.abs(g["J_21"][both].astype(np.float64) - o["J_21"][both])
    scale = float(np.sqrt(np.mean(o["J_21"][both].astype(np.float64) ** 2)))
    rel = dJ / np.maximum(np.abs(o["J_21"][both]), 1e-30)
    print(f"J_21 rel diff max {rel.max() if rel.size else 0:.2e}, max |d|/rms {dJ.max() / scale:.2e}", flush=True)
    assert (dJ <= 1e-4 * np.abs(o["J_21"][both]) + 1e-6 * scale).mean() > 0.999
    assert abs(g["volume_weighted_global_J_21"] - o["volume_weighted_global_J_21"]) <= 1e-4 * abs(
        o["volume_weighted_global_J_21"]) + 1e-30
    carried = z_in > 0
    assert np.array_equal(g["z_at_ionization"][carried], z_in[carried])
    assert np.array_equal(g["J_21_at_ionization"][carried], J_in[carried])


def test_spin_temperature_field(ctx):
    """x_e filtered as a fourth field (Flag_IncludeSpinTemp, find_HII_bubbles.c:144-152)."""
    N = 64
    d, s, w = synth.snapshot_inputs(N, 8)
    rng = np.random.default_rng(2)
    xe = synth.to_padded((0.02 + 0.05 * rng.random((N, N, N))).astype(np.float32))
    p = P.hii_params(N, 100.0, 8.0, efficiency=40.0, include_spin_temp=1)
    g, o = run_pair(ctx, p, d, s, w, xe=xe)
    check_parity(g, o, N, "x_e", eff=40.0)


def test_192_box_with_x_e(ctx):
    """A grid with a factor of 3 (192^3, 100 Mpc/h: 48 radii), top-hat, UVB on,
    x_e as the fourth field."""
    N = 192
    print(f"{N}^3 with x_e: oracle and GPU", flush=True)
    d, s, w = synth.snapshot_inputs(N, 9)
    rng = np.random.default_rng(5)
    xe = synth.to_padded((0.02 + 0.05 * rng.random((N, N, N))).astype(np.float32))
    p = P.hii_params(N, 100.0, 8.0, efficiency=30.0, uvb_flag=1, include_spin_temp=1)
    g, o = run_pair(ctx, p, d, s, w, xe=xe)
    assert g["n_radii"] == o["n_radii"] == 48
    check_parity(g, o, N, "192 x_e", eff=30.0)


def test_config2_like_128_sharpk(ctx):
    """Sharp-k with UVB on a 128^3, 200 Mpc/h box (the 256^3 case follows)."""
    N = 128
    d, s, w = synth.snapshot_inputs(N, 0)
    p = P.hii_params(N, 200.0, 9.0, filter_type=1, efficiency=20.0, uvb_flag=1)
    g, o = run_pair(ctx, p, d, s, w)
    check_parity(g, o, N, "C2-like 128")


def test_config2_256_sharpk(ctx):
    """The 256^3, 200 Mpc/h, z = 9 configuration itself: sharp-k, UVB on, J_21
    out, 44 radii.  The parity contract as in check_parity; on the partially
    neutral cells (n_f - eff f_coll from the float round trip of the
    unfiltered fields) the HIP path must be as close to the FP64 oracle as the
    float oracle is.  The two oracle runs take about a minute on the host."""
    N = 256
    print(f"config 2 at {N}^3: GPU, float and FP64 oracle", flush=True)
    d, s, w = synth.snapshot_inputs(N, 0)
    p = P.hii_params(N, 200.0, 9.0, filter_type=1, efficiency=20.0, uvb_flag=1)
    g, o = run_pair(ctx, p, d, s, w)
    o64 = O.find_hii_bubbles(p, d, s, w, precision=64)
    assert g["n_radii"] == o["n_radii"] == 44
    flips = check_parity(g, o, N, "C2 256", neutral_vs_fp64=True)
    keep = ~flips
    ion = (o["xH"] < 1e-5) & keep
    crossed = ion & (o["r_bubble"] > 0) & (g["r_bubble"] > 0)
    assert np.array_equal(g["r_bubble"][crossed], o["r_bubble"][crossed])
    neutral = (g["xH"] >= 1e-5) & (o["xH"] >= 1e-5) & (o64["xH"] >= 1e-5)
    e_g = float(np.abs(g["xH"][neutral].astype(np.float64) - o64["xH"][neutral]).max())
    e_32 = float(np.abs(o["xH"][neutral].astype(np.float64) - o64["xH"][neutral]).max())
    print(f"C2 256 partial xH vs FP64: gpu {e_g:.2e}, float oracle {e_32:.2e}", flush=True)
    assert e_g <= 2 * e_32 + 1e-6
    # J_21 where both sides ionised the cell, as in test_uvb_and_carried_state
    dJ = np.abs(g["J_21"][ion].astype(np.float64) - o["J_21"][ion])
    scale = float(np.sqrt(np.mean(o["J_21"][ion].astype(np.float64) ** 2)))
    print(f"C2 256 J_21 max |d|/rms {dJ.max() / scale:.2e}", flush=True)
    assert (dJ <= 1e-4 * np.abs(o["J_21"][ion]) + 1e-6 * scale).mean() > 0.999
    assert abs(g["volume_weighted_global_J_21"] - o["volume_weighted_global_J_21"]) <= 1e-4 * abs(
        o["volume_weighted_global_J_21"]) + 1e-30
    # the X-ray source grid, L_X = SFR x 1e40 erg/s per (Msun/yr), filtered with the
    # same ladder as a fourth field (parity check 1; it does not enter the
    # threshold): stored in units of 1e30 erg/s to stay inside float range, a
    # source grid's clamp (max(v, 0), zero below 1e-8)
    sfr_msun_yr = w.astype(np.float64) * (P.UNIT_MASS_G / P.units()["unit_time_s"]) / (P.SOLAR_MASS / 3.155e7)
    xray = (sfr_msun_yr * 1e40 / 1e30).astype(np.float32)
    R = P.radius_ladder(p)
    for k in (0, len(R) // 2, len(R) - 2, len(R) - 1):
        apply = 1 if k < len(R) - 1 else 0
        gx = reion.filtered_field(ctx, xray, p.box_size, 1, R[k], apply_filter=apply, clamp_kind=1)
        ox = O.filtered_field(N, p.box_size, 1, xray, R[k], apply_filter=apply, clamp_kind=1)
        a, b = gx[:, :, :N].astype(np.float64), ox[:, :, :N].astype(np.float64)
        e = np.linalg.norm(a - b) / np.linalg.norm(b)
        print(f"C2 256 X-ray field R[{k}] = {R[k]:.3f}: rel L2 {e:.2e}", flush=True)
        assert e <= 1e-5


def test_device_pointers_match_host(ctx):
    import torch

    N = 64
    d, s, w = synth.snapshot_inputs(N, 21)
    p = P.hii_params(N, 100.0, 8.0, efficiency=30.0, uvb_flag=1)
    h = reion.find_hii_bubbles(ctx, p, d, s, w)
    dd = [torch.from_numpy(a).cuda() for a in (d, s, w)]
    g = reion.find_hii_bubbles(ctx, p, *dd)
    torch.cuda.synchronize()
    for k in ("xH", "r_bubble", "J_21", "z_at_ionization", "J_21_at_ionization"):
        assert np.array_equal(g[k].cpu().numpy(), h[k]), k
    assert g["volume_weighted_global_xH"] == h["volume_weighted_global_xH"]


def test_invalid_arguments_raise(ctx):
    d, s, w = synth.snapshot_inputs(32, 1)
    p = P.hii_params(32, 50.0, 8.0, filter_type=7)
    with pytest.raises(reion.MrxError) as e:
        reion.find_hii_bubbles(ctx, p, d, s, w)
    assert e.value.code == -5
    p = P.hii_params(40, 50.0, 8.0)
    with pytest.raises(reion.MrxError) as e:
        reion.find_hii_bubbles(ctx, p, d, s, w)
    assert e.value.code == -4
