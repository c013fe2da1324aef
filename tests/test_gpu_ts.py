"""GPU parity of the spin-temperature step (ComputeTs.c:78-1020) against the oracle.

A short run of snapshots from z = 20 (the homogeneous RECFAST branch) through
the inhomogeneous branch with SFR histories, both sides carrying their own
state.  Tolerances: the smoothed SFR shells are float FFT outputs (rel L2
<= 1e-5 vs the oracle); x_e, T_K and T_S are FP64 evolutions of those shells
rounded to float, compared at <= 1e-4 relative per cell, averages at 1e-5.
"""
import numpy as np
import pytest

from meraxes_amd import params as P
from meraxes_amd import reion
from meraxes_amd.snapshots import SyntheticRun
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-30)


@pytest.mark.parametrize("N,L", [(32, 25.0), (64, 50.0), (128, 100.0)])
def test_ts_snapshot_sequence(ctx, N, L):
    run = SyntheticRun(N, L, seed=3)
    nh = 8
    hist = np.zeros((nh, N, N, 2 * (N // 2 + 1)), np.float32)
    fc_g = np.zeros(len(run.z))
    fc_o = np.zeros(len(run.z))
    xe_g = np.zeros((N, N, 2 * (N // 2 + 1)), np.float32)
    Tk_g = np.zeros((N, N, N), np.float32)
    xe_o, Tk_o = xe_g.copy(), Tk_g.copy()
    for snap in range(5):
        d, s, f = run.inputs(snap)
        hist = np.roll(hist, 1, axis=0)
        hist[0] = f
        pg = P.ts_params(N, L, snap, run.z, fc_g, nh)
        po = P.ts_params(N, L, snap, run.z, fc_o, nh)
        o = O.compute_ts(po, d, s, f, hist, xe_o, Tk_o)
        g = reion.compute_ts(ctx, pg, d, s, f, hist, xe_g, Tk_g)
        xe_o, Tk_o = o["x_e_box_prev"], o["Tk_box"]
        assert g["homogeneous"] == o["homogeneous"] == (snap == 0)
        rx = rel(g["x_e_box_prev"][:, :, :N], xe_o[:, :, :N])
        rT = rel(g["Tk_box"], Tk_o)
        rS = rel(g["TS_box"], o["TS_box"])
        print(f"N={N} snap={snap} z={run.z[snap]:.3f} fcoll gpu {g['collapse_fraction']:.6e} oracle "
              f"{o['collapse_fraction']:.6e} | x_e max rel {rx.max():.2e} T_K {rT.max():.2e} T_S {rS.max():.2e} "
              f"| <T_S> {g['volume_ave_TS']:.5f} vs {o['volume_ave_TS']:.5f}", flush=True)
        assert abs(g["collapse_fraction"] - o["collapse_fraction"]) <= 1e-12 + 1e-9 * abs(o["collapse_fraction"])
        assert fc_g[snap] == pytest.approx(fc_o[snap], rel=1e-9, abs=1e-15)
        assert rx.max() <= 1e-4 and rT.max() <= 1e-4 and rS.max() <= 1e-4
        for k in ("volume_ave_TS", "volume_ave_TK", "volume_ave_xe", "volume_ave_J_alpha", "volume_ave_Xheat"):
            assert abs(g[k] - o[k]) <= 1e-5 * abs(o[k]) + 1e-30, k
        assert np.array_equal(g["x_e_box"], g["x_e_box_prev"])


def test_ts_shells_match_oracle(ctx):
    """The smoothed SFR shells themselves (float FFT outputs) against the oracle's."""
    N, L = 32, 25.0
    run = SyntheticRun(N, L, seed=5)
    nh = 8
    hist = np.zeros((nh, N, N, 2 * (N // 2 + 1)), np.float32)
    for snap in range(3):
        d, s, f = run.inputs(snap)
        hist = np.roll(hist, 1, axis=0)
        hist[0] = f
    fc = np.zeros(len(run.z))
    fc[:2] = [1e-3, 2e-3]
    p = P.ts_params(N, L, 2, run.z, fc, nh)
    xe = np.full((N, N, 2 * (N // 2 + 1)), 2e-4, np.float32)
    Tk = np.full((N, N, N), 10.0, np.float32)
    o = O.compute_ts(p, d, s, f, hist, xe, Tk, want_shells=True)
    sh = o["shells"]
    assert np.isfinite(sh).all() and sh.max() > 0
    g = reion.compute_ts(ctx, P.ts_params(N, L, 2, run.z, fc.copy(), nh), d, s, f, hist, xe.copy(), Tk.copy())
    assert rel(g["TS_box"], o["TS_box"]).max() <= 1e-4
