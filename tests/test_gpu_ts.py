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


def shell_rel_l2(g, o):
    """Per-shell relative L2 error (absolute where the oracle's shell is all zero)."""
    out = []
    for c in range(o.shape[0]):
        oc = o[c].astype(np.float64)
        d = np.linalg.norm(g[c].astype(np.float64) - oc)
        n = np.linalg.norm(oc)
        out.append(d / n if n > 0 else d)
    return np.array(out)


@pytest.mark.parametrize("N,L,ftype,quirks", [(32, 25.0, 0, 0), (32, 25.0, 1, 1), (32, 25.0, 2, 0),
                                              (64, 50.0, 0, 1), (64, 50.0, 1, 0), (64, 50.0, 2, 0)])
def test_ts_shells_match_oracle(ctx, N, L, ftype, quirks):
    """The 40 smoothed SFR shells (ComputeTs.c:379-503, TsHeatingFilterType
    0 top-hat, 1 sharp-k, 2 Gaussian) over a 4-snapshot sequence: each shell
    within rel L2 1e-5 of the oracle's, then x_e, T_K, T_S per cell within 1e-4;
    with reproduce_quirks the volume-averaged x_e is the reference's stale-index
    read (ComputeTs.c:957), the last cell's x_e."""
    run = SyntheticRun(N, L, seed=5)
    nh = 8
    hist = np.zeros((nh, N, N, 2 * (N // 2 + 1)), np.float32)
    fc_g = np.zeros(len(run.z))
    fc_o = np.zeros(len(run.z))
    xe_g = np.full((N, N, 2 * (N // 2 + 1)), 2e-4, np.float32)
    Tk_g = np.full((N, N, N), 10.0, np.float32)
    xe_o, Tk_o = xe_g.copy(), Tk_g.copy()
    for snap in range(4):
        d, s, f = run.inputs(snap)
        hist = np.roll(hist, 1, axis=0)
        hist[0] = f
        kw = dict(heating_filter_type=ftype, reproduce_quirks=quirks)
        pg = P.ts_params(N, L, snap, run.z, fc_g, nh, **kw)
        po = P.ts_params(N, L, snap, run.z, fc_o, nh, **kw)
        print(f"shells case N={N} filter={ftype} quirks={quirks} snap={snap}", flush=True)
        o = O.compute_ts(po, d, s, f, hist, xe_o, Tk_o, want_shells=True)
        g = reion.compute_ts(ctx, pg, d, s, f, hist, xe_g, Tk_g)
        xe_o, Tk_o = o["x_e_box_prev"], o["Tk_box"]
        assert g["homogeneous"] == o["homogeneous"] == (snap == 0)
        if snap == 0:
            with pytest.raises(reion.MrxError):
                reion.ts_last_shells(ctx, N, 40)
            continue
        e = shell_rel_l2(reion.ts_last_shells(ctx, N, 40), o["shells"])
        rS = rel(g["TS_box"], o["TS_box"])
        print(f"  shell rel L2 max {e.max():.2e} (shell {int(e.argmax())}) | T_S max rel {rS.max():.2e} | "
              f"<x_e> gpu {g['volume_ave_xe']:.6e} oracle {o['volume_ave_xe']:.6e}", flush=True)
        assert e.max() <= 1e-5
        assert rS.max() <= 1e-4 and rel(g["Tk_box"], Tk_o).max() <= 1e-4
        assert rel(g["x_e_box_prev"][:, :, :N], xe_o[:, :, :N]).max() <= 1e-4
        assert abs(g["volume_ave_xe"] - o["volume_ave_xe"]) <= 1e-5 * abs(o["volume_ave_xe"])
        if quirks:
            assert g["volume_ave_xe"] == pytest.approx(float(g["x_e_box_prev"][-1, -1, N - 1]), rel=1e-12)


def test_ts_many_history_operations_per_shell(ctx):
    """Closely spaced snapshots (1000 from z = 20 to 8): a shell needs more
    SFR-history operations (ComputeTs.c:296-376) than one device launch takes
    (24), so they run in several launches, in plan order; shells and T_S
    against the oracle."""
    N, L = 32, 100.0
    z = np.linspace(20.0, 8.0, 1000)
    snap, nh = 700, 256
    rng = np.random.default_rng(11)
    from meraxes_amd import synth

    d, s, f = synth.snapshot_inputs(N, 4)
    hist = np.empty((nh, N, N, 2 * (N // 2 + 1)), np.float32)
    for k in range(nh):
        hist[k] = f * np.float32(rng.uniform(0.5, 1.5))
    hist[0] = f
    fc_g = np.linspace(0.0, 2e-3, len(z))
    fc_o = fc_g.copy()
    pg = P.ts_params(N, L, snap, z, fc_g, nh)
    po = P.ts_params(N, L, snap, z, fc_o, nh)
    plan_ops = np.bincount([op[0] for op in O.ts_host.host(po).plan().ops])
    assert plan_ops.max() > 24
    xe = np.full((N, N, 2 * (N // 2 + 1)), 2e-4, np.float32)
    Tk = np.full((N, N, N), 30.0, np.float32)
    print(f"many-ops case: max {plan_ops.max()} operations in a shell", flush=True)
    o = O.compute_ts(po, d, s, f, hist, xe, Tk, want_shells=True)
    g = reion.compute_ts(ctx, pg, d, s, f, hist, xe.copy(), Tk.copy())
    e = shell_rel_l2(reion.ts_last_shells(ctx, N, 40), o["shells"])
    rS = rel(g["TS_box"], o["TS_box"])
    print(f"  shell rel L2 max {e.max():.2e} | T_S max rel {rS.max():.2e}", flush=True)
    assert e.max() <= 1e-5 and rS.max() <= 1e-4
