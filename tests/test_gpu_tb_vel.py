"""GPU: delta_T with peculiar velocities (mrx_brightness_temp_vel) against the oracle.

The converted velocity grid is bit-identical (the same float/double steps).
delta_T for pec_vels 1, 2 and 3 is compared per cell: the velocity gradient
goes through two different float FFTs (GPU and oracle), so the bar is rel-L2
<= 1e-5 with every cell within 1e-4 relative + 1e-5 of the field's rms (the
re-binned sums of mixed sign cancel), and <delta_T> within 1e-6, for velocity
gradients below the
Hubble rate.  With gradients beyond it the optical depth is ill-conditioned
(|dv/dz / H + 1| near 0), and the re-binning is checked bit for bit on the
GPU's own input field instead.  Host and device pointers; 64^3 and 128^3.
"""
import numpy as np
import pytest

from meraxes_amd import params as P
from meraxes_amd import reion, synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


MPC = 3.086e24


def inputs(N, L, seed, trees, disp_rms=0.7, width=4.0):
    """xH, deltax, T_S and a smooth velocity field whose displacement v / H(z) has
    rms `disp_rms` cells over a Gaussian correlation width of `width` cells."""
    d, s, w = synth.snapshot_inputs(N, seed)
    rng = np.random.default_rng(seed + 100)
    xH = np.clip(rng.uniform(-0.2, 1.0, (N, N, N)), 0, 1).astype(np.float32)  # some fully ionised cells
    TS = rng.uniform(15.0, 300.0, (N, N, N)).astype(np.float32)
    gk = np.fft.rfftn(rng.normal(0, 1, (N, N, N)))
    k2 = (np.fft.fftfreq(N)[:, None, None] ** 2 + np.fft.fftfreq(N)[None, :, None] ** 2 +
          np.fft.rfftfreq(N)[None, None, :] ** 2)
    gk *= np.exp(-0.5 * k2 * (2 * np.pi * width) ** 2)
    D = np.fft.irfftn(gk, s=(N, N, N), axes=(0, 1, 2))
    D *= disp_rms / D.std()  # displacement in cells
    p = P.tb_vel_params(N, L, 8.0, trees_id=trees)
    z = float(np.float32(p.tb.redshift))
    Hz = 3.2407789e-18 * P.HUBBLE_H * np.sqrt(P.OMEGA_M * (1 + z) ** 3 + P.OMEGA_L)
    unit = {1: np.sqrt(1 / (1 + z)) / 1000.0, 0: 1 + z, 2: 1 + z}[trees] * 1e5 / MPC * (1 + z)
    vel = np.zeros((N, N, N + 2), np.float32)
    vel[:, :, :N] = D * ((L / P.HUBBLE_H) / N) * Hz / unit
    return xH, d, TS, vel


def check(g, o, what):
    g = np.asarray(g, np.float64)
    o = np.asarray(o, np.float64)
    rms = np.sqrt((o ** 2).mean())
    rel_l2 = np.sqrt(((g - o) ** 2).sum() / (o ** 2).sum())
    worst = np.abs(g - o) / (np.abs(o) + 0.1 * rms)  # floor: re-binned sums of mixed sign cancel
    print(f"{what}: rel-L2 {rel_l2:.2e}, max rel {worst.max():.2e}", flush=True)
    assert rel_l2 <= 1e-5 and worst.max() <= 1e-4, what
    assert (np.abs(g - o) <= 1e-4 * np.abs(o) + 1e-5 * rms).all(), what


def run_gpu(ctx, p, xH, d, TS, vel, on_dev):
    if on_dev:
        tv = torch.from_numpy(vel.copy()).cuda()
        dT, ave = reion.brightness_temp_vel(ctx, p, *(torch.from_numpy(a).cuda() for a in (xH, d, TS)), tv)
        torch.cuda.synchronize()
        return dT.cpu().numpy(), ave, tv.cpu().numpy()
    hv = vel.copy()
    dT, ave = reion.brightness_temp_vel(ctx, p, xH, d, TS, hv)
    return dT, ave, hv


@pytest.mark.parametrize("N,pec,trees,on_dev,disp", [(64, 1, 0, True, 0.7), (64, 2, 1, False, 0.7),
                                                     (64, 3, 0, True, 0.7), (128, 3, 2, True, 0.7),
                                                     (64, 3, 1, False, 0.7), (64, 3, 0, True, 0.2),
                                                     (128, 3, 1, True, 0.25)])
def test_tb_vel_matches_oracle(ctx, N, pec, trees, on_dev, disp):
    """Velocity gradients below the Hubble rate (|dv/dz| << H(z), the physical regime),
    displacements of 0.2-0.7 cells rms."""
    print(f"tb_vel N={N} pec={pec} trees={trees} on_dev={on_dev} disp={disp}", flush=True)
    L = 100.0 * N / 64
    xH, d, TS, vel = inputs(N, L, 7 + N + pec, trees, disp_rms=disp)
    p = P.tb_vel_params(N, L, 8.0, pec_vels=pec, trees_id=trees)
    o_dT, o_ave, o_vel = O.brightness_temp_vel(p, xH, d, TS, vel)
    g_dT, ave, g_vel = run_gpu(ctx, p, xH, d, TS, vel, on_dev)
    assert np.array_equal(g_vel.view(np.uint32), o_vel.view(np.uint32))
    check(g_dT, o_dT, f"delta_T pec={pec} N={N}")
    assert abs(ave - o_ave) <= 1e-6 * abs(o_ave) + 1e-9
    if pec == 3:
        # the re-binning moved signal between cells, and conserved each column's sum
        p2 = P.tb_vel_params(N, L, 8.0, pec_vels=2, trees_id=trees)
        o2, _, _ = O.brightness_temp_vel(p2, xH, d, TS, vel)
        assert not np.allclose(o2, o_dT)
        kept = (np.abs(o2) >= 1e-7) & (xH >= 1e-7)
        np.testing.assert_allclose(g_dT.astype(np.float64).sum(axis=2), np.where(kept, o2, 0).sum(axis=2),
                                   rtol=1e-4, atol=1e-5 * np.abs(o2).max())


def test_tb_vel_large_displacements_rebinning_exact(ctx):
    """Displacements of several cells and gradients beyond H(z): the optical depth's
    |dv/dz / H + 1| nears 0 in places, where the two float FFTs' rounding is
    amplified, so the re-binning is checked on the GPU's own pec_vels 2 field:
    the oracle's re-binning of it must equal the GPU's pec_vels 3 bit for bit."""
    N, L, trees = 64, 100.0, 0
    xH, d, TS, vel = inputs(N, L, 31, trees, disp_rms=3.0, width=1.5)
    p2 = P.tb_vel_params(N, L, 8.0, pec_vels=2, trees_id=trees)
    p3 = P.tb_vel_params(N, L, 8.0, pec_vels=3, trees_id=trees)
    g2, _, g_vel = run_gpu(ctx, p2, xH, d, TS, vel, True)
    g3, ave3, _ = run_gpu(ctx, p3, xH, d, TS, vel, True)
    o3 = O.rsd_los(p3, xH, g_vel, g2)
    assert np.array_equal(g3.view(np.uint32), o3.view(np.uint32)), int((g3 != o3).sum())
    moved = np.abs(g3 - np.where((np.abs(g2) >= 1e-7) & (xH >= 1e-7), g2, 0)).max()
    assert moved > 0.1 * np.abs(g2).max()  # multi-cell moves happened


def test_tb_vel_without_spin_temperature_is_plain_tb(ctx):
    N = 32
    xH, d, TS, vel = inputs(N, 50.0, 3, 1)
    p = P.tb_vel_params(N, 50.0, 8.0, pec_vels=3)
    p.tb.include_spin_temp = 0
    dev = [torch.from_numpy(a).cuda() for a in (xH, d, TS, vel)]
    dT, ave = reion.brightness_temp_vel(ctx, p, *dev)
    ref, ref_ave = reion.brightness_temp(ctx, P.tb_params(N, 8.0), dev[0], dev[1])
    torch.cuda.synchronize()
    assert torch.equal(dT, ref) and ave == ref_ave
    assert torch.equal(dev[3], torch.from_numpy(vel).cuda())  # velocities untouched


def test_tb_vel_rejects_bad_flag(ctx):
    N = 32
    xH, d, TS, vel = inputs(N, 50.0, 4, 1)
    p = P.tb_vel_params(N, 50.0, 8.0, pec_vels=4)
    with pytest.raises(reion.MrxError) as e:
        reion.brightness_temp_vel(ctx, p, xH, d, TS, vel)
    assert e.value.code == -1
