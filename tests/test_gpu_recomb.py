"""GPU: the bubble finder with inhomogeneous recombinations vs the oracle.

N_rec is filtered as one more field (four fields without, five with the spin
temperature's x_e), the threshold becomes n_f (1 + N_rec_R / (1 + delta_R)),
Gamma12 is written at a cell's first crossing and N_rec is updated after the
ladder from the splined recombination rate.  The parity bar of the plain
bubble finder applies (flips only within 1e-4 of the threshold); Gamma12 and
N_rec come from float-filtered fields, so they are compared at 1e-4 relative
(+ 1e-6 of their rms) on the cells whose ionisation agrees.  Two snapshots in a
row carry N_rec and Gamma12 as the reference's state.
"""
import numpy as np
import pytest

from meraxes_amd import params as P
from meraxes_amd import reion, synth
from oracle import oracle as O

from test_gpu_hii import check_parity

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def nrec_field(N, seed, scale=0.3):
    rng = np.random.default_rng(seed)
    g = np.fft.irfftn(np.fft.rfftn(rng.normal(0, 1, (N, N, N))) *
                      np.exp(-0.5 * (np.fft.fftfreq(N)[:, None, None] ** 2 + np.fft.fftfreq(N)[None, :, None] ** 2 +
                                     np.fft.rfftfreq(N)[None, None, :] ** 2) * (2 * np.pi * 3) ** 2),
                      s=(N, N, N), axes=(0, 1, 2))
    out = np.zeros((N, N, N + 2), np.float32)
    out[:, :, :N] = np.abs(g) / np.abs(g).std() * scale
    return out


def close(a, b, what, keep):
    a = np.asarray(a, np.float64)[keep]
    b = np.asarray(b, np.float64)[keep]
    rms = np.sqrt((b ** 2).mean()) if b.size else 0.0
    ok = np.abs(a - b) <= 1e-4 * np.abs(b) + 1e-6 * rms
    print(f"{what}: {ok.mean():.6f} of {b.size} cells within 1e-4", flush=True)
    assert ok.mean() >= 0.999, what


@pytest.mark.parametrize("spin,on_dev", [(0, False), (1, True)])
def test_recombinations_two_snapshots(ctx, spin, on_dev):
    N, L = 64, 100.0
    print(f"recombinations N={N} spin={spin} on_dev={on_dev}", flush=True)
    d, s, w = synth.snapshot_inputs(N, 11)
    xe = None
    if spin:
        xe = np.zeros_like(d)
        xe[:, :, :N] = np.random.default_rng(3).uniform(0.0, 0.2, (N, N, N))
    nrec0 = nrec_field(N, 5, 0.05)
    g12 = np.zeros((N, N, N), np.float32)
    state_o = {"N_rec": nrec0.copy(), "Gamma12": g12.copy(), "z": None}
    state_g = {"N_rec": nrec0.copy(), "Gamma12": g12.copy(), "z": None}
    if on_dev:
        state_g = {k: (torch.from_numpy(v).cuda() if v is not None else None) for k, v in state_g.items()}
    for snap, (z, zprev) in enumerate(((8.5, 8.8), (8.2, 8.5))):
        p = P.hii_params(N, L, z, efficiency=150.0, uvb_flag=1, include_spin_temp=spin, include_recombinations=1,
                         prev_redshift=zprev, r_bubble_max=33.9)
        o = O.find_hii_bubbles(p, d, s, w, x_e=xe, z_at_ion=state_o["z"], want_ratio=True, N_rec=state_o["N_rec"],
                               Gamma12=state_o["Gamma12"])
        if on_dev:
            dev = [torch.from_numpy(a).cuda() for a in (d, s, w)]
            dxe = torch.from_numpy(xe).cuda() if xe is not None else None
            gr = reion.find_hii_bubbles(ctx, p, *dev, x_e=dxe, z_at_ionization=state_g["z"], N_rec=state_g["N_rec"],
                                        Gamma12=state_g["Gamma12"])
            torch.cuda.synchronize()
            state_g["z"] = gr["z_at_ionization"]
            g = {k: (v.cpu().numpy() if hasattr(v, "cpu") else v) for k, v in gr.items()}
        else:
            g = reion.find_hii_bubbles(ctx, p, d, s, w, x_e=xe, z_at_ionization=state_g["z"], N_rec=state_g["N_rec"],
                                       Gamma12=state_g["Gamma12"])
            state_g["z"] = g["z_at_ionization"]
        flips = check_parity(g, o, N, f"recomb spin={spin} snapshot {snap}", eff=150.0)
        keep = ~flips
        # Gamma12: cells that crossed (in both) during this call
        crossed = keep & (o["Gamma12"] != state_o["Gamma12"])
        assert crossed.sum() > 50
        close(g["Gamma12"], o["Gamma12"], f"Gamma12 snapshot {snap}", crossed)
        close(g["N_rec"][:, :, :N], o["N_rec"][:, :, :N], f"N_rec snapshot {snap}", keep)
        assert (np.asarray(g["N_rec"])[:, :, :N] >= nrec0[:, :, :N] - 1e-7).all()
        state_o = {"N_rec": o["N_rec"], "Gamma12": o["Gamma12"], "z": o["z_at_ionization"]}
        if not on_dev:
            state_g["N_rec"], state_g["Gamma12"] = g["N_rec"], g["Gamma12"]


def test_recombinations_need_their_grids_and_no_batch(ctx):
    N = 32
    d, s, w = synth.snapshot_inputs(N, 2)
    p = P.hii_params(N, 50.0, 8.0, include_recombinations=1, prev_redshift=8.2)
    with pytest.raises(ValueError):
        reion.find_hii_bubbles(ctx, p, d, s, w)
    dd = [torch.from_numpy(a).cuda() for a in (d, s, w)]
    with pytest.raises(reion.MrxError) as e:
        reion.HiiBatch(ctx, [{"params": p, "deltax": dd[0], "stars": dd[1], "weighted_sfr": dd[2]}])
    assert e.value.code == -4
