"""GPU: galaxy gridding (mrx_construct_baryon_grids) against the oracle.

Bit-identical stars / weighted_sfr / sfr grids and SFR history ring for
catalogues with many galaxies per cell (including cells of a thousand
galaxies, where the float accumulation order decides the last bits), dead
galaxies, black holes either side of the mass limit, with and without the
spin-temperature grids, from device and from host memory; a 512^3 catalogue
of four million galaxies; the error path of a negative grid index, which
leaves the outputs untouched.
"""
import numpy as np
import pytest

from meraxes_amd import params as P
from meraxes_amd import reion
from oracle import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def catalogue(n, L, seed, clustered=True):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0, L, size=(n, 3)).astype(np.float32)
    if clustered:
        # a tenth of the galaxies in 16 clumps of a few cells
        k = n // 10
        centres = rng.uniform(0, L, size=(16, 3))
        pos[:k] = (centres[rng.integers(0, 16, k)] + rng.normal(0, L / 400, size=(k, 3))) % L
        pos[:k] = np.clip(pos[:k], 0, np.nextafter(np.float32(L), 0))
    type_ = rng.choice(np.array([0, 1, 2, 3], np.int32), size=n, p=[0.5, 0.3, 0.15, 0.05])
    fesc = rng.lognormal(-3, 2, n)
    gross = rng.lognormal(-2, 2, n)
    bhm = rng.lognormal(-4, 1.5, n)
    ebhm = rng.lognormal(-6, 2, n)
    return pos, type_.astype(np.int32), fesc, gross, bhm, ebhm


def dev(*arrs):
    return [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in arrs]


@pytest.mark.parametrize("spin,bh,on_dev", [(1, 1, True), (0, 1, True), (1, 0, False), (0, 0, False)])
def test_gridding_bit_identical_to_oracle(ctx, spin, bh, on_dev):
    N, L, ns = 64, 100.0, 4
    print(f"gridding N={N} spin={spin} bh={bh} on_dev={on_dev}", flush=True)
    p = P.grid_params(N, L, 8.0, include_spin_temp=spin, n_store_sfr=ns, bh_feedback=bh,
                      bh_mass_limit_reion=np.exp(-4.0))
    pos, type_, fesc, gross, bhm, ebhm = catalogue(300_000, L, 1)
    # one cell with a thousand galaxies in a fixed random order
    pos[:1000] = np.float32([50.0, 25.0, 75.0])
    shape = (N, N, N + 2)
    hist = np.random.default_rng(2).random((ns,) + shape).astype(np.float32) if spin else None
    o = O.construct_baryon_grids(p, pos, type_, fesc, gross, bhm, ebhm, hist)
    if on_dev:
        g = reion.construct_baryon_grids(ctx, p, *dev(pos, type_, fesc, gross, bhm, ebhm),
                                         sfr_histories=dev(hist)[0] if spin else None)
        torch.cuda.synchronize()
        get = (lambda t: t.cpu().numpy())
    else:
        hh = hist.copy() if spin else None
        g = reion.construct_baryon_grids(ctx, p, pos, type_, fesc, gross, bhm, ebhm, sfr_histories=hh)
        get = (lambda a: a)
    for k in ("stars", "weighted_sfr") + (("sfr", "sfr_histories") if spin else ()):
        a, b = get(g[k]), o[k]
        assert a.shape == b.shape, k
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (k, int((a != b).sum()))
    for k in ("n_gridded", "n_dead", "n_bh_below_limit"):
        assert g[k] == o[k], k
    assert o["stars"][32, 16, 48] > 0  # the crowded cell


def test_gridding_512_four_million_galaxies(ctx):
    N, L = 512, 400.0
    print("gridding 512^3, 4M galaxies", flush=True)
    p = P.grid_params(N, L, 8.4, include_spin_temp=1, n_store_sfr=2, bh_feedback=1)
    pos, type_, fesc, gross, bhm, ebhm = catalogue(4_000_000, L, 5)
    hist = np.zeros((2, N, N, N + 2), np.float32)
    o = O.construct_baryon_grids(p, pos, type_, fesc, gross, bhm, ebhm, hist)
    dh = torch.zeros((2, N, N, N + 2), dtype=torch.float32, device="cuda")
    g = reion.construct_baryon_grids(ctx, p, *dev(pos, type_, fesc, gross, bhm, ebhm), sfr_histories=dh)
    torch.cuda.synchronize()
    for k in ("stars", "weighted_sfr", "sfr", "sfr_histories"):
        a = g[k].cpu().numpy()
        assert np.array_equal(a.view(np.uint32), o[k].view(np.uint32)), k
    # size-independent: the stars grid holds every kept galaxy's mass (float sums, double total)
    kept = type_ <= 2
    total = float(np.sum(fesc[kept]) + np.sum(ebhm[kept]))
    assert abs(float(g["stars"].double().sum()) - total) <= 1e-5 * total


def test_negative_index_rejected_outputs_untouched(ctx):
    N, L = 32, 50.0
    p = P.grid_params(N, L, 8.0, bh_feedback=0)
    pos, type_, fesc, gross, bhm, ebhm = catalogue(1000, L, 3, clustered=False)
    type_[:] = 0
    pos[517] = np.float32([-2.0, 1.0, 1.0])
    stars = torch.full((N, N, N + 2), 7.0, device="cuda")
    wsfr = torch.full((N, N, N + 2), 7.0, device="cuda")
    with pytest.raises(reion.MrxError) as e:
        reion.construct_baryon_grids(ctx, p, *dev(pos, type_, fesc), stars=stars, weighted_sfr=wsfr)
    assert e.value.code == -1 and "galaxy 517" in str(e.value)
    torch.cuda.synchronize()
    assert bool((stars == 7.0).all()) and bool((wsfr == 7.0).all())


def test_no_galaxies_zero_grids_and_shifted_ring(ctx):
    N, ns = 16, 3
    p = P.grid_params(N, 20.0, 8.0, include_spin_temp=1, n_store_sfr=ns)
    hist = torch.stack([torch.full((N, N, N + 2), float(s + 1), device="cuda") for s in range(ns)])
    empty = [torch.empty((0, 3), dtype=torch.float32, device="cuda"), torch.empty(0, dtype=torch.int32, device="cuda")]
    f0 = torch.empty(0, dtype=torch.float64, device="cuda")
    g = reion.construct_baryon_grids(ctx, p, empty[0], empty[1], f0, f0, sfr_histories=hist)
    torch.cuda.synchronize()
    assert not bool(g["stars"].any()) and not bool(g["sfr"].any())
    assert bool((hist[1] == 1.0).all()) and bool((hist[2] == 2.0).all())
    assert bool((hist[0][:, :, :N] == 0).all()) and bool((hist[0][:, :, N:] == 1.0).all())
