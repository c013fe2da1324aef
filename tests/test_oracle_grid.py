"""Oracle: galaxy gridding (construct_baryon_grids, reionization.c:1353-1615).

Pinned by closed forms and designed cases rather than by restating the loop:
nearest-grid-point placement (misc_tools.c:163-173: ties to even, the wrap of
an index above N-1, the rejected negative index), the float accumulation in
galaxy order (a cell where adding the small values first or last gives
different floats), dead galaxies, the black-hole terms and their count, the
clamp of a negative cell, the SFR timescale (ReionSfrTimescale / H(z),
virial_properties.c:79-93), the history ring's shift and the zero padding.
"""
import math

import numpy as np
import pytest

from meraxes_amd import params as P
from oracle import oracle as O


def one(p, pos, fesc=1.0, gross=1.0, bhm=0.0, ebhm=0.0, type_=0, hist=None):
    pos = np.asarray(pos, np.float32).reshape(-1, 3)
    n = len(pos)
    arr = (lambda v: np.full(n, v, np.float64) if np.isscalar(v) else np.asarray(v, np.float64))
    return O.construct_baryon_grids(p, pos, np.full(n, type_, np.int32), arr(fesc), arr(gross), arr(bhm), arr(ebhm),
                                    hist)


def cells(grid, N):
    nz = np.argwhere(grid[:, :, :N] != 0)
    return [tuple(int(v) for v in c) for c in nz]


def test_nearest_grid_point_ties_to_even_and_wraps():
    N, L = 8, 8.0  # one cell per Mpc/h: pos/L*N = pos
    p = P.grid_params(N, L, 8.0, bh_feedback=0)
    # x: 2.5 -> 2 (tie to even), y: 3.5 -> 4, z: 7.5 -> 8 -> wraps to 0
    r = one(p, [2.5, 3.5, 7.5])
    assert cells(r["stars"], N) == [(2, 4, 0)]
    # just above / below a half: 2.5000002 -> 3, 0.4999 -> 0; L itself wraps to 0
    r = one(p, [2.5000002, 0.4999, 8.0])
    assert cells(r["stars"], N) == [(3, 0, 0)]
    # a tiny negative coordinate rounds to (-)0 and is kept
    r = one(p, [-0.3, 1.0, 1.0])
    assert cells(r["stars"], N) == [(0, 1, 1)]


def test_negative_index_is_rejected():
    p = P.grid_params(8, 8.0, 8.0, bh_feedback=0)
    with pytest.raises(ValueError):
        one(p, [-0.6, 1.0, 1.0])
    # a dead galaxy is not placed, so its position is not checked
    r = one(p, [-5.0, 1.0, 1.0], type_=3)
    assert r["n_dead"] == 1 and r["n_gridded"] == 0 and not r["stars"].any()


def test_random_placement_matches_rint():
    rng = np.random.default_rng(3)
    N, L = 16, 37.0
    p = P.grid_params(N, L, 8.0, bh_feedback=0)
    pos = rng.uniform(0, L, size=(300, 3)).astype(np.float32)
    fesc = 2.0 ** rng.integers(-4, 4, size=300)  # exact in float, sums exact while small
    r = one(p, pos, fesc=fesc)
    idx = np.rint(pos.astype(np.float64) / L * N).astype(np.int64)
    idx[idx > N - 1] = 0
    expect = np.zeros((N, N, N))
    np.add.at(expect, (idx[:, 0], idx[:, 1], idx[:, 2]), fesc)
    assert np.array_equal(r["stars"][:, :, :N], expect.astype(np.float32))
    assert not r["stars"][:, :, N:].any()  # padding stays zero


def test_accumulation_is_sequential_float_in_galaxy_order():
    N, L = 4, 4.0
    p = P.grid_params(N, L, 8.0, bh_feedback=0)
    small = [1e-8] * 100
    # 1 first: every later addition is below half an ulp of 1 and is lost
    r = one(p, [[1.0, 1.0, 1.0]] * 101, fesc=[1.0] + small)
    assert r["stars"][1, 1, 1] == np.float32(1.0)
    # small values first: they accumulate before the 1 is added
    r = one(p, [[1.0, 1.0, 1.0]] * 101, fesc=small + [1.0])
    acc = np.float32(0.0)
    for v in small + [1.0]:
        acc = np.float32(np.float64(acc) + v)
    assert r["stars"][1, 1, 1] == acc and acc > np.float32(1.0)


def test_dead_galaxies_black_holes_and_clamp():
    N, L = 4, 4.0
    p = P.grid_params(N, L, 8.0, bh_feedback=1, bh_mass_limit_reion=10.0, alpha_uv_bh=1.5, alpha_uv=6.0)
    pos = [[0, 0, 0], [0, 0, 0], [1, 0, 0], [2, 0, 0]]
    r = one(p, pos, fesc=[1.0, 2.0, 4.0, -3.0], bhm=[20.0, 5.0, 10.0, 1.0], ebhm=[8.0, 8.0, 16.0, 1.0],
            type_=0)
    # cell (0,0,0): 1 + 8 (BH above limit) + 2 (BH below limit: not added)
    assert r["stars"][0, 0, 0] == np.float32(11.0)
    # cell (1,0,0): the limit itself counts (>=)
    assert r["stars"][1, 0, 0] == np.float32(20.0)
    # cell (2,0,0): negative total is clamped to 0 in both grids
    assert r["stars"][2, 0, 0] == 0.0 and r["weighted_sfr"][2, 0, 0] == 0.0
    assert r["n_bh_below_limit"] == 2
    # weighted_sfr adds EffectiveBHM * alpha_uv_bh / alpha_uv, then divides by the timescale
    H = p.hubble * math.sqrt(p.omega_m * 9.0**3 + p.omega_lambda)
    t = p.reion_sfr_timescale / H
    w00 = np.float32(np.float32(1.0 + 8.0 * 1.5 / 6.0) + 2.0)
    assert r["weighted_sfr"][0, 0, 0] == np.float32(np.float64(w00) / t)
    # type 3 is dead, type 2 is kept
    r = one(p, [[3, 3, 3], [3, 3, 3]], fesc=[1.0, 2.0], bhm=20.0, ebhm=0.0, type_=0)
    r2 = O.construct_baryon_grids(p, np.full((2, 3), 3, np.float32), np.array([2, 3], np.int32),
                                  np.array([1.0, 2.0]), None, np.full(2, 20.0), np.zeros(2))
    assert r2["stars"][3, 3, 3] == 1.0 and r2["n_dead"] == 1 and r2["n_gridded"] == 1
    assert r["stars"][3, 3, 3] == 3.0


def test_sfr_timescale_and_history_ring():
    N, L, z = 4, 4.0, 7.0
    ns = 3
    p = P.grid_params(N, L, z, include_spin_temp=1, n_store_sfr=ns, bh_feedback=0, reion_sfr_timescale=0.5)
    hist = np.stack([np.full((N, N, N + 2), 10.0 * (s + 1), np.float32) for s in range(ns)])
    r = one(p, [[1, 2, 3]], fesc=1.0, gross=6.0, hist=hist)
    # H(z) in internal units: HUBBLE [h/s] * UnitTime_in_s * E(z)
    u = P.units()
    E = math.sqrt(P.OMEGA_M * (1 + z) ** 3 + P.OMEGA_L)
    t = 0.5 / (P.HUBBLE * u["unit_time_s"] * E)
    assert r["sfr"][1, 2, 3] == np.float32(6.0 / t)
    h = r["sfr_histories"]
    # slot s+1 <- slot s; the oldest slot is dropped
    assert (h[1] == 10.0).all() and (h[2] == 20.0).all()
    # slot 0: the new sfr at every real cell, its padding untouched
    assert np.array_equal(h[0][:, :, :N], r["sfr"][:, :, :N])
    assert (h[0][:, :, N:] == 10.0).all()
    assert not r["sfr"][:, :, N:].any()
