"""Pins the oracle's brightness temperature against the closed form per cell
(BrightnessTemperature.c:49-124: dTb = C(z) xH (1+delta) (1 - T_CMB(1+z)/T_S))."""
import numpy as np
import pytest

from meraxes_amd import params as P
from meraxes_amd import synth
from oracle import oracle as O


@pytest.mark.parametrize("spin", [0, 1])
def test_closed_form(spin):
    N = 16
    rng = np.random.default_rng(4)
    d = synth.gaussian_field(N, 3)
    xH = rng.random((N, N, N)).astype(np.float32)
    TS = (3.0 + 50 * rng.random((N, N, N))).astype(np.float32)
    z = 9.0
    p = P.tb_params(N, z, include_spin_temp=spin)
    got, ave = O.brightness_temp(p, xH, synth.to_padded(d), TS if spin else None)
    h, om, ob = P.HUBBLE_H, P.OMEGA_M, P.OMEGA_M * P.BARYON_FRAC
    C = 27.0 * (ob * h * h / 0.023) * np.sqrt(0.15 / om / h / h * (1 + z) / 10.0)
    base = C * xH * (1.0 + d.astype(np.float64))
    ref, scale = base, np.abs(base)
    if spin:
        ref = base * (1 - P.TCMB * (1 + z) / TS)
        scale = np.abs(base) * (1 + P.TCMB * (1 + z) / TS)  # the Ts factor can cancel
    # const_factor and the intermediate are float in the reference: a few ulp of the scale
    assert np.all(np.abs(got - ref) <= 4e-7 * scale + 1e-12)
    assert np.isclose(ave, got.astype(np.float64).mean(), rtol=1e-12)


def test_neutral_mean_density_value():
    N = 16
    p = P.tb_params(N, 8.0)
    got, ave = O.brightness_temp(p, np.ones((N, N, N), np.float32), synth.to_padded(np.zeros((N, N, N), np.float32)))
    h, om, ob = P.HUBBLE_H, P.OMEGA_M, P.OMEGA_M * P.BARYON_FRAC
    C = 27.0 * (ob * h * h / 0.023) * np.sqrt(0.15 / om / h / h * 9.0 / 10.0)
    assert np.allclose(got, C, rtol=1e-6)
