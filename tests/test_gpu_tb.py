"""GPU parity of the brightness-temperature grid (BrightnessTemperature.c:24-124, :512-545)."""
import numpy as np
import pytest

from meraxes_amd import params as P
from meraxes_amd import reion, synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("spin", [0, 1])
def test_brightness_temp_bitexact(ctx, spin):
    N = 64
    rng = np.random.default_rng(3)
    d = synth.to_padded(synth.gaussian_field(N, 2))
    xH = rng.random((N, N, N)).astype(np.float32)
    TS = (5.0 + 100.0 * rng.random((N, N, N))).astype(np.float32) if spin else None
    p = P.tb_params(N, 9.5, include_spin_temp=spin)
    g, ave_g = reion.brightness_temp(ctx, p, xH, d, TS)
    o, ave_o = O.brightness_temp(p, xH, d, TS)
    assert np.array_equal(g, o)
    assert abs(ave_g - ave_o) <= 1e-12 * abs(ave_o)
