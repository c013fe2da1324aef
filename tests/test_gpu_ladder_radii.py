"""The bubble ladder's multi-radius x pass is the single-radius one, bit for bit.

With tabulated windows (top-hat, Gaussian) the inverse x pass covers several
radii per launch (MRX_LADDER_RADII, read when a context is created, 2 by
default).  Each radius's spectra are those of the one-radius pass, so every
output grid and global must be identical for 1, 3 (groups that do not divide
the ladder) and 4 radii per pass, from 64^3 to the full 512^3.
"""
import os

import numpy as np
import pytest

from meraxes_amd import params as P
from meraxes_amd import reion, synth
from meraxes_amd._lib import Context

pytestmark = pytest.mark.gpu


def run_with(radii, p, d, s, w, xe):
    old = os.environ.get("MRX_LADDER_RADII")
    os.environ["MRX_LADDER_RADII"] = str(radii)
    try:
        c = Context(0)
    finally:
        if old is None:
            del os.environ["MRX_LADDER_RADII"]
        else:
            os.environ["MRX_LADDER_RADII"] = old
    try:
        return reion.find_hii_bubbles(c, p, d, s, w, x_e=xe)
    finally:
        c.close()


@pytest.mark.parametrize("N,filt,spin", [(64, 0, 0), (64, 2, 1), (128, 0, 1), (512, 0, 1)])
def test_radii_per_pass_do_not_change_results(ctx, N, filt, spin):
    print(f"ladder radii N={N} filter={filt} spin={spin}", flush=True)
    d, s, w = synth.snapshot_inputs(N, 17)
    xe = None
    if spin:
        rng = np.random.default_rng(3)
        xe = synth.to_padded((0.02 + 0.05 * rng.random((N, N, N))).astype(np.float32))
    p = P.hii_params(N, 100.0, 8.0, filter_type=filt, uvb_flag=1, include_spin_temp=spin)
    ref = run_with(1, p, d, s, w, xe)
    for k in (3, 4):
        got = run_with(k, p, d, s, w, xe)
        for name in ("xH", "r_bubble", "J_21", "z_at_ionization", "J_21_at_ionization"):
            assert np.array_equal(got[name], ref[name]), (k, name)
        for name in ("volume_weighted_global_xH", "mass_weighted_global_xH", "volume_weighted_global_J_21"):
            assert got[name] == ref[name], (k, name)
