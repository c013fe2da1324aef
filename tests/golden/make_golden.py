"""Writes the golden fixtures of tests/golden/ from the oracle.

The reference publishes no golden vectors for this path, so these fixtures are
the oracle's own outputs on seeded synthetic inputs: they pin the oracle
against regressions and give the GPU tests a fixed target that does not need
the oracle at run time.  Inputs are regenerated from the seeds (Philox), so
only outputs are stored.

Run from the repository root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from meraxes_amd import params as P  # noqa: E402
from meraxes_amd import synth  # noqa: E402
from meraxes_amd.snapshots import SyntheticRun  # noqa: E402
from oracle import oracle as O  # noqa: E402


def hii_case(N=32, L=50.0, z=8.0, ftype=0, eff=60.0, seed=0):
    d, s, w = synth.snapshot_inputs(N, seed)
    p = P.hii_params(N, L, z, filter_type=ftype, efficiency=eff, uvb_flag=1)
    o = O.find_hii_bubbles(p, d, s, w, want_ratio=True)
    return {k: o[k] for k in ("xH", "r_bubble", "J_21", "z_at_ionization", "J_21_at_ionization", "max_ratio")}, {
        "volume_weighted_global_xH": o["volume_weighted_global_xH"],
        "mass_weighted_global_xH": o["mass_weighted_global_xH"],
        "volume_weighted_global_J_21": o["volume_weighted_global_J_21"], "n_radii": o["n_radii"]}


def ts_case(N=32, L=25.0, nsnap=3, seed=3):
    run = SyntheticRun(N, L, seed=seed)
    nh = 8
    hist = np.zeros((nh, N, N, 2 * (N // 2 + 1)), np.float32)
    fc = np.zeros(len(run.z))
    xe = np.zeros((N, N, 2 * (N // 2 + 1)), np.float32)
    Tk = np.zeros((N, N, N), np.float32)
    out = {}
    for snap in range(nsnap):
        d, s, f = run.inputs(snap)
        hist = np.roll(hist, 1, axis=0)
        hist[0] = f
        o = O.compute_ts(P.ts_params(N, L, snap, run.z, fc, nh), d, s, f, hist, xe, Tk)
        xe, Tk = o["x_e_box_prev"], o["Tk_box"]
        out[f"TS_{snap}"] = o["TS_box"]
        out[f"Tk_{snap}"] = Tk
        out[f"xe_{snap}"] = xe[:, :, :N].copy()
    return out


def main():
    for ftype in (0, 1, 2):
        arrays, scal = hii_case(ftype=ftype)
        np.savez_compressed(os.path.join(HERE, f"hii_n32_type{ftype}.npz"), **arrays,
                            **{k: np.array(v) for k, v in scal.items()})
    np.savez_compressed(os.path.join(HERE, "ts_n32.npz"), **ts_case())
    print("written")


if __name__ == "__main__":
    main()
