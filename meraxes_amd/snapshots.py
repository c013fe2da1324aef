"""A run of snapshots through the reionization path: the caller-side state the
reference keeps in run_globals between calls (SFR history ring, collapse
fraction history, x_e / T_K / z_at_ionization), driven in the reference's order
per snapshot: spin temperature, then the bubble finder, then the brightness
temperature (src/core/dracarys.c:352-390, reionization.c:244-330).

Synthetic inputs follow the full-snapshot configuration: one seeded k^-2 field
grown linearly as 1/(1+z), sources from the same generator.
"""
import numpy as np

from . import params as P
from . import synth


class SyntheticRun:
    """Host-side generator of per-snapshot input grids (padded, float32)."""

    def __init__(self, N, box_size, seed=0, redshifts=None, n_history=None):
        self.N = N
        self.L = box_size
        self.seed = seed
        self.z = P.snapshot_redshifts() if redshifts is None else np.asarray(redshifts, np.float64)
        self.d0 = synth.gaussian_field(N, seed)
        self.n_history = n_history or len(self.z)

    def inputs(self, snap):
        z = self.z[snap]
        g = np.float32((1 + self.z[0]) / (1 + z))
        d = np.maximum(self.d0 * g, np.float32(-0.99)).astype(np.float32)
        stars, sfr = synth.source_grids(d, self.seed + 1000 * snap)
        return synth.to_padded(d), synth.to_padded(stars), synth.to_padded(sfr)
