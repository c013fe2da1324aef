"""Seeded synthetic input grids for the measured configurations.

deltax: Gaussian random field with P(k) proportional to k^-2, the k = 0 mode
zero, scaled to rms 0.6 per cell and clipped at -0.99.  Stellar mass:
1e8 Msun * exp(deltax) * Poisson(2)/2 in internal units (1e10 Msun/h).  SFR:
stellar mass / 1e8 yr in internal units (UnitTime_in_s = UnitLength/UnitVelocity,
src/core/init.c:188).  Random numbers come from numpy's counter-based Philox,
so a (seed, N) pair gives the same grids on any host.

Grids are returned in the reference's padded real layout (INDEX_PADDED,
src/core/misc_tools.c:206): shape (N, N, 2(N/2+1)), the padding zero.
"""
import numpy as np

from . import params as P


def rng(seed):
    return np.random.Generator(np.random.Philox(seed))


def _fft():
    try:
        import scipy.fft as sf

        return sf, {"workers": -1}
    except Exception:  # pragma: no cover - scipy is part of the image
        return np.fft, {}


def to_padded(real):
    """(N,N,N) -> (N,N,2(N/2+1)) float32, padding zero."""
    N = real.shape[0]
    out = np.zeros((N, N, 2 * (N // 2 + 1)), dtype=np.float32)
    out[:, :, :N] = real
    return out


def from_padded(padded):
    N = padded.shape[0]
    return np.ascontiguousarray(padded[:, :, :N])


def gaussian_field(N, seed, rms=0.6, clip=-0.99, growth=1.0):
    """k^-2 Gaussian random field on an N^3 grid, float32, unpadded."""
    g = rng(seed)
    w = g.standard_normal((N, N, N), dtype=np.float32)
    sf, kw = _fft()
    W = sf.rfftn(w, **kw)
    del w
    kx = np.fft.fftfreq(N, 1.0 / N).astype(np.float32)
    kz = np.arange(N // 2 + 1, dtype=np.float32)
    k2 = kx[:, None, None] ** 2 + kx[None, :, None] ** 2 + kz[None, None, :] ** 2
    k2[0, 0, 0] = 1.0
    W *= (1.0 / np.sqrt(k2)).astype(np.float32)
    W[0, 0, 0] = 0.0
    del k2
    d = sf.irfftn(W, s=(N, N, N), **kw).astype(np.float32)
    del W
    d *= np.float32(rms / float(np.sqrt(np.mean(d.astype(np.float64) ** 2))))
    d *= np.float32(growth)
    np.maximum(d, np.float32(clip), out=d)
    return d


def source_grids(deltax, seed):
    """(stars, sfr) in internal units from an unpadded deltax."""
    g = rng(seed + 7919)
    pois = g.poisson(2.0, size=deltax.shape).astype(np.float32)
    mstar_msun = np.float32(1e8) * np.exp(deltax) * pois * np.float32(0.5)
    stars = (mstar_msun * np.float32(P.HUBBLE_H / 1e10)).astype(np.float32)
    unit_time_s = P.units()["unit_time_s"]
    sfr_timescale = 1e8 * P.SEC_PER_YEAR / unit_time_s  # 1e8 yr in internal time units
    sfr = (stars / np.float32(sfr_timescale)).astype(np.float32)
    return stars, sfr


def snapshot_inputs(N, seed, growth=1.0):
    """Padded (deltax, stars, weighted_sfr) for one seeded box."""
    d = gaussian_field(N, seed, growth=growth)
    stars, sfr = source_grids(d, seed)
    return to_padded(d), to_padded(stars), to_padded(sfr)


def snapshot_inputs_torch(N, seed, device, growth=1.0):
    """Padded (deltax, stars, weighted_sfr) generated on the GPU with torch.

    The same recipe as snapshot_inputs (k^-2 field, rms 0.6, clip -0.99,
    1e8 Msun exp(delta) Poisson(2)/2, SFR over 1e8 yr) from torch's generator,
    so the numbers differ from the numpy stream; used where many boxes are
    needed quickly (the batched benchmark), never as a parity input.
    """
    import torch

    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    w = torch.randn((N, N, N), generator=g, device=device, dtype=torch.float32)
    W = torch.fft.rfftn(w)
    del w
    kx = torch.fft.fftfreq(N, 1.0 / N, device=device)
    kz = torch.arange(N // 2 + 1, device=device, dtype=torch.float32)
    k2 = kx[:, None, None] ** 2 + kx[None, :, None] ** 2 + kz[None, None, :] ** 2
    k2[0, 0, 0] = 1.0
    W *= torch.rsqrt(k2)
    W[0, 0, 0] = 0.0
    d = torch.fft.irfftn(W, s=(N, N, N)).float()
    del W, k2
    d *= rms_scale(d) * float(growth)
    d.clamp_(min=-0.99)
    lam = torch.full((N, N, N), 2.0, device=device)
    pois = torch.poisson(lam, generator=g)
    mstar = 1e8 * torch.exp(d) * pois * 0.5
    stars = mstar * float(P.HUBBLE_H / 1e10)
    sfr_timescale = 1e8 * P.SEC_PER_YEAR / P.units()["unit_time_s"]
    sfr = stars / float(sfr_timescale)

    def pad(x):
        out = torch.zeros((N, N, 2 * (N // 2 + 1)), device=device, dtype=torch.float32)
        out[:, :, :N] = x
        return out

    return pad(d), pad(stars), pad(sfr)


def rms_scale(d, rms=0.6):
    return float(rms / float((d.double() ** 2).mean().sqrt()))
