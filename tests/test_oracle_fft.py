"""Pins the oracle's FFT: brute-force DFT, numpy.fft, and the identities a DFT obeys."""
import numpy as np
import pytest

from meraxes_amd import synth
from oracle import oracle as O


def dft(x, sign):
    n = len(x)
    j = np.arange(n)
    W = np.exp(sign * 2j * np.pi * np.outer(j, j) / n)
    return W @ x


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 12, 15, 16, 24, 25, 30])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft1d_brute_force_fp64(n, sign):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    got = O.fft1d(x.astype(np.complex128), sign)
    ref = dft(x, sign)
    assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("n", [17, 31, 48, 64, 96, 100, 192, 256, 384, 512, 1024])
def test_fft1d_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert np.allclose(O.fft1d(x.astype(np.complex128), -1), np.fft.fft(x), rtol=0, atol=1e-12 * np.sqrt(n) * 10)
    assert np.allclose(O.fft1d(x.astype(np.complex128), +1), np.fft.ifft(x) * n, rtol=0, atol=1e-11 * np.sqrt(n))
    # float32 transform: error of a well-conditioned float FFT, ~1e-7 relative
    g32 = O.fft1d(x.astype(np.complex64), -1)
    ref = np.fft.fft(x.astype(np.complex64).astype(np.complex128))
    assert np.linalg.norm(g32 - ref) / np.linalg.norm(ref) < 5e-7


@pytest.mark.parametrize("N", [4, 6, 8, 12, 16, 32])
def test_r2c_c2r_3d_numpy(N):
    rng = np.random.default_rng(N)
    f = rng.standard_normal((N, N, N))
    pad = np.zeros((N, N, 2 * (N // 2 + 1)))
    pad[:, :, :N] = f
    H = O.r2c3d(pad)
    assert H.shape == (N, N, N // 2 + 1)
    assert np.allclose(H, np.fft.rfftn(f), atol=1e-10)
    back = O.c2r3d(H)
    assert np.allclose(back[:, :, :N], f * N**3, atol=1e-9)
    assert np.all(back[:, :, N:] == 0)
    # float32 path
    H32 = O.r2c3d(pad.astype(np.float32))
    assert np.linalg.norm(H32 - np.fft.rfftn(f)) / np.linalg.norm(np.fft.rfftn(f)) < 1e-6


def test_identities():
    N = 16
    rng = np.random.default_rng(0)
    f = rng.standard_normal((N, N, N)).astype(np.float32).astype(np.float64)
    pad = synth.to_padded(f).astype(np.float64)
    H = O.r2c3d(pad)
    # Parseval over the full spectrum (half spectrum, doubling the interior kz)
    w = np.full(N // 2 + 1, 2.0)
    w[0] = w[-1] = 1.0
    assert np.isclose((np.abs(H) ** 2 * w).sum() / N**3, (f**2).sum(), rtol=1e-12)
    # Hermitian symmetry of the kz = 0 and Nyquist planes
    for kz in (0, N // 2):
        P = H[:, :, kz]
        Pm = np.roll(np.flip(P, (0, 1)), 1, (0, 1))
        assert np.allclose(P, np.conj(Pm), atol=1e-10)
    # delta -> flat spectrum; plane wave -> one bin
    d = np.zeros((N, N, N))
    d[0, 0, 0] = 1.0
    assert np.allclose(O.r2c3d(synth.to_padded(d).astype(np.float64)), 1.0)
    x = np.arange(N)
    pw = np.cos(2 * np.pi * (3 * x[:, None, None] + 5 * x[None, :, None] + 2 * x[None, None, :]) / N)
    Hp = O.r2c3d(synth.to_padded(pw).astype(np.float64))
    big = np.argwhere(np.abs(Hp) > 1e-3)  # float32 input rounding leaves ~1e-5 elsewhere
    assert [tuple(b) for b in big] == [(3, 5, 2)]  # its conjugate has kz < 0, outside the half spectrum
    assert np.isclose(Hp[3, 5, 2], N**3 / 2)


def test_c2r_ignores_imaginary_dc_and_nyquist_like_fftw():
    N = 8
    rng = np.random.default_rng(3)
    H = O.r2c3d(synth.to_padded(rng.standard_normal((N, N, N))).astype(np.float64))
    H2 = H.copy()
    H2[0, 0, 0] += 5j
    H2[0, 0, N // 2] += 3j
    assert np.allclose(O.c2r3d(H), O.c2r3d(H2), atol=1e-10)
