"""Pins the oracle's filter() restatement (reionization.c:2110-2170) against closed forms."""
import numpy as np
import pytest

from meraxes_amd import synth
from oracle import oracle as O

EPS32 = 2.0**-23


def kmag(nx, ny, nz, N, L):
    dk = 2 * np.pi / L
    c = lambda n: (n - N if n > N // 2 else n) * dk  # noqa: E731
    return np.sqrt(c(nx) ** 2 + c(ny) ** 2 + (nz * dk) ** 2)


def test_tophat_closed_form():
    N, L = 64, 100.0
    rng = np.random.default_rng(0)
    for R in (0.9, 3.0, 12.0, 50.0):
        for _ in range(300):
            nx, ny, nz = rng.integers(0, N), rng.integers(0, N), rng.integers(0, N // 2 + 1)
            x = kmag(nx, ny, nz, N, L) * R
            w = O.window_f32(int(nx), int(ny), int(nz), N, R, 0, L)
            if x <= 1e-4:
                assert w == 1.0
                continue
            exact = 3 * (np.sin(x) - x * np.cos(x)) / x**3
            # the reference's float expression cancels: error ~ a few ulp of 1/x^2
            tol = 12 * EPS32 * (1 + 1 / x**2) + 4 * EPS32 * abs(exact) + 1e-6 * x * R / L
            assert abs(w - exact) <= tol, (nx, ny, nz, R, w, exact)


def test_tophat_small_kR_untouched_and_dc():
    for t in (0, 1, 2):
        assert O.window_f32(0, 0, 0, 32, 5.0, t, 50.0) == 1.0
    # tiny R: every kR below 1e-4 leaves the top-hat mode untouched
    assert O.window_f32(1, 0, 0, 32, 1e-5, 0, 50.0) == 1.0


def test_sharpk_edge():
    N, L = 32, 50.0
    k = kmag(3, 0, 0, N, L)
    R_edge = 1.0 / (0.413566994 * k)
    assert O.window_f32(3, 0, 0, N, R_edge * 0.999, 1, L) == 1.0
    assert O.window_f32(3, 0, 0, N, R_edge * 1.001, 1, L) == 0.0
    # lower mode kept at the same radius
    assert O.window_f32(2, 0, 0, N, R_edge * 1.001, 1, L) == 1.0


def test_gaussian_closed_form():
    N, L = 32, 50.0
    for R in (0.5, 2.0, 8.0):
        for n in range(0, N // 2 + 1, 3):
            x = kmag(n, 1, 2, N, L) * R
            w = O.window_f32(n, 1, 2, N, R, 2, L)
            assert abs(w - np.exp(-((0.643 * x) ** 2) / 2)) <= 1e-6


def test_constant_field_filters_to_itself():
    N, L = 16, 40.0
    f = np.full((N, N, N), 0.37, np.float32)
    for t in (0, 1, 2):
        out = O.filtered_field(N, L, t, synth.to_padded(f), 7.0, 1, -1)
        assert np.allclose(out[:, :, :N], 0.37, atol=1e-6)


def test_slab_offsets_equal_whole():
    N, L = 16, 40.0
    H = O.r2c3d(synth.to_padded(synth.gaussian_field(N, 1)))
    whole = O.filter_f32(H, 3.0, 0, L)
    a = O.filter_f32(H[:5], 3.0, 0, L, local_ix_start=0)
    b = O.filter_f32(H[5:], 3.0, 0, L, local_ix_start=5)
    assert np.array_equal(np.concatenate([a, b]), whole)


@pytest.mark.parametrize("t", [0, 1, 2])
def test_filter_multiplies_by_window(t):
    N, L, R = 12, 30.0, 2.5
    H = O.r2c3d(synth.to_padded(synth.gaussian_field(N, 2)))
    F = O.filter_f32(H, R, t, L)
    for nx in range(N):
        for ny in range(N):
            for nz in range(N // 2 + 1):
                w = O.window_f32(nx, ny, nz, N, R, t, L)
                assert F[nx, ny, nz] == np.complex64(H[nx, ny, nz] * np.float32(w))


def test_invalid_filter_type_is_an_error():
    H = np.zeros((8, 8, 5), np.complex64)
    with pytest.raises(RuntimeError):
        O.filter_f32(H, 1.0, 3, 10.0)


def test_fp64_window_matches_closed_form():
    N, L, R = 16, 40.0, 4.0
    H = np.ones((N, N, N // 2 + 1), np.complex128)
    F = O.filter_f64(H, R, 0, L)
    x = kmag(3, 4, 5, N, L) * R
    assert np.isclose(F[3, 4, 5].real, 3 * (np.sin(x) - x * np.cos(x)) / x**3, rtol=1e-12)
