"""Pins the oracle FFT: the closed-form cases of the reference's
test/math/test_fft.cpp:21-126 and numpy.fft (pocketfft, the library the
reference calls through src/math/fft.cppm:36-41)."""
import numpy as np
import pytest

import oracle


def test_delta_r2g_and_roundtrip():          # test_fft.cpp:27-41
    n = 8
    x = np.zeros(n, complex); x[0] = 1
    X = oracle.fft1d(x, True)
    assert np.allclose(X, 1.0 / n, atol=1e-12)
    assert np.allclose(oracle.fft1d(X, False), x, atol=1e-14)


def test_cos_peak():                          # test_fft.cpp:43-51
    n = 8
    x = np.cos(2 * np.pi * np.arange(n) / n).astype(complex)
    X = oracle.fft1d(x, True)
    assert abs(abs(X[1]) - 0.5) < 1e-12 and abs(abs(X[n - 1]) - 0.5) < 1e-12


def test_constant_and_n2():                   # test_fft.cpp:53-69
    X = oracle.fft1d(np.ones(8, complex), True)
    assert abs(abs(X[0]) - 1) < 1e-12 and np.abs(X[1:]).max() < 1e-12
    y = oracle.fft1d(np.array([1, 0], complex), False)
    assert np.allclose(y, [1, 1], atol=1e-15)


@pytest.mark.parametrize("n", [6, 45])        # test_fft.cpp:71-95
def test_roundtrip_1d(n):
    x = np.arange(n) * 0.1 + 1j * np.sin(0.1 * np.arange(n))
    assert np.allclose(oracle.fft1d(oracle.fft1d(x, True), False), x, atol=1e-12)


def test_roundtrip_3d():                      # test_fft.cpp:98-126
    n1, n2, n3 = 6, 8, 5
    i = np.arange(n1 * n2 * n3)
    g = ((i % 7) + 1j * ((i * i) % 11)).reshape(n1, n2, n3)
    back = oracle.fft3d(oracle.fft3d(g, False), True)
    assert np.allclose(back, g, atol=1e-12)


@pytest.mark.parametrize("n", [2, 3, 5, 7, 11, 13, 36, 45, 48, 90, 77, 128])
def test_matches_pocketfft(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    # R2G: forward exp(-i) with 1/n; G2R: backward exp(+i) unscaled
    assert np.allclose(oracle.fft1d(x, True), np.fft.fft(x) / n, rtol=0, atol=1e-14 * np.abs(x).sum())
    assert np.allclose(oracle.fft1d(x, False), np.fft.ifft(x) * n, rtol=0, atol=1e-13 * np.abs(x).sum())


def test_3d_matches_pocketfft():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((12, 10, 9)) + 1j * rng.standard_normal((12, 10, 9))
    assert np.allclose(oracle.fft3d(x, True), np.fft.fftn(x) / x.size, atol=1e-15 * np.abs(x).sum())
    assert np.allclose(oracle.fft3d(x, False), np.fft.ifftn(x) * x.size, atol=1e-12 * np.abs(x).sum())
