"""Closed forms of the spectral kurtosis estimator's restatement (oracle/kurtosis.py),
Mode::calculateAndAverageKurtosis (mode.cpp:1405-1450)."""
import numpy as np

import kurtosis as ku


def test_constant_modulus_is_minus_one():
    """|X| constant over FFTs: N s2 / s1^2 = 1, so sk = (1 - 1) N/(N-1) - 1 = -1."""
    rng = np.random.default_rng(1)
    N, C = 64, 32
    amp = rng.uniform(0.5, 2.0, C)
    X = amp * np.exp(2j * np.pi * rng.random((N, C)))
    s1, s2 = ku.sums(X)
    np.testing.assert_allclose(ku.estimator(s1, s2, N, C), -1.0, atol=1e-12)


def test_single_nonzero_fft():
    """Power in one FFT of N: N s2 / s1^2 = N, so sk = (N - 1) N/(N - 1) - 1 = N - 1."""
    N, C = 50, 8
    X = np.zeros((N, C), complex)
    X[7] = 3.0 + 4.0j
    s1, s2 = ku.sums(X)
    np.testing.assert_allclose(ku.estimator(s1, s2, N, C), N - 1.0, rtol=1e-12)


def test_gaussian_mean_is_zero_and_averaging():
    """Complex Gaussian spectra: E|X|^4 = 2 (E|X|^2)^2, so sk averages to 0 within its
    standard error; averaging to C/4 channels sums groups of 4 divided by 4; an all-zero
    band gives no record."""
    rng = np.random.default_rng(2)
    N, C = 400, 256
    X = rng.standard_normal((N, C)) + 1j * rng.standard_normal((N, C))
    s1, s2 = ku.sums(X)
    sk = ku.estimator(s1, s2, N, C)
    assert abs(sk.mean()) < 4 * 2.0 / np.sqrt(N) / np.sqrt(C)
    avg = ku.estimator(s1, s2, N, C // 4)
    np.testing.assert_allclose(avg, sk.reshape(C // 4, 4).mean(axis=1), rtol=1e-12)
    assert ku.estimator(np.zeros(C), np.zeros(C), N, C).size == 0
