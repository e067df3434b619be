"""CPU restatement of the spectral kurtosis estimator (test infrastructure only).

This module is a CHECKER, not the product: only tests/ may import it.  Mode::process sums
s1 = sum |X|^2 and s2 = sum |X|^4 over the FFTs of a block range per recorded band and channel
(mpifxcorr/src/mode.cpp:1237-1253); Mode::calculateAndAverageKurtosis (mode.cpp:1405-1450)
forms sk = (numblocks s2 / s1^2 - 1) numblocks / (numblocks - 1) - 1 and, when fewer channels
are wanted, sums groups of C / maxchannels channels divided by the group size.  A band whose
s1 is zero everywhere gives no record.

Pins (tests/test_kurtosis_cpu.py): constant-modulus spectra give sk = -1 in every channel;
a single nonzero FFT gives sk = numblocks - 1; Gaussian spectra average to 0.
"""
from __future__ import annotations

import numpy as np


def sums(spectra: np.ndarray):
    """spectra [numblocks, C] -> (s1, s2) in FP64."""
    p = np.abs(spectra) ** 2
    return p.sum(axis=0), (p * p).sum(axis=0)


def estimator(s1: np.ndarray, s2: np.ndarray, numblocks: int, maxchannels: int) -> np.ndarray:
    if not np.any(s1 > 0):
        return np.zeros(0)
    with np.errstate(divide="ignore", invalid="ignore"):
        sk = (numblocks * s2 / (s1 * s1) - 1.0) * numblocks / (numblocks - 1.0) - 1.0
    C = len(sk)
    if maxchannels < C:
        k = C // maxchannels
        sk = (sk[: maxchannels * k] / k).reshape(maxchannels, k).sum(axis=1)
    return sk
