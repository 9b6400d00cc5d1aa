"""Windowed-sinc FIR tap design (host side, float64).

The reference has no FIR designer (its windows are Hann, Blackman-Harris and
flat-top only, crates/scir-signal/src/window.rs:23-34), so the taps used by
the benchmark configurations are designed here: a windowed sinc with the
``scipy.signal.firwin`` conventions -- cutoffs as fractions of Nyquist,
symmetric windows, and ``scale=True`` normalisation (unit gain at DC for a
low-pass, at the band centre for a band-pass).  tests/test_design.py pins the
output against scipy-generated golden taps in tests/golden/.

Taps come out in lfilter order ``b``; they are symmetric, so they equal the
reference's ``taps`` order as well.
"""
from __future__ import annotations

import math

import numpy as np


def bessel_i0(z) -> np.ndarray:
    """Modified Bessel function of the first kind, order 0, by its power
    series sum_k ((z/2)^k / k!)^2, summed until the term is below 1e-17 of
    the partial sum (converges for every real z; the windows use z <= 20)."""
    z = np.asarray(z, dtype=np.float64)
    q = (z * 0.5) ** 2
    term = np.ones_like(z)
    total = np.ones_like(z)
    k = 0
    while True:
        k += 1
        term = term * q / (k * k)
        total = total + term
        if np.all(term <= 1e-17 * total):
            break
    return total


def window(name, numtaps: int) -> np.ndarray:
    """Symmetric window of length ``numtaps``.

    ``"hamming"``: 0.54 - 0.46 cos(2 pi n / (M-1));
    ``("kaiser", beta)``: I0(beta sqrt(1 - ((n - a)/a)^2)) / I0(beta), a = (M-1)/2.
    """
    M = int(numtaps)
    if M < 1:
        raise ValueError("numtaps must be >= 1")
    if M == 1:
        return np.ones(1)
    n = np.arange(M, dtype=np.float64)
    if name == "hamming":
        return 0.54 - 0.46 * np.cos(2.0 * math.pi * n / (M - 1))
    if isinstance(name, tuple) and name[0] == "kaiser":
        beta = float(name[1])
        a = (M - 1) / 2.0
        r = (n - a) / a
        return bessel_i0(beta * np.sqrt(np.clip(1.0 - r * r, 0.0, None))) / bessel_i0(np.float64(beta))
    raise ValueError(f"unsupported window {name!r}")


def firwin(numtaps: int, cutoff, window_spec="hamming", pass_zero: bool = True) -> np.ndarray:
    """Windowed-sinc FIR taps (float64), firwin-style.

    ``cutoff`` is a scalar (low-pass when ``pass_zero``) or a pair
    ``[lo, hi]`` (band-pass when ``pass_zero`` is False), as fractions of
    Nyquist in (0, 1)."""
    cut = np.atleast_1d(np.asarray(cutoff, dtype=np.float64))
    if np.any(cut <= 0) or np.any(cut >= 1):
        raise ValueError("cutoff must lie strictly between 0 and 1 (fraction of Nyquist)")
    if np.any(np.diff(cut) <= 0):
        raise ValueError("cutoffs must be strictly increasing")
    edges = np.concatenate(([0.0] if pass_zero else [], cut))
    if edges.size % 2 == 1:
        # The last band runs to Nyquist (a high-pass response).
        if numtaps % 2 == 0:
            raise ValueError("a filter with gain at Nyquist needs an odd number of taps")
        edges = np.concatenate((edges, [1.0]))
    bands = edges.reshape(-1, 2)
    alpha = 0.5 * (numtaps - 1)
    m = np.arange(numtaps, dtype=np.float64) - alpha
    h = np.zeros(numtaps, dtype=np.float64)
    for left, right in bands:
        h += right * np.sinc(right * m)
        h -= left * np.sinc(left * m)
    h *= window(window_spec, numtaps)
    left, right = bands[0]
    if left == 0.0:
        f0 = 0.0
    elif right == 1.0:
        f0 = 1.0
    else:
        f0 = 0.5 * (left + right)
    h /= np.sum(h * np.cos(math.pi * m * f0))
    return h


def lowpass(numtaps: int, cutoff: float = 0.2) -> np.ndarray:
    """Hamming-windowed low-pass, unit DC gain, as float32."""
    return firwin(numtaps, cutoff).astype(np.float32)


def bandpass_kaiser(numtaps: int, band=(0.1, 0.3), beta: float = 8.0) -> np.ndarray:
    """Kaiser(beta)-windowed band-pass, unit gain at the band centre, float32."""
    return firwin(numtaps, list(band), ("kaiser", beta), pass_zero=False).astype(np.float32)


def resample_filter(up: int, down: int, numtaps: int = 31) -> np.ndarray:
    """Anti-aliasing filter of the polyphase resampler (float64): a Hamming
    low-pass at 1/max(up, down) of Nyquist with gain ``up``.  For up=2,
    down=3 and 31 taps this is the fixed table of the reference's
    resample_poly (crates/scir-signal/src/lib.rs:315-346) to within one ulp."""
    q = max(int(up), int(down))
    if q < 2:
        raise ValueError("resampling needs up or down >= 2")
    return float(up) * firwin(numtaps, 1.0 / q)
