"""Tap design (scir_amd/design.py) against scipy.signal.firwin golden taps."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import load_golden
from scir_amd import design

CASES = {
    "lp33": (33, 0.2, "hamming", True),
    "lp129": (129, 0.2, "hamming", True),
    "lp255": (255, 0.2, "hamming", True),
    "bp512_kaiser8": (512, [0.1, 0.3], ("kaiser", 8.0), False),
    "bp2048_kaiser8": (2048, [0.1, 0.3], ("kaiser", 8.0), False),
    "lp15_c005": (15, 0.05, "hamming", True),
    "lp1023_c045": (1023, 0.45, "hamming", True),
    "lp64_c0317": (64, 0.317, "hamming", True),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_firwin_matches_golden(name):
    n, cut, win, pz = CASES[name]
    h = design.firwin(n, cut, win, pass_zero=pz)
    ref = load_golden(f"firwin_{name}.npy")
    np.testing.assert_allclose(h, ref, rtol=1e-10, atol=1e-13)


def test_bessel_i0_known_values():
    # I0(0) = 1, I0(1) = 1.2660658777520082, I0(8) = 427.56411572180474
    np.testing.assert_allclose(design.bessel_i0(np.array([0.0, 1.0, 8.0])),
                               [1.0, 1.2660658777520082, 427.56411572180474], rtol=1e-14)


def test_normalisation():
    assert abs(design.lowpass(33).astype(np.float64).sum() - 1.0) < 1e-6
    h = design.firwin(512, [0.1, 0.3], ("kaiser", 8.0), pass_zero=False)
    m = np.arange(512) - 255.5
    assert abs(np.sum(h * np.cos(np.pi * m * 0.2)) - 1.0) < 1e-12


def test_rejects_bad_cutoffs():
    with pytest.raises(ValueError):
        design.firwin(33, 1.2)
    with pytest.raises(ValueError):
        design.firwin(32, 0.3, pass_zero=False)  # high-pass with even length
