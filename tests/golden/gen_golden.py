"""Write the golden vectors under tests/golden/ (run in the build container).

Inputs and expected outputs come from scipy 1.15.3 / numpy 2.2.x, a
third-party library, not from the reference: the reference ships no FIR or
lfilter fixture (its fixture script, scripts/gen_signal_fixtures.py, covers
IIR and resample_poly only).  The reference pins scipy 1.13.1
(requirements.txt:31); lfilter's FIR arithmetic and firwin's design rules are
the same in both versions.

Files (all float64 unless noted, saved with allow_pickle=False):
  firwin_<name>.npy        taps from scipy.signal.firwin
  lfilter_x.npy            float32 inputs [B, N]
  lfilter_b<K>.npy         float32 lfilter-order taps used with lfilter_x
  lfilter_y<K>.npy         scipy.signal.lfilter(b, [1.0], x) in float64
  resample_x.npy           np.linspace(0, 1, 32, endpoint=False), the input of the
                           reference's fixture script (scripts/gen_signal_fixtures.py:28)
  resample_poly_2_3.npy    scipy.signal.resample_poly(resample_x, 2, 3), the expected
                           output of the reference's resample_poly test
                           (crates/scir-signal/src/lib.rs:655-668, atol 2e-2)
"""
import pathlib

import numpy as np
from scipy import signal

OUT = pathlib.Path(__file__).resolve().parent

FIRWIN = {
    "lp33": dict(numtaps=33, cutoff=0.2),
    "lp129": dict(numtaps=129, cutoff=0.2),
    "lp255": dict(numtaps=255, cutoff=0.2),
    "bp512_kaiser8": dict(numtaps=512, cutoff=[0.1, 0.3], window=("kaiser", 8.0), pass_zero=False),
    "bp2048_kaiser8": dict(numtaps=2048, cutoff=[0.1, 0.3], window=("kaiser", 8.0), pass_zero=False),
    "lp15_c005": dict(numtaps=15, cutoff=0.05),
    "lp1023_c045": dict(numtaps=1023, cutoff=0.45),
    "lp64_c0317": dict(numtaps=64, cutoff=0.317),
    "lp31_c0333": dict(numtaps=31, cutoff=1.0 / 3.0),
}


def main() -> None:
    for name, kw in FIRWIN.items():
        np.save(OUT / f"firwin_{name}.npy", signal.firwin(**kw), allow_pickle=False)

    rng = np.random.default_rng(20251202)
    x = rng.standard_normal((3, 300)).astype(np.float32)
    x[1, ::37] *= 1000.0  # a few large samples, as in the long-tap configuration
    np.save(OUT / "lfilter_x.npy", x, allow_pickle=False)
    for K in (1, 2, 3, 17, 33, 129):
        b = rng.standard_normal(K).astype(np.float32)  # asymmetric on purpose
        y = signal.lfilter(b.astype(np.float64), [1.0], x.astype(np.float64), axis=-1)
        np.save(OUT / f"lfilter_b{K}.npy", b, allow_pickle=False)
        np.save(OUT / f"lfilter_y{K}.npy", y, allow_pickle=False)

    xr = np.linspace(0, 1, 32, endpoint=False)
    np.save(OUT / "resample_x.npy", xr, allow_pickle=False)
    np.save(OUT / "resample_poly_2_3.npy", signal.resample_poly(xr, 2, 3), allow_pickle=False)


if __name__ == "__main__":
    main()
