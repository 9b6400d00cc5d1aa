"""Synthetic inputs of the benchmark workloads (seeded, reproducible).

Dense rows come from the library's counter-based N(0,1) generator on the GPU;
the ragged workload's lengths, tap counts and cutoffs come from numpy's
PCG64 with a fixed seed, so host and device agree on the layout.
"""
from __future__ import annotations

import numpy as np

from . import design

SEED_BASE = 0x5C120000


def ragged_layout(S=10000, min_len=256, max_len=1 << 20, k_lo=15, k_hi=1023, seed=SEED_BASE + 3):
    """Lengths log-uniform in [min_len, max_len], K uniform in [k_lo, k_hi],
    cutoffs uniform in [0.05, 0.45] (fraction of Nyquist).
    Returns (offs, b_offs, taps float32 concatenated in lfilter order)."""
    rng = np.random.default_rng(seed)
    lens = np.round(np.exp(rng.uniform(np.log(min_len), np.log(max_len), S))).astype(np.int64)
    ks = rng.integers(k_lo, k_hi + 1, S).astype(np.int64)
    cut = rng.uniform(0.05, 0.45, S)
    offs = np.concatenate(([0], np.cumsum(lens))).astype(np.int64)
    b_offs = np.concatenate(([0], np.cumsum(ks))).astype(np.int64)
    taps = np.concatenate([design.firwin(int(k), float(c)).astype(np.float32)
                           for k, c in zip(ks, cut)])
    return offs, b_offs, taps
