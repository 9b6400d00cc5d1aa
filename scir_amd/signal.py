"""Mirror of the ``scir_signal::gpu`` forwarder
(crates/scir-signal/src/lib.rs:364-375): batched FIR with device selection."""
from __future__ import annotations

from .gpu import Device, fir1d_batched_f32_auto


def fir1d_batched_f32(x, taps, device: Device):
    """Batched FIR for f32 with device selection (forwards to the dispatcher)."""
    return fir1d_batched_f32_auto(x, taps, device)
