"""Ragged (CSR) batches of signals with per-signal taps, and HIP-graph
capture of repeated calls.

``offs`` (int64, S+1) delimits signal s as ``x[offs[s]:offs[s+1]]``;
``b_offs`` (int64, S+1) delimits its lfilter-order taps in ``b``.  A
:class:`RaggedPlan` is built once from the host offsets (cost-sorted work
items + a device work queue); :meth:`RaggedPlan.execute` is one persistent
launch and allocates nothing, so it can be captured with :class:`Graph`.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib


def _i64p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))


def _stream(device=None):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


class RaggedPlan:
    def __init__(self, offs, b_offs, precision="f32"):
        self.offs = np.ascontiguousarray(offs, dtype=np.int64)
        self.b_offs = np.ascontiguousarray(b_offs, dtype=np.int64)
        if self.offs.ndim != 1 or self.offs.shape != self.b_offs.shape or self.offs.size < 1:
            raise ValueError("offs and b_offs must be 1-D int64 arrays of equal length S+1")
        self.S = self.offs.size - 1
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().scir_hip_ragged_plan_create(
            _i64p(self.offs), _i64p(self.b_offs), self.S, _lib.precision_code(precision),
            ctypes.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def _check(self, x, b, y):
        for name, t, n in (("x", x, self.offs[-1]), ("y", y, self.offs[-1]),
                           ("b", b, self.b_offs[-1])):
            if t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous():
                raise TypeError(f"{name} must be a contiguous float32 tensor on a HIP device")
            if t.numel() < n:
                raise ValueError(f"{name} holds {t.numel()} values, offsets need {n}")

    def execute(self, x, b, y=None):
        """One persistent launch over every signal (on torch's current stream)."""
        if y is None:
            y = torch.empty_like(x)
        self._check(x, b, y)
        _lib.check(_lib.lib().scir_hip_ragged_plan_execute(
            self._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(b.data_ptr()),
            ctypes.c_void_p(y.data_ptr()), _stream(x.device)))
        return y

    def execute_per_signal(self, x, b, y=None):
        """The baseline: one dense-kernel launch per signal."""
        if y is None:
            y = torch.empty_like(x)
        self._check(x, b, y)
        _lib.check(_lib.lib().scir_hip_ragged_plan_execute_per_signal(
            self._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(b.data_ptr()),
            ctypes.c_void_p(y.data_ptr()), _stream(x.device)))
        return y

    def close(self):
        if self._h:
            _lib.check(_lib.lib().scir_hip_ragged_plan_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """Capture the library calls issued inside ``with Graph(stream) as g:``
    on ``stream`` and replay them with :meth:`launch`."""

    def __init__(self, stream: torch.cuda.Stream):
        self.stream = stream
        self._h = None

    def __enter__(self):
        self._ctx = torch.cuda.stream(self.stream)
        self._ctx.__enter__()
        _lib.check(_lib.lib().scir_hip_graph_begin(ctypes.c_void_p(self.stream.cuda_stream)))
        return self

    def __exit__(self, exc_type, exc, tb):
        h = ctypes.c_void_p()
        status = _lib.lib().scir_hip_graph_end(ctypes.c_void_p(self.stream.cuda_stream),
                                               ctypes.byref(h))
        self._ctx.__exit__(exc_type, exc, tb)
        if exc_type is None:
            _lib.check(status)
            self._h = h
        elif status == _lib.OK and h:
            # The capture body raised: drop whatever was captured.
            _lib.lib().scir_hip_graph_destroy(h)
        return False

    def launch(self):
        _lib.check(_lib.lib().scir_hip_graph_launch(self._h,
                                                    ctypes.c_void_p(self.stream.cuda_stream)))

    def close(self):
        if self._h:
            _lib.check(_lib.lib().scir_hip_graph_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
