"""Uneven batches over many events in one persistent launch.

Each event is a GravitationalWaveTransient (time-marginalised, FP32 bulk);
their parameter batches of different sizes are evaluated by one persistent
kernel that drains a longest-first work list, then one epilogue launch.  The
pair can be captured into a hipGraph and replayed for every new set of
parameter rows.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import LIB, check


class MultiEventLikelihood:
    def __init__(self, likelihoods, max_rows, capture=False, device=0):
        self.likelihoods = list(likelihoods)
        self.max_rows = [int(m) for m in max_rows]
        self.device = device
        dev = torch.device("cuda:%d" % device)
        self.params = [torch.zeros((max(1, m), _lib.NPARAM), dtype=torch.float64, device=dev)
                       for m in self.max_rows]
        self.outs = [torch.zeros(max(1, m), dtype=torch.float64, device=dev) for m in self.max_rows]
        # pinned host staging of each event's rows (asynchronous uploads)
        self._staging = [torch.zeros((max(1, m), _lib.NPARAM), dtype=torch.float64).pin_memory()
                         for m in self.max_rows]
        n = len(self.likelihoods)
        evs = (ctypes.c_void_p * n)(*[lk._event.value for lk in self.likelihoods])
        pp = (ctypes.c_void_p * n)(*[t.data_ptr() for t in self.params])
        op = (ctypes.c_void_p * n)(*[t.data_ptr() for t in self.outs])
        mr = (ctypes.c_int64 * n)(*self.max_rows)
        from .likelihood import get_context
        h = ctypes.c_void_p()
        check(LIB.gwl_multi_create(get_context(device), n, evs, pp, op, mr, ctypes.byref(h)),
              "gwl_multi_create")
        self._m = h
        self.stream = torch.cuda.Stream(device=dev)
        self.rows = [0] * n
        self.captured = False
        if capture:
            self.capture()

    def __del__(self):
        try:
            if getattr(self, "_m", None) is not None:
                LIB.gwl_multi_destroy(self._m)
                self._m = None
        except Exception:
            pass

    def set_tile(self, steps_per_item):
        arr = (ctypes.c_int32 * len(self.likelihoods))(*[int(s) for s in steps_per_item])
        check(LIB.gwl_multi_set_tile(self._m, arr), "gwl_multi_set_tile")

    def set_parameters(self, e, parameters):
        """Rows for event ``e``: a dict of arrays, an array [n, 16] or a GPU
        tensor [n, 16]."""
        lk = self.likelihoods[e]
        if isinstance(parameters, torch.Tensor) and parameters.is_cuda:
            rows = parameters.to(torch.float64)
        elif isinstance(parameters, dict):
            rows = lk.rows_from_parameters(parameters)
        else:
            rows = np.ascontiguousarray(parameters, dtype=np.float64)
        n = rows.shape[0]
        if n > self.max_rows[e]:
            raise ValueError("event %d: %d rows > max_rows %d" % (e, n, self.max_rows[e]))
        # the previous launch may still read the staging buffer and rows
        self.stream.synchronize()
        with torch.cuda.stream(self.stream):
            if isinstance(rows, torch.Tensor):
                self.params[e][:n].copy_(rows, non_blocking=True)
            else:
                stage = self._staging[e]
                stage[:n].numpy()[...] = rows
                self.params[e][:n].copy_(stage[:n], non_blocking=True)
        self.rows[e] = n

    def capture(self):
        self._set_rows()
        check(LIB.gwl_multi_capture(self._m, ctypes.c_void_p(self.stream.cuda_stream)),
              "gwl_multi_capture")
        self.captured = True

    def _set_rows(self):
        arr = (ctypes.c_int64 * len(self.rows))(*self.rows)
        check(LIB.gwl_multi_set_rows(self._m, arr, ctypes.c_void_p(self.stream.cuda_stream)),
              "gwl_multi_set_rows")

    def run(self, synchronize=True):
        """Evaluate every event's current rows; returns GPU tensors."""
        self.stream.synchronize()
        self._set_rows()
        check(LIB.gwl_multi_launch(self._m, ctypes.c_void_p(self.stream.cuda_stream)),
              "gwl_multi_launch")
        if synchronize:
            self.stream.synchronize()
        return [o[:n] for o, n in zip(self.outs, self.rows)]
