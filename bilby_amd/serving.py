"""Continuous-batching likelihood server on one GPU.

Many samplers submit requests (an event name and N parameter rows); the
scheduler in libgwl (gwl_server_*) keeps at most one batch in flight per
event (per group of events evaluated together): whenever an event has
pending requests and no batch running, everything pending for it (up to
``max_batch_rows`` rows) is packed into pinned staging, evaluated on the
event's stream and copied back, while other events' batches keep running.
The scheduler is native, so it never waits for the interpreter lock;
callers block in ``Future.result()`` with the lock released.  Event data,
PSDs, splines and ROQ weights stay resident; registration is admitted
against an HBM budget.

bilby's own parallelism is a process pool of single-dict callers
(bilby/core/sampler/base_sampler.py:772-800); here the callers share one
device-resident likelihood per event through this queue.
"""
import ctypes
import os
import threading
import time

import numpy as np

from . import _lib
from ._lib import LIB, check


class HBMBudgetExceeded(MemoryError):
    pass


def _device_bytes(lk):
    if getattr(lk, "_relbin", None) is not None:
        out = ctypes.c_int64()
        check(LIB.gwl_relbin_device_bytes(lk._relbin, ctypes.byref(out)), "gwl_relbin_device_bytes")
        ev = ctypes.c_int64()
        check(LIB.gwl_event_device_bytes(lk._event, ctypes.byref(ev)), "gwl_event_device_bytes")
        return int(out.value) + int(ev.value)
    if getattr(lk, "_event", None) is not None:
        out = ctypes.c_int64()
        check(LIB.gwl_event_device_bytes(lk._event, ctypes.byref(out)), "gwl_event_device_bytes")
        if getattr(lk, "_calib", None) is not None:
            cal = ctypes.c_int64()
            check(LIB.gwl_calib_device_bytes(lk._calib, ctypes.byref(cal)),
                  "gwl_calib_device_bytes")
            return int(out.value) + int(cal.value)
        return int(out.value)
    return int(getattr(lk, "hbm_bytes", 0))


class ServedResult:
    """Future of one request: ``result()`` blocks (without the interpreter
    lock) until the scheduler has written the request's values."""

    def __init__(self, server, ticket, event, rows, out):
        self._server = server
        self._ticket = ticket
        self.event = event
        self._rows = rows          # kept alive until the request completes
        self._out = out
        self.t_submit = time.perf_counter()
        self.latency = None
        self._rc = None
        self._msg = ""
        self._lock = threading.Lock()

    def result(self, timeout=None):
        with self._lock:
            if self._rc is None:
                lat = ctypes.c_double()
                rc = LIB.gwl_server_wait(self._server._s, self._ticket, ctypes.byref(lat))
                if rc != 0:
                    self._msg = LIB.gwl_last_error().decode(errors="replace")
                self._rc = rc
                self.latency = lat.value
                self._rows = None
                self._server._outstanding.pop(self._ticket, None)
                self._server.completed.append((self.event, len(self._out), self.t_submit,
                                               self.t_submit + self.latency))
        if self._rc != 0:
            raise _lib.GWLError("request on %r failed (%d): %s" % (self.event, self._rc, self._msg))
        return self._out

    def done(self):
        return self._rc is not None


class _EmptyResult:
    """Future of a request with no rows: already done."""

    def __init__(self, event):
        self.event = event
        self.latency = 0.0

    def result(self, timeout=None):
        return np.zeros(0)

    def done(self):
        return True


class LikelihoodServer:
    """Request queue over device-resident likelihoods.

    ``hbm_budget``: bytes of device memory the resident events and the
    server's buffers may use.  ``max_batch_rows``: rows of one event per
    launch (and the largest request)."""

    def __init__(self, hbm_budget=200e9, max_batch_rows=8192, device=0):
        from .likelihood import get_context
        self.hbm_budget = float(hbm_budget)
        self.max_batch_rows = int(max_batch_rows)
        self.device = device
        h = ctypes.c_void_p()
        check(LIB.gwl_server_create(get_context(device), self.max_batch_rows, ctypes.byref(h)),
              "gwl_server_create")
        self._s = h
        self.events = {}
        self._keep = []
        self._outstanding = {}   # ticket -> future: its buffers live until consumed
        self.event_bytes = 0
        self.completed = []

    @property
    def resident_bytes(self):
        b = ctypes.c_int64()
        check(LIB.gwl_server_device_bytes(self._s, ctypes.byref(b)), "gwl_server_device_bytes")
        return self.event_bytes + int(b.value)

    def _admit(self, name, nbytes):
        if self.resident_bytes + nbytes > self.hbm_budget:
            raise HBMBudgetExceeded("event %r needs %.3g GB; %.3g of %.3g GB in use"
                                    % (name, nbytes / 1e9, self.resident_bytes / 1e9,
                                       self.hbm_budget / 1e9))

    # ------------------------------------------------------------------ events
    def register(self, name, likelihood):
        """Make an event resident (its workspace sized for max_batch_rows) and
        admit its device bytes plus the server's staging for it against the
        budget; over budget raises HBMBudgetExceeded."""
        staging = self.max_batch_rows * (_lib.NPARAM + 1) * 8
        if getattr(likelihood, "_calib", None) is not None:
            check(LIB.gwl_calib_reserve(likelihood._calib, self.max_batch_rows),
                  "gwl_calib_reserve")
            nbytes = _device_bytes(likelihood)
            self._admit(name, nbytes + staging)
            i = ctypes.c_int32()
            check(LIB.gwl_server_add_calib(self._s, likelihood._calib,
                                           int(likelihood.calibration_mode), ctypes.byref(i)),
                  "gwl_server_add_calib")
        elif getattr(likelihood, "_relbin", None) is not None:
            check(LIB.gwl_relbin_reserve(likelihood._relbin, self.max_batch_rows),
                  "gwl_relbin_reserve")
            nbytes = _device_bytes(likelihood)
            self._admit(name, nbytes + staging)
            i = ctypes.c_int32()
            check(LIB.gwl_server_add_relbin(self._s, likelihood._relbin, int(likelihood.precision),
                                            ctypes.byref(i)), "gwl_server_add_relbin")
        elif getattr(likelihood, "_event", None) is not None:
            check(LIB.gwl_event_reserve(likelihood._event, self.max_batch_rows), "gwl_event_reserve")
            nbytes = _device_bytes(likelihood)
            self._admit(name, nbytes + staging)
            i = ctypes.c_int32()
            check(LIB.gwl_server_add_event(self._s, likelihood._event, int(likelihood.precision),
                                           ctypes.byref(i)), "gwl_server_add_event")
        else:
            if getattr(likelihood, "_multi_basis", False):
                raise NotImplementedError(
                    "a ROQ likelihood with several bases is served per basis pair: register "
                    "one likelihood per pair (the rows' basis selection is the caller's)")
            check(LIB.gwl_roq_reserve(likelihood._roq, self.max_batch_rows), "gwl_roq_reserve")
            nbytes = _device_bytes(likelihood)
            self._admit(name, nbytes + staging)
            i = ctypes.c_int32()
            check(LIB.gwl_server_add_roq(self._s, likelihood._roq, int(likelihood.precision),
                                         ctypes.byref(i)), "gwl_server_add_roq")
        self.events[name] = i.value
        self.event_bytes += nbytes
        self._keep.append(likelihood)

    def register_multi(self, names, likelihoods):
        """Register time-marginalised TaylorF2 events that are evaluated
        together: every launch of the group takes all their pending rows
        through one multi-event launch."""
        n = len(likelihoods)
        nbytes = 0
        for name, lk in zip(names, likelihoods):
            check(LIB.gwl_event_reserve(lk._event, 1), "gwl_event_reserve")
            nbytes += _device_bytes(lk)
        self._admit(names[0], nbytes)
        evs = (ctypes.c_void_p * n)(*[lk._event.value for lk in likelihoods])
        first = ctypes.c_int32()
        check(LIB.gwl_server_add_group(self._s, n, evs, ctypes.byref(first)), "gwl_server_add_group")
        for k, name in enumerate(names):
            self.events[name] = first.value + k
        self.event_bytes += nbytes
        self._keep.extend(likelihoods)
        if self.resident_bytes > self.hbm_budget:
            raise HBMBudgetExceeded("group of %d events exceeds the budget (%.3g GB in use)"
                                    % (n, self.resident_bytes / 1e9))

    # ------------------------------------------------------------------ requests
    def submit(self, event, rows):
        """rows: float64 array [n, 16] (gwl column order).  Returns a future
        resolving to an array of n log-likelihood ratios."""
        if event not in self.events:
            raise KeyError(event)
        rows = np.ascontiguousarray(rows, dtype=np.float64)
        if rows.ndim != 2 or rows.shape[1] != _lib.NPARAM:
            raise ValueError("rows must be [n, %d]" % _lib.NPARAM)
        if rows.shape[0] == 0:
            return _EmptyResult(event)
        out = np.empty(rows.shape[0])
        t = ctypes.c_int64()
        check(LIB.gwl_server_submit(self._s, self.events[event], _lib.dptr(rows), rows.shape[0],
                                    _lib.dptr(out), ctypes.byref(t)), "gwl_server_submit")
        fut = ServedResult(self, t.value, event, rows, out)
        self._outstanding[t.value] = fut
        return fut

    def listen(self, path):
        """Also serve other processes on a Unix socket at ``path``
        (gwl_server_listen; bilby_amd.sampling.RemoteLikelihood is the client)."""
        check(LIB.gwl_server_listen(self._s, os.fsencode(path)), "gwl_server_listen")

    def stats(self):
        b, r = ctypes.c_int64(), ctypes.c_int64()
        check(LIB.gwl_server_stats(self._s, ctypes.byref(b), ctypes.byref(r)), "gwl_server_stats")
        return int(b.value), int(r.value)

    @property
    def rounds(self):
        return self.stats()[0]

    @property
    def rows_done(self):
        return self.stats()[1]

    def close(self):
        """Finish every outstanding request, then free the server."""
        if self._s is not None:
            for fut in list(self._outstanding.values()):
                try:
                    fut.result()
                except Exception:
                    pass
            LIB.gwl_server_destroy(self._s)
            self._s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_load(server, event_mix, n_samplers=32, rate=2000.0, duration=5.0, rows_range=(1, 1024),
             seed=0):
    """Seeded Poisson load: ``n_samplers`` threads each emit requests at
    rate / n_samplers per second for ``duration`` seconds.  ``event_mix`` is
    a list of (weight, [(event name, row pool [M, 16]), ...]); each request
    picks a group by weight, an event uniformly within it, and 1..1024 rows
    from that event's pool.  Returns (requests, wall seconds)."""
    weights = np.array([w for w, _ in event_mix], float)
    weights /= weights.sum()
    issued = []
    lock = threading.Lock()
    t0 = time.perf_counter()
    t_end = t0 + duration

    def sampler(i):
        rng = np.random.default_rng((seed, i))
        t = time.perf_counter()
        while True:
            t += rng.exponential(n_samplers / rate)
            if t > t_end:
                return
            delay = t - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            grp = event_mix[rng.choice(len(event_mix), p=weights)][1]
            name, pool = grp[rng.integers(len(grp))]
            n = int(rng.integers(rows_range[0], rows_range[1] + 1))
            idx = rng.integers(0, pool.shape[0], n)
            fut = server.submit(name, pool[idx])
            with lock:
                issued.append((name, idx, fut))

    threads = [threading.Thread(target=sampler, args=(i,)) for i in range(n_samplers)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    for _, _, f in issued:
        f.result()
    return issued, time.perf_counter() - t0


def latency_summary(server, since=0.0, prefix=""):
    sel = [c for c in server.completed if c[2] >= since and c[0].startswith(prefix)]
    lat = np.array([(t1 - t0) for (_, _, t0, t1) in sel])
    rows = np.array([n for (_, n, _, _) in sel])
    if len(lat) == 0:
        return {}
    return dict(requests=int(len(lat)), rows=int(rows.sum()),
                p50_ms=float(np.percentile(lat, 50) * 1e3), p99_ms=float(np.percentile(lat, 99) * 1e3),
                mean_ms=float(lat.mean() * 1e3))
