"""Continuous-batching likelihood server on one GPU.

Many samplers submit requests (an event name and N parameter rows); one
scheduler thread forms a batch per event from everything pending, launches
the events' batches on separate HIP streams so they overlap on the device,
and completes each request's future as soon as its batch finishes.  Requests
that arrive while batches are in flight join the next round.  Event data,
PSDs, splines and ROQ weights stay resident; registration and in-flight
work are admitted against an HBM budget.

bilby's own parallelism is a process pool of single-dict callers
(bilby/core/sampler/base_sampler.py:772-800); here the callers share one
device-resident likelihood per event through this queue.
"""
import ctypes
import threading
import time
from collections import deque
from concurrent.futures import Future

import numpy as np
import torch

from . import _lib
from ._lib import LIB, check


class HBMBudgetExceeded(MemoryError):
    pass


class _Request:
    __slots__ = ("event", "rows", "future", "t_submit", "t_done")

    def __init__(self, event, rows):
        self.event = event
        self.rows = rows
        self.future = Future()
        self.t_submit = time.perf_counter()
        self.t_done = None


def _device_bytes(lk):
    if hasattr(lk, "_event") and lk._event is not None:
        out = ctypes.c_int64()
        check(LIB.gwl_event_device_bytes(lk._event, ctypes.byref(out)), "gwl_event_device_bytes")
        return int(out.value)
    return int(getattr(lk, "hbm_bytes", 0))


class LikelihoodServer:
    """Request loop over device-resident likelihoods.

    ``hbm_budget``: bytes of device memory the resident events plus the
    in-flight batches may use.  ``max_batch_rows``: rows of one event per
    launch; larger request sets wait for the next round."""

    def __init__(self, hbm_budget=200e9, max_batch_rows=8192, n_streams=4, device=0):
        self.hbm_budget = float(hbm_budget)
        self.max_batch_rows = int(max_batch_rows)
        self.device = device
        dev = torch.device("cuda:%d" % device)
        self._dev = dev
        self.streams = [torch.cuda.Stream(device=dev) for _ in range(n_streams)]
        self.events = {}
        self.resident_bytes = 0
        self._pending = deque()
        self._cv = threading.Condition()
        self._stop = False
        self.completed = []
        self.rounds = 0
        self.rows_done = 0
        self._thread = threading.Thread(target=self._loop, name="gwl-server", daemon=True)
        self._thread.start()

    # ------------------------------------------------------------------ events
    def register(self, name, likelihood, row_bytes=None):
        """Make an event resident.  Its device bytes (data, spline or ROQ
        weights, and a workspace sized for ``max_batch_rows``) are admitted
        against the budget; over budget raises HBMBudgetExceeded."""
        if hasattr(likelihood, "_event"):
            check(LIB.gwl_event_reserve(likelihood._event, self.max_batch_rows), "gwl_event_reserve")
        nbytes = _device_bytes(likelihood)
        extra = 0
        if row_bytes is not None:
            extra = int(row_bytes) * self.max_batch_rows
        if self.resident_bytes + nbytes + extra > self.hbm_budget:
            raise HBMBudgetExceeded("event %r needs %.3g GB; %.3g of %.3g GB in use"
                                    % (name, (nbytes + extra) / 1e9, self.resident_bytes / 1e9,
                                       self.hbm_budget / 1e9))
        self.events[name] = dict(likelihood=likelihood, bytes=nbytes + extra)
        self.resident_bytes += nbytes + extra

    def register_multi(self, names, likelihoods):
        """Register time-marginalised events that are evaluated together:
        every round, all their pending rows go through one persistent
        multi-event launch (replayed from a captured hipGraph)."""
        from .multi import MultiEventLikelihood
        for name, lk in zip(names, likelihoods):
            self.register(name, lk)
        multi = MultiEventLikelihood(likelihoods, [self.max_batch_rows] * len(likelihoods),
                                     device=self.device)
        multi.capture()
        gid = len(getattr(self, "_multis", []))
        if not hasattr(self, "_multis"):
            self._multis = []
        self._multis.append(multi)
        for k, name in enumerate(names):
            self.events[name]["multi"] = (gid, k)
        extra = sum(int(t.numel() * 8) for t in multi.params + multi.outs)
        self.resident_bytes += extra
        return multi

    def unregister(self, name):
        ev = self.events.pop(name)
        self.resident_bytes -= ev["bytes"]

    # ------------------------------------------------------------------ requests
    def submit(self, event, rows):
        """rows: float64 array [n, 16] (gwl column order).  Returns a Future
        resolving to an array of n log-likelihood ratios."""
        if event not in self.events:
            raise KeyError(event)
        rows = np.ascontiguousarray(rows, dtype=np.float64)
        if rows.ndim != 2 or rows.shape[1] != _lib.NPARAM:
            raise ValueError("rows must be [n, %d]" % _lib.NPARAM)
        req = _Request(event, rows)
        with self._cv:
            self._pending.append(req)
            self._cv.notify()
        return req.future

    def close(self):
        with self._cv:
            self._stop = True
            self._cv.notify()
        self._thread.join()

    # ------------------------------------------------------------------ loop
    def _take_round(self):
        """All pending requests, grouped by event, each group capped at
        max_batch_rows (a request is never split); the rest stays queued."""
        groups = {}
        keep = deque()
        while self._pending:
            req = self._pending.popleft()
            g = groups.setdefault(req.event, [])
            n = sum(r.rows.shape[0] for r in g)
            if g and n + req.rows.shape[0] > self.max_batch_rows:
                keep.append(req)
            else:
                g.append(req)
        self._pending = keep
        return groups

    def _loop(self):
        while True:
            with self._cv:
                while not self._pending and not self._stop:
                    self._cv.wait()
                if self._stop and not self._pending:
                    return
                groups = self._take_round()
            launched = []
            multi_groups = {}
            for name in list(groups):
                m = self.events[name].get("multi")
                if m is not None:
                    multi_groups.setdefault(m[0], []).append((m[1], name, groups.pop(name)))
            for gid, members in multi_groups.items():
                multi = self._multis[gid]
                multi.rows = [0] * len(multi.rows)
                for k, name, reqs in members:
                    rows = np.concatenate([r.rows for r in reqs]) if len(reqs) > 1 else reqs[0].rows
                    multi.set_parameters(k, rows)
                outs = multi.run(synchronize=False)
                for k, name, reqs in members:
                    with torch.cuda.stream(multi.stream):
                        out_host = torch.empty(outs[k].shape, dtype=torch.float64).pin_memory()
                        out_host.copy_(outs[k], non_blocking=True)
                        done = torch.cuda.Event()
                        done.record(multi.stream)
                    launched.append((reqs, out_host, done, outs[k]))
            for k, (name, reqs) in enumerate(groups.items()):
                stream = self.streams[k % len(self.streams)]
                lk = self.events[name]["likelihood"]
                rows = np.concatenate([r.rows for r in reqs]) if len(reqs) > 1 else reqs[0].rows
                with torch.cuda.stream(stream):
                    host = torch.from_numpy(rows).pin_memory()
                    dev_rows = host.to(self._dev, non_blocking=True)
                    out = lk.log_likelihood_ratio_batch(dev_rows, stream=stream.cuda_stream)
                    out_host = torch.empty(out.shape, dtype=torch.float64).pin_memory()
                    out_host.copy_(out, non_blocking=True)
                    done = torch.cuda.Event()
                    done.record(stream)
                launched.append((reqs, out_host, done, host, dev_rows, out))
            for reqs, out_host, done, *_keep in launched:
                done.synchronize()
                t = time.perf_counter()
                vals = out_host.numpy()
                o = 0
                for r in reqs:
                    n = r.rows.shape[0]
                    r.t_done = t
                    r.future.set_result(vals[o:o + n].copy())
                    o += n
                    self.completed.append((r.event, n, r.t_submit, r.t_done))
                self.rows_done += o
            self.rounds += 1


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


def latency_summary(server, since=0.0):
    lat = np.array([(t1 - t0) for (_, _, t0, t1) in server.completed if t0 >= since])
    rows = np.array([n for (_, n, t0, _) in server.completed if t0 >= since])
    if len(lat) == 0:
        return {}
    return dict(requests=int(len(lat)), rows=int(rows.sum()),
                p50_ms=float(np.percentile(lat, 50) * 1e3), p99_ms=float(np.percentile(lat, 99) * 1e3),
                mean_ms=float(lat.mean() * 1e3))
