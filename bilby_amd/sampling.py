"""Sampler-side batching and worker-process clients for the GPU likelihood.

bilby's samplers call the likelihood one parameter dict at a time: dynesty
through ``_log_likelihood_wrapper`` (bilby/core/sampler/dynesty.py:40-53),
mapped over a queue of ``queue_size`` proposals with ``pool.map``
(dynesty.py:750-753, 808-810); emcee, ptemcee and bilby_mcmc through
``LikePriorEvaluator`` (bilby/core/sampler/base_sampler.py:1108-1185).  bilby
parallelises with a ``multiprocessing.Pool`` whose workers each receive the
likelihood, priors and search keys through an initializer
(base_sampler.py:44-67, 772-803).  A GPU likelihood cannot be shipped into
forked workers, and one dict per kernel launch leaves the device idle.

``BatchedPool``
    replaces that pool; bilby takes a user pool from ``kwargs["pool"]``
    (base_sampler.py:773-775).  ``map`` over a likelihood wrapper evaluates
    the whole queue of proposals as ONE batched GPU call, with the same
    constraint check and ``nan_to_num(-inf)`` for rejected points as the
    wrapper; any other function (prior transforms) is mapped in the calling
    process.

``LikelihoodService`` / ``RemoteLikelihood``
    keep the GPU in one process and serve other processes through the native
    continuous-batching scheduler (``serving.LikelihoodServer``) on a Unix
    socket that the scheduler itself reads.  ``RemoteLikelihood`` is the
    picklable stub a sampler's worker processes hold instead of the
    likelihood: the same method names, a dict in and a float out, or a dict of
    arrays in and an array out; it converts dicts to rows in its own process.

This module imports neither torch nor the HIP library, so worker processes
that only hold a ``RemoteLikelihood`` never touch the GPU.
"""
import os
import shutil
import socket
import struct
import sys
import tempfile
import threading
import time

import numpy as np

from .abi import NPARAM, WIRE_REPLY, WIRE_REQUEST

_REJECTED = float(np.nan_to_num(-np.inf))


class SamplingState:
    """What bilby keeps in ``_sampling_convenience_dump``
    (base_sampler.py:44-49): the likelihood, the priors, the sampled keys,
    whether the ratio is used, and the full parameter dict (fixed values)."""

    def __init__(self, likelihood=None, priors=None, search_parameter_keys=None, use_ratio=True,
                 parameters=None):
        self.likelihood = likelihood
        self.priors = priors
        self.search_parameter_keys = list(search_parameter_keys or [])
        self.use_ratio = use_ratio
        self.parameters = dict(parameters or {})


_state = SamplingState()


def initialize_global_variables(likelihood, priors, search_parameter_keys, use_ratio, parameters):
    """base_sampler.py:52-67: the process-wide state the wrappers below read."""
    _state.likelihood = likelihood
    _state.priors = priors
    _state.search_parameter_keys = list(search_parameter_keys)
    _state.use_ratio = use_ratio
    _state.parameters = dict(parameters or {})


def evaluate_points(state, thetas):
    """Log-likelihood (ratio) of a queue of points in one batched call.

    dynesty.py:40-53 point by point: the sampled values overwrite the fixed
    parameters; a point that fails ``priors.evaluate_constraints`` (on the
    sampled values) gets ``nan_to_num(-inf)`` without being evaluated; else
    ``log_likelihood_ratio`` (``use_ratio``) or ``log_likelihood``.
    thetas: [n, len(search_parameter_keys)] -> float64 [n].
    """
    keys = state.search_parameter_keys
    thetas = np.asarray(thetas, dtype=float)
    if thetas.ndim == 1:
        thetas = thetas.reshape(-1, len(keys)) if len(keys) else thetas.reshape(-1, 0)
    if thetas.ndim != 2 or thetas.shape[1] != len(keys):
        raise ValueError("points must be [n, %d] (one value per search parameter)" % len(keys))
    n = thetas.shape[0]
    out = np.full(n, _REJECTED)
    if n == 0:
        return out
    sampled = {k: thetas[:, j] for j, k in enumerate(keys)}
    ok = np.ones(n, dtype=bool)
    if state.priors is not None and hasattr(state.priors, "evaluate_constraints") and keys:
        ok = np.broadcast_to(np.asarray(state.priors.evaluate_constraints(dict(sampled)),
                                        dtype=bool), (n,)).copy()
    if not ok.any():
        return out
    params = {k: v for k, v in state.parameters.items() if k not in sampled and v is not None}
    params.update({k: v[ok] for k, v in sampled.items()})
    if not keys:
        params = {k: np.full(int(ok.sum()), v) for k, v in params.items()}
    lnl = np.asarray(state.likelihood.log_likelihood_ratio_batch(params), dtype=float)
    if not state.use_ratio:
        lnl = lnl + state.likelihood.noise_log_likelihood()
    out[ok] = lnl
    return out


def log_likelihood_wrapper(theta):
    """dynesty.py:40-53 against the state set by initialize_global_variables
    (one point; BatchedPool.map evaluates a queue of them at once)."""
    return float(evaluate_points(_state, [theta])[0])


def _bilby_state(func):
    """The module-level container bilby fills before sampling
    (base_sampler.py:44-67, 797-803), found from the wrapper's own package."""
    module = getattr(func, "__module__", "") or ""
    package = module.rpartition(".")[0]
    base = sys.modules.get(package + ".base_sampler") if package else None
    dump = getattr(base, "_sampling_convenience_dump", None)
    if dump is None or getattr(dump, "likelihood", None) is None:
        raise RuntimeError("%s.%s: bilby's sampling state is not initialised"
                           % (module, getattr(func, "__name__", func)))
    return SamplingState(dump.likelihood, dump.priors, dump.search_parameter_keys,
                         dump.use_ratio, dump.parameters)


class BatchedPool:
    """Pool for bilby samplers whose ``map`` batches likelihood calls.

    Recognised likelihood functions: this module's ``log_likelihood_wrapper``,
    bilby's dynesty ``_log_likelihood_wrapper``, and ``LikePriorEvaluator``
    instances (``__call__`` returns ``(logl, logp)``; base_sampler.py:1173-1185).
    With ``likelihood`` given, those use this pool's own state; otherwise the
    state set by ``initialize_global_variables`` (or bilby's own container
    for bilby's wrapper).  Every other function is mapped in this process.
    """

    def __init__(self, likelihood=None, priors=None, search_parameter_keys=None, use_ratio=True,
                 parameters=None):
        self._own = (SamplingState(likelihood, priors, search_parameter_keys, use_ratio, parameters)
                     if likelihood is not None else None)
        self.size = 1
        self.batches = 0
        self.points = 0

    def _state_for(self, func, kind):
        if self._own is not None:
            return self._own
        if func is log_likelihood_wrapper:
            if _state.likelihood is None:
                raise RuntimeError("initialize_global_variables has not been called")
            return _state
        return _bilby_state(type(func).__call__ if kind == "logl_logp" else func)

    @staticmethod
    def _kind(func):
        if func is log_likelihood_wrapper or getattr(func, "__name__", "") in (
                "_log_likelihood_wrapper", "log_likelihood_wrapper"):
            return "logl"
        if type(func).__name__ == "LikePriorEvaluator" and hasattr(func, "logp"):
            return "logl_logp"
        return None

    def map(self, func, iterable, chunksize=None):
        kind = self._kind(func)
        if kind is None:
            return [func(x) for x in iterable]
        # one array conversion of the whole queue: converting point by point
        # costs more host time than the batched evaluation itself
        thetas = np.asarray(iterable if isinstance(iterable, np.ndarray) else list(iterable),
                            dtype=float)
        if thetas.ndim == 0 or len(thetas) == 0:
            return []
        thetas = thetas.reshape(len(thetas), -1)
        state = self._state_for(func, kind)
        self.batches += 1
        self.points += len(thetas)
        if kind == "logl":
            return evaluate_points(state, thetas).tolist()
        # LikePriorEvaluator.__call__ (base_sampler.py:1173-1185): the prior per
        # point; the likelihood only where the prior is finite (else 0)
        lp = np.array([func.logp(x) for x in thetas], dtype=float)
        if np.isnan(lp).any():
            raise ValueError("Prior function returned NaN.")
        live = lp != -np.inf
        ll = np.zeros(len(thetas))
        if live.any():
            ll[live] = evaluate_points(state, thetas[live])
            if np.isnan(ll).any():
                raise ValueError("Log likelihood function returned NaN.")
        return list(zip(ll.tolist(), lp.tolist()))

    def imap(self, func, iterable, chunksize=None):
        return iter(self.map(func, iterable))

    def starmap(self, func, iterable, chunksize=None):
        return [func(*args) for args in iterable]

    def close(self):
        pass

    def join(self):
        pass

    def terminate(self):
        pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False


class LikelihoodService:
    """Serve GPU likelihoods to sampler worker processes.

    ``likelihoods``: name -> GravitationalWaveTransient or
    ROQGravitationalWaveTransient, made resident on the GPU of this process
    and registered with the native scheduler, which listens on a Unix socket
    in a private directory (gwl_server_listen).  Requests from every client
    process go straight to the native scheduler: no Python runs per request
    on this side, and concurrent requests of an event share one launch.
    """

    def __init__(self, likelihoods, hbm_budget=200e9, max_batch_rows=8192, device=0):
        from .serving import LikelihoodServer
        self.server = LikelihoodServer(hbm_budget=hbm_budget, max_batch_rows=max_batch_rows,
                                       device=device)
        self._dir = tempfile.mkdtemp(prefix="gwl-service-")
        self.address = os.path.join(self._dir, "socket")
        self._clients = {}
        try:
            for name, lk in dict(likelihoods).items():
                self.server.register(name, lk)
                info = dict(noise_log_likelihood=float(lk.noise_log_likelihood()),
                            marginalized_parameters=list(getattr(lk, "marginalized_parameters",
                                                                 [])),
                            meta_data=_plain(getattr(lk, "meta_data", {})),
                            max_rows=int(max_batch_rows))
                self._clients[name] = (self.server.events[name], info, lk.row_builder())
            self.server.listen(self.address)
        except Exception:
            self.close()
            raise

    def client(self, name):
        """A picklable stub for ``name`` to hand to worker processes."""
        unit, info, rows = self._clients[name]
        return RemoteLikelihood(self.address, unit, info, rows)

    def close(self):
        if self.server is not None:
            self.server.close()
            self.server = None
        shutil.rmtree(self._dir, ignore_errors=True)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False


def _plain(d):
    out = {}
    for k, v in dict(d).items():
        if isinstance(v, (bool, int, float, str, type(None))):
            out[k] = v
        else:
            out[k] = repr(v)
    return out


_HEADER = struct.Struct("=Iiq")


def _recv_exact(sock, nbytes):
    buf = bytearray(nbytes)
    view = memoryview(buf)
    got = 0
    while got < nbytes:
        k = sock.recv_into(view[got:], nbytes - got)
        if k == 0:
            raise ConnectionError("likelihood service closed the connection")
        got += k
    return buf


class RemoteLikelihood:
    """Client stub of one served likelihood (bilby/core/likelihood.py:13-73
    method names): ``log_likelihood_ratio(dict) -> float``,
    ``log_likelihood``, ``noise_log_likelihood``, ``marginalized_parameters``,
    ``meta_data`` and ``log_likelihood_ratio_batch``.  Rows are built in the
    calling process (``abi.RowBuilder``) and sent in the wire format of
    gwl_server_listen (include/gwl.h).  Pickling drops the connection; each
    process and thread opens its own on first use."""

    def __init__(self, address, unit, info, rows):
        self._address = address
        self.unit = int(unit)
        self._info = dict(info)
        self.rows_from_parameters = rows
        self._local = threading.local()

    def __getstate__(self):
        d = dict(self.__dict__)
        d.pop("_local")
        return d

    def __setstate__(self, d):
        self.__dict__.update(d)
        self._local = threading.local()

    def _socket(self):
        sock = getattr(self._local, "sock", None)
        if sock is None or getattr(self._local, "pid", None) != os.getpid():
            sock = socket.socket(socket.AF_UNIX, socket.SOCK_STREAM)
            sock.connect(self._address)
            self._local.sock = sock
            self._local.pid = os.getpid()
        return sock

    def _evaluate_rows(self, rows):
        rows = np.ascontiguousarray(rows, dtype=np.float64)
        if rows.ndim != 2 or rows.shape[1] != NPARAM:
            raise ValueError("rows must be [n, %d]" % NPARAM)
        out = np.empty(rows.shape[0])
        step = self._info["max_rows"]
        sock = self._socket()
        for a in range(0, rows.shape[0], step):
            chunk = rows[a:a + step]
            sock.sendall(_HEADER.pack(WIRE_REQUEST, self.unit, chunk.shape[0]) + chunk.tobytes())
            magic, status, n = _HEADER.unpack(_recv_exact(sock, _HEADER.size))
            body = _recv_exact(sock, n if status else 8 * n)
            if magic != WIRE_REPLY:
                raise ConnectionError("likelihood service: bad reply header")
            if status:
                raise RuntimeError("likelihood service (status %d): %s"
                                   % (status, bytes(body).decode(errors="replace")))
            out[a:a + chunk.shape[0]] = np.frombuffer(body, dtype=np.float64)
        return out

    def log_likelihood_ratio_batch(self, parameters):
        """dict of 1-D arrays (bilby names) or float64 rows [n, 16] -> [n]."""
        if isinstance(parameters, dict):
            parameters = self.rows_from_parameters(parameters)
        return self._evaluate_rows(parameters)

    def log_likelihood_ratio(self, parameters):
        """base.py:419-446 through the service (the caller's dict is not
        modified)."""
        return float(self._evaluate_rows(self.rows_from_parameters(parameters))[0])

    def log_likelihood(self, parameters):
        return self.log_likelihood_ratio(parameters) + self.noise_log_likelihood()

    def noise_log_likelihood(self):
        return self._info["noise_log_likelihood"]

    @property
    def marginalized_parameters(self):
        return list(self._info["marginalized_parameters"])

    @property
    def meta_data(self):
        return dict(self._info["meta_data"])

    def close(self):
        sock = getattr(self._local, "sock", None)
        if sock is not None:
            sock.close()
            self._local.sock = None


def client_load(likelihood, draws, seconds, start_index=0):
    """One sampler process calling ``likelihood.log_likelihood_ratio`` one
    dict at a time, as bilby's pool workers do, for ``seconds``: returns the
    number of calls, their latencies (s) and the first values (row index ->
    lnL) for checking."""
    keys = list(draws)
    n = len(draws[keys[0]])
    lat = []
    values = {}
    i = start_index
    t_end = time.perf_counter() + seconds
    while time.perf_counter() < t_end:
        p = {k: float(draws[k][i % n]) for k in keys}
        t0 = time.perf_counter()
        v = likelihood.log_likelihood_ratio(p)
        lat.append(time.perf_counter() - t0)
        if len(values) < 64:
            values[i % n] = v
        i += 1
    return len(lat), np.asarray(lat), values
