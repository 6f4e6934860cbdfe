"""Sampler-side batching (bilby_amd.sampling) on the CPU, with a host-only
stand-in likelihood: BatchedPool.map against the per-point semantics of
bilby's wrappers (dynesty.py:40-53, base_sampler.py:1148-1185), and the
worker-process client protocol through spawned processes."""
import multiprocessing as mp
import pickle
import sys
import types

import numpy as np
import pytest

REJECTED = np.nan_to_num(-np.inf)


class Quadratic:
    """Host-only stand-in with the likelihood's batch API."""

    def log_likelihood_ratio_batch(self, p):
        x = np.asarray(p["x"], float)
        y = np.asarray(p["y"], float)
        return -0.5 * ((x - 1.0) ** 2 + (y + 2.0) ** 2) + np.asarray(p.get("c", 0.0), float)

    def log_likelihood_ratio(self, p):
        return float(self.log_likelihood_ratio_batch({k: np.atleast_1d(v) for k, v in p.items()})[0])

    def noise_log_likelihood(self):
        return -10.0


def _priors():
    from bilby_amd.prior import Constraint, PriorDict, Uniform
    return PriorDict(x=Uniform(-5, 5), y=Uniform(-5, 5), x_plus_y=Constraint(-3.0, 3.0),
                     conversion_function=_add_sum)


def _add_sum(sample):
    out = dict(sample)
    out["x_plus_y"] = np.asarray(sample["x"]) + np.asarray(sample["y"])
    return out


def _per_point(lk, priors, keys, fixed, theta, use_ratio):
    """dynesty.py:40-53, one point."""
    sampling = {k: t for k, t in zip(keys, theta)}
    params = dict(fixed)
    params.update(sampling)
    if not priors.evaluate_constraints(sampling):
        return REJECTED
    return lk.log_likelihood_ratio(params) if use_ratio else (
        lk.log_likelihood_ratio(params) + lk.noise_log_likelihood())


@pytest.mark.parametrize("use_ratio", [True, False])
def test_batched_pool_matches_per_point_wrapper(use_ratio):
    from bilby_amd.sampling import BatchedPool, initialize_global_variables, log_likelihood_wrapper
    lk, priors, keys, fixed = Quadratic(), _priors(), ["x", "y"], {"x": 0.0, "y": 0.0, "c": 0.25}
    thetas = np.random.default_rng(3).uniform(-5, 5, (257, 2))
    expect = [_per_point(lk, priors, keys, fixed, t, use_ratio) for t in thetas]
    assert REJECTED in expect and any(v != REJECTED for v in expect)
    initialize_global_variables(lk, priors, keys, use_ratio, fixed)
    pool = BatchedPool()
    got = pool.map(log_likelihood_wrapper, thetas)
    assert pool.batches == 1 and pool.points == len(thetas)
    assert np.allclose(got, expect, rtol=0, atol=1e-12)
    assert [log_likelihood_wrapper(t) for t in thetas[:5]] == pytest.approx(expect[:5], abs=1e-12)
    own = BatchedPool(lk, priors, keys, use_ratio, fixed)
    assert np.allclose(own.map(log_likelihood_wrapper, list(thetas)), expect, rtol=0, atol=1e-12)
    # other functions (a prior transform) are mapped point by point
    assert own.map(lambda t: priors.rescale(keys, t).tolist(), [[0.5, 0.0]]) == [[0.0, -5.0]]
    assert np.allclose(own.map(log_likelihood_wrapper, (tuple(t) for t in thetas)), expect,
                       rtol=0, atol=1e-12)
    assert own.map(log_likelihood_wrapper, []) == []
    with pytest.raises(ValueError):
        own.map(log_likelihood_wrapper, [[0.0, 1.0, 2.0]])
    with pytest.raises(ValueError):  # a ragged queue
        own.map(log_likelihood_wrapper, [[0.0, 1.0], [2.0]])


def test_pool_reads_bilby_sampling_state():
    """bilby's dynesty wrapper and LikePriorEvaluator read
    base_sampler._sampling_convenience_dump; the pool finds it the same way."""
    from bilby_amd.sampling import BatchedPool
    lk, priors, keys, fixed = Quadratic(), _priors(), ["x", "y"], {"c": 0.0}
    names = ["fakebilby", "fakebilby.core", "fakebilby.core.sampler",
             "fakebilby.core.sampler.base_sampler", "fakebilby.core.sampler.dynesty"]
    mods = {n: types.ModuleType(n) for n in names}
    dump = types.SimpleNamespace(likelihood=lk, priors=priors, search_parameter_keys=keys,
                                 use_ratio=True, parameters=fixed)
    mods["fakebilby.core.sampler.base_sampler"]._sampling_convenience_dump = dump

    def _log_likelihood_wrapper(theta):
        return _per_point(lk, priors, keys, fixed, theta, True)
    _log_likelihood_wrapper.__module__ = "fakebilby.core.sampler.dynesty"

    class LikePriorEvaluator:
        def logp(self, v):
            return -np.inf if v[0] > 4.0 else float(np.log(0.01))

        def logl(self, v):
            return _per_point(lk, priors, keys, fixed, v, True)

        def __call__(self, x):  # base_sampler.py:1173-1185
            lp = self.logp(x)
            return (0, lp) if lp == -np.inf else (self.logl(x), lp)
    LikePriorEvaluator.__module__ = "fakebilby.core.sampler.base_sampler"
    LikePriorEvaluator.__call__.__module__ = "fakebilby.core.sampler.base_sampler"

    saved = {n: sys.modules.get(n) for n in names}
    sys.modules.update(mods)
    try:
        thetas = np.random.default_rng(5).uniform(-5, 5, (64, 2))
        pool = BatchedPool()
        got = pool.map(_log_likelihood_wrapper, thetas)
        assert np.allclose(got, [_log_likelihood_wrapper(t) for t in thetas], rtol=0, atol=1e-12)
        ev = LikePriorEvaluator()
        pairs = pool.map(ev, thetas)
        want = [ev(t) for t in thetas]
        assert [p[1] for p in pairs] == [w[1] for w in want]
        assert np.allclose([p[0] for p in pairs], [w[0] for w in want], rtol=0, atol=1e-12)
        dump.likelihood = None
        with pytest.raises(RuntimeError):
            pool.map(_log_likelihood_wrapper, thetas)
    finally:
        for n, m in saved.items():
            if m is None:
                sys.modules.pop(n, None)
            else:
                sys.modules[n] = m


def _stand_in_server(path, ready):
    """A Python stand-in that speaks gwl_server_listen's wire format
    (include/gwl.h): unit 3 answers -((ra - 1)^2 + (dec + 2)^2) / 2 per row,
    any other unit an error reply."""
    import socket
    import struct
    import threading
    from bilby_amd.abi import NPARAM, P_DEC, P_RA, WIRE_REPLY, WIRE_REQUEST
    hdr = struct.Struct("=Iiq")
    lsock = socket.socket(socket.AF_UNIX, socket.SOCK_STREAM)
    lsock.bind(path)
    lsock.listen(8)
    ready.set()

    def exact(c, n):
        buf = b""
        while len(buf) < n:
            part = c.recv(n - len(buf))
            if not part:
                raise EOFError
            buf += part
        return buf

    def serve(c):
        try:
            while True:
                magic, unit, n = hdr.unpack(exact(c, 16))
                assert magic == WIRE_REQUEST
                rows = np.frombuffer(exact(c, 8 * NPARAM * n)).reshape(n, NPARAM)
                if unit != 3:
                    msg = b"unknown id"
                    c.sendall(hdr.pack(WIRE_REPLY, 2, len(msg)) + msg)
                    continue
                out = -0.5 * ((rows[:, P_RA] - 1.0) ** 2 + (rows[:, P_DEC] + 2.0) ** 2)
                c.sendall(hdr.pack(WIRE_REPLY, 0, n) + out.tobytes())
        except EOFError:
            pass
        finally:
            c.close()

    while True:
        try:
            c, _ = lsock.accept()
        except OSError:
            return
        threading.Thread(target=serve, args=(c,), daemon=True).start()


def _client_worker(client, points):
    single = [client.log_likelihood_ratio({"ra": x, "dec": y}) for x, y in points]
    batch = client.log_likelihood_ratio_batch({"ra": points[:, 0], "dec": points[:, 1]})
    return single, batch.tolist(), client.log_likelihood({"ra": 1.0, "dec": -2.0})


def test_remote_client_wire_format_in_worker_processes(tmp_path):
    import threading
    from bilby_amd.abi import RowBuilder
    from bilby_amd.sampling import RemoteLikelihood
    from bilby_amd.waveform_generator import convert_to_lal_binary_black_hole_parameters
    path = str(tmp_path / "socket")
    ready = threading.Event()
    threading.Thread(target=_stand_in_server, args=(path, ready), daemon=True).start()
    assert ready.wait(10)
    info = dict(noise_log_likelihood=-10.0, marginalized_parameters=["geocent_time"],
                meta_data={}, max_rows=7)
    rows = RowBuilder(convert_to_lal_binary_black_hole_parameters, {"geocent_time": 5.0})
    client = RemoteLikelihood(path, 3, info, rows)
    assert client.log_likelihood_ratio({"ra": 1.0, "dec": -2.0}) == 0.0
    clone = pickle.loads(pickle.dumps(client))
    assert clone.marginalized_parameters == ["geocent_time"]
    rng = np.random.default_rng(9)
    jobs = [(clone, rng.uniform(-3, 3, (20, 2))) for _ in range(3)]   # 20 rows > max_rows: chunks
    with mp.get_context("spawn").Pool(3) as pool:
        results = pool.starmap(_client_worker, jobs)
    for (_, pts), (single, batch, ll) in zip(jobs, results):
        want = -0.5 * ((pts[:, 0] - 1.0) ** 2 + (pts[:, 1] + 2.0) ** 2)
        assert np.array_equal(single, want) and np.array_equal(batch, want)
        assert ll == -10.0
    bad = RemoteLikelihood(path, 4, info, rows)
    with pytest.raises(RuntimeError, match="unknown id"):
        bad.log_likelihood_ratio({"ra": 0.0, "dec": 0.0})
    with pytest.raises(ValueError):
        client.log_likelihood_ratio_batch(np.zeros((2, 5)))
    client.close()


def test_row_builder_columns_and_defaults():
    """Conversion (chirp mass / mass ratio, a_i cos tilt_i for aligned spins;
    TaylorF2 refuses in-plane spin), marginalised-key defaults that an
    explicit value overrides, and the gwl_param columns."""
    from bilby_amd import abi
    from bilby_amd.waveform_generator import convert_to_lal_binary_black_hole_parameters
    rb = pickle.loads(pickle.dumps(abi.RowBuilder(convert_to_lal_binary_black_hole_parameters,
                                                  {"geocent_time": 7.0, "phase": 0.0})))
    q, mc = np.array([0.5, 0.8]), np.array([20.0, 30.0])
    rows = rb(dict(chirp_mass=mc, mass_ratio=q, a_1=0.3, tilt_1=np.pi, ra=1.0, dec=0.5,
                   phase=np.array([1.0, 2.0])))
    with pytest.raises(ValueError):
        rb(dict(chirp_mass=mc, mass_ratio=q, a_1=0.3, tilt_1=np.pi / 3, ra=1.0, dec=0.5))
    total = mc * (1 + q) ** 1.2 / q ** 0.6
    assert rows.shape == (2, abi.NPARAM)
    assert np.allclose(rows[:, abi.P_MASS_1], total / (1 + q), rtol=1e-14)
    assert np.allclose(rows[:, abi.P_MASS_2], total / (1 + q) * q, rtol=1e-14)
    assert np.allclose(rows[:, abi.P_CHI_1], -0.3, rtol=1e-14)
    assert np.all(rows[:, abi.P_GEOCENT_TIME] == 7.0)
    assert np.all(rows[:, abi.P_PHASE] == [1.0, 2.0])
    assert np.all(rows[:, abi.P_RA] == 1.0) and np.all(rows[:, abi.P_CHI_2] == 0.0)


def test_row_builder_scalar_and_array_inputs_agree():
    """One dict of Python floats, numpy float64 / float32 scalars, ints, or
    1-element arrays gives the same row, and a dict of arrays gives the rows of
    its points one by one."""
    from bilby_amd import abi
    from bilby_amd.waveform_generator import convert_to_lal_binary_black_hole_parameters
    rb = abi.RowBuilder(convert_to_lal_binary_black_hole_parameters, {"geocent_time": 5.0})
    rng = np.random.default_rng(2)
    n = 7
    arrays = dict(mass_1=rng.uniform(20, 40, n), mass_2=rng.uniform(5, 20, n),
                  chi_1=rng.uniform(-0.5, 0.5, n), chi_2=np.zeros(n),
                  luminosity_distance=rng.uniform(100, 900, n), theta_jn=rng.uniform(0, 3, n),
                  phase=rng.uniform(0, 6, n), ra=rng.uniform(0, 6, n), dec=rng.uniform(-1, 1, n),
                  psi=rng.uniform(0, 3, n), geocent_time=np.full(n, 1.25))
    rows = rb(arrays)
    assert rows.shape == (n, abi.NPARAM)
    for i in range(n):
        point = {k: float(v[i]) for k, v in arrays.items()}
        np.testing.assert_array_equal(rb(point), rows[i:i + 1])
        np.testing.assert_array_equal(rb({k: np.float64(v) for k, v in point.items()}),
                                      rows[i:i + 1])
        np.testing.assert_array_equal(rb({k: np.array([v]) for k, v in point.items()}),
                                      rows[i:i + 1])
    f32 = {k: np.float32(v[0]) for k, v in arrays.items()}
    np.testing.assert_array_equal(rb(f32), rb({k: float(v) for k, v in f32.items()}))
    mixed = {k: float(v[0]) for k, v in arrays.items()}
    mixed["geocent_time"] = 2
    assert rb(mixed)[0, abi.PARAM_COLUMNS["geocent_time"]] == 2.0
