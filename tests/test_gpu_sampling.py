"""GPU checks of the sampler-side callers (bilby_amd.sampling): a queue of
proposals through BatchedPool.map equals the same points called one dict at
a time, constraint-rejected points get nan_to_num(-inf); spawned worker
processes holding RemoteLikelihood stubs get the values of a direct batched
call from the GPU-owning LikelihoodService."""
import multiprocessing as mp
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ["mass_1", "mass_2", "ra", "dec", "theta_jn", "psi"]


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def event():
    from bilby_amd.configs import make_taylorf2_marginalised
    from bilby_amd.likelihood import GravitationalWaveTransient
    from tests.helpers import fixed_marginalised
    s = make_taylorf2_marginalised(n_draws=96, seed=23, duration=4.0, sampling_frequency=2048.0)
    priors = s["priors"].copy()
    lk = GravitationalWaveTransient(s["interferometers"], s["waveform_generator"],
                                    time_marginalization=True, distance_marginalization=True,
                                    phase_marginalization=True, priors=priors, precision="fp64")
    return lk, fixed_marginalised(s["draws"], priors)


def test_batched_pool_equals_single_dict_calls(event):
    from bilby_amd.prior import Constraint, PriorDict
    from bilby_amd.sampling import BatchedPool, initialize_global_variables, log_likelihood_wrapper
    lk, draws = event
    say("batched pool vs single dicts")
    fixed = {k: float(v[0]) for k, v in draws.items() if k not in KEYS}
    thetas = np.stack([draws[k] for k in KEYS], axis=1)
    priors = PriorDict(mass_1=Constraint(0.0, 40.0))
    for use_ratio in (True, False):
        initialize_global_variables(lk, priors, KEYS, use_ratio, fixed)
        pool = BatchedPool()
        got = np.array(pool.map(log_likelihood_wrapper, thetas))
        want = []
        for t in thetas:
            p = dict(fixed)
            p.update(zip(KEYS, t))
            if not p["mass_1"] < 40.0:
                want.append(np.nan_to_num(-np.inf))
            else:
                want.append(lk.log_likelihood_ratio(p) if use_ratio else lk.log_likelihood(p))
        want = np.array(want)
        rejected = thetas[:, 0] >= 40.0
        say("use_ratio", use_ratio, "rejected", int(rejected.sum()), "max diff",
            np.max(np.abs(got[~rejected] - want[~rejected])))
        assert rejected.any() and not rejected.all()
        assert np.all(got[rejected] == np.nan_to_num(-np.inf))
        assert np.max(np.abs(got[~rejected] - want[~rejected])) < 1e-8
        assert pool.batches == 1


def _worker(client, dicts):
    return [client.log_likelihood_ratio(p) for p in dicts], client.noise_log_likelihood()


def test_worker_processes_through_service(event):
    from bilby_amd.sampling import LikelihoodService, RemoteLikelihood
    from tests.helpers import draws_as_dicts
    lk, draws = event
    say("service with 4 spawned client processes")
    direct = lk.log_likelihood_ratio_batch(draws)
    dicts = draws_as_dicts(draws, range(len(direct)))
    with LikelihoodService({"c2": lk}, hbm_budget=20e9, max_batch_rows=64) as svc:
        client = svc.client("c2")
        parts = [list(range(w, len(dicts), 4)) for w in range(4)]
        with mp.get_context("spawn").Pool(4) as pool:
            res = pool.starmap(_worker, [(client, [dicts[i] for i in part]) for part in parts])
        batch = client.log_likelihood_ratio_batch(draws)     # 96 rows > 64: two requests
        bad = RemoteLikelihood(svc.address, 999, client._info, client.rows_from_parameters)
        with pytest.raises(RuntimeError, match="unknown id"):
            bad.log_likelihood_ratio(dicts[0])
        client.close()
    served = np.empty(len(direct))
    for part, (vals, noise) in zip(parts, res):
        served[part] = vals
        assert noise == lk.noise_log_likelihood()
    say("served vs direct", np.max(np.abs(served - direct)), "batch", np.max(np.abs(batch - direct)))
    assert np.max(np.abs(served - direct)) < 1e-9
    assert np.max(np.abs(batch - direct)) < 1e-9
