"""GPU checks of the continuous-batching server: results are the same alone
and in a mixed load (bitwise without time marginalisation, 1e-9 with it),
equal the direct batched call, and admission honours the HBM budget."""
import sys
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def events():
    from bilby_amd.configs import make_sine_gaussian_plain, make_taylorf2_marginalised
    from bilby_amd.likelihood import GravitationalWaveTransient
    from tests.helpers import fixed_marginalised
    sg = make_sine_gaussian_plain(n_draws=512, seed=3)
    lk_sg = GravitationalWaveTransient(sg["interferometers"], sg["waveform_generator"],
                                       precision="fp64")
    tf = make_taylorf2_marginalised(n_draws=512, seed=4, duration=4.0, sampling_frequency=2048.0,
                                    detectors=("H1", "L1"))
    pri = tf["priors"].copy()
    lk_tf = GravitationalWaveTransient(tf["interferometers"], tf["waveform_generator"],
                                       time_marginalization=True, distance_marginalization=True,
                                       phase_marginalization=True, priors=pri)
    rows_sg = lk_sg.rows_from_parameters(sg["draws"])
    rows_tf = lk_tf.rows_from_parameters(fixed_marginalised(tf["draws"], pri))
    return dict(sg=(lk_sg, rows_sg), tf=(lk_tf, rows_tf))


def test_same_alone_and_in_the_mix(events):
    from bilby_amd.serving import LikelihoodServer, run_load
    say("serving alone vs mix")
    server = LikelihoodServer(hbm_budget=20e9, max_batch_rows=2048)
    try:
        for name, (lk, _) in events.items():
            server.register(name, lk)
        probe = {name: rows[:37] for name, (_, rows) in events.items()}
        alone = {name: server.submit(name, r).result() for name, r in probe.items()}
        direct = {name: events[name][0].log_likelihood_ratio_batch(r) for name, r in probe.items()}
        mix = [(1.0, [(n, rows) for n, (_, rows) in events.items()])]
        results = {}

        def probe_during_load():
            for name, r in probe.items():
                results[name] = server.submit(name, r).result()

        th = threading.Thread(target=lambda: run_load(server, mix, n_samplers=8, rate=400.0,
                                                      duration=1.5, rows_range=(1, 256), seed=1))
        th.start()
        probe_during_load()
        th.join()
        for name in probe:
            tol = 0.0 if name == "sg" else 1e-9
            assert np.max(np.abs(results[name] - alone[name])) <= tol, name
            assert np.max(np.abs(direct[name] - alone[name])) <= tol, name
        say("rounds", server.rounds, "rows", server.rows_done)
        assert server.rounds > 1
    finally:
        server.close()


def test_hbm_budget_admission(events):
    from bilby_amd.serving import HBMBudgetExceeded, LikelihoodServer
    server = LikelihoodServer(hbm_budget=1e5, max_batch_rows=64)
    try:
        with pytest.raises(HBMBudgetExceeded):
            server.register("tf", events["tf"][0])
    finally:
        server.close()
