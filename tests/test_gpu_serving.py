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


def test_group_and_roq_units_match_direct_calls():
    """Events served as one multi-event group, and a ROQ likelihood, give the
    values of their direct batched calls (1e-8; the group splits each row's
    transform into a different number of float64 partial groups), for
    requests interleaved across the units."""
    from bilby_amd.configs import make_roq, make_uneven_events
    from bilby_amd.likelihood import GravitationalWaveTransient
    from bilby_amd.roq import ROQGravitationalWaveTransient
    from bilby_amd.serving import LikelihoodServer
    from tests.helpers import fixed_marginalised
    say("serving group + roq")
    events = make_uneven_events(n_events=3, seed=8, rows_range=(20, 40), durations=(4, 8))
    names, lks, pools = [], [], []
    for k, ev in enumerate(events):
        pri = ev["priors"].copy()
        lk = GravitationalWaveTransient(ev["interferometers"], ev["waveform_generator"],
                                        time_marginalization=True, distance_marginalization=True,
                                        phase_marginalization=True, priors=pri)
        names.append("g%d" % k)
        lks.append(lk)
        pools.append(lk.rows_from_parameters(fixed_marginalised(ev["draws"], pri)))
    s = make_roq(n_draws=64, seed=9)
    priors = s["priors"].copy()
    roq = ROQGravitationalWaveTransient(s["interferometers"], s["waveform_generator"], priors,
                                        weights=s["weights"], time_marginalization=True,
                                        phase_marginalization=True, delta_tc=s["delta_tc"])
    roq_rows = roq.rows_from_parameters(fixed_marginalised(s["draws"], priors))
    server = LikelihoodServer(hbm_budget=50e9, max_batch_rows=256)
    try:
        server.register_multi(names, lks)
        server.register("roq", roq)
        futs = []
        for rep in range(4):
            for name, pool in zip(names, pools):
                futs.append((name, pool[rep::4], server.submit(name, pool[rep::4])))
            futs.append(("roq", roq_rows[rep::4], server.submit("roq", roq_rows[rep::4])))
        direct = dict(zip(names, lks))
        direct["roq"] = roq
        for name, rows, fut in futs:
            ref = direct[name].log_likelihood_ratio_batch(rows)
            got = fut.result()
            assert np.max(np.abs(got - ref)) < 1e-8, name
        b, r = server.stats()
        say("batches", b, "rows", r)
        assert r == sum(len(x[1]) for x in futs)
    finally:
        server.close()
