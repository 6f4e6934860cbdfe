"""GPU checks of the multi-event persistent launch (uneven batches): agrees
with the per-event path and the FP64 oracle, and a captured hipGraph replays
correctly with new rows between replays."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def _events():
    from bilby_amd.configs import make_uneven_events
    return make_uneven_events(n_events=5, seed=5, rows_range=(9, 70), durations=(4, 16, 32),
                              chi_range=0.8)


def _likelihoods(events):
    from bilby_amd.likelihood import GravitationalWaveTransient
    lks = []
    for ev in events:
        pri = ev["priors"].copy()
        lks.append(GravitationalWaveTransient(ev["interferometers"], ev["waveform_generator"],
                                              time_marginalization=True,
                                              distance_marginalization=True,
                                              phase_marginalization=True, priors=pri,
                                              precision="fp32"))
        ev["fixed_priors"] = pri
    return lks


def test_multi_event_matches_per_event_and_oracle():
    """Each event's rows through the persistent multi-event launch equal the
    per-event batched path to 1e-8 (only float64 summation order differs)
    and the FP64 oracle to 1e-3 (spot check of 3 rows per event)."""
    from bilby_amd.multi import MultiEventLikelihood
    from tests.helpers import draws_as_dicts, fixed_marginalised, oracle_from_setup
    say("multi-event vs per-event")
    events = _events()
    lks = _likelihoods(events)
    multi = MultiEventLikelihood(lks, [ev["n_rows"] for ev in events])
    single = []
    for e, (ev, lk) in enumerate(zip(events, lks)):
        d = fixed_marginalised(ev["draws"], ev["fixed_priors"])
        multi.set_parameters(e, d)
        single.append(lk.log_likelihood_ratio_batch(d))
    outs = [o.cpu().numpy() for o in multi.run()]
    for e in range(len(events)):
        assert np.max(np.abs(outs[e] - single[e])) < 1e-8, e
    for e, ev in enumerate(events[:3]):
        olk, _ = oracle_from_setup(ev)
        d = fixed_marginalised(ev["draws"], ev["fixed_priors"])
        ref = np.array([olk.log_likelihood_ratio(p) for p in draws_as_dicts(d, range(3))])
        say("event", e, "T", ev["interferometers"].duration, "max |dlnL|", np.max(np.abs(outs[e][:3] - ref)))
        assert np.max(np.abs(outs[e][:3] - ref)) < 1e-3


def test_multi_event_graph_replays():
    """A captured graph replayed three times, with new rows and row counts
    set between replays (and eager work and synchronisations in between),
    matches the eager per-event results every time."""
    import torch
    from bilby_amd.multi import MultiEventLikelihood
    from tests.helpers import fixed_marginalised
    say("multi-event graph replay")
    events = _events()
    lks = _likelihoods(events)
    multi = MultiEventLikelihood(lks, [ev["n_rows"] for ev in events])
    full = [fixed_marginalised(ev["draws"], ev["fixed_priors"]) for ev in events]
    for e in range(len(events)):
        multi.set_parameters(e, full[e])
    multi.capture()
    rng = np.random.default_rng(0)
    for rep in range(3):
        sub = []
        for e, ev in enumerate(events):
            n = int(rng.integers(1, ev["n_rows"] + 1))
            idx = rng.permutation(ev["n_rows"])[:n]
            d = {k: v[idx] for k, v in full[e].items()}
            multi.set_parameters(e, d)
            sub.append(d)
        outs = [o.cpu().numpy().copy() for o in multi.run()]
        eager = [lk.log_likelihood_ratio_batch(d) for lk, d in zip(lks, sub)]
        torch.cuda.synchronize()
        for e in range(len(events)):
            assert np.max(np.abs(outs[e] - eager[e])) < 1e-8, (rep, e)
