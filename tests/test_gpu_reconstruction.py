"""Posterior reconstruction of the marginalised parameters
(generate_posterior_sample_from_marginalized_likelihood, base.py:502-742) on
the GPU path against the FP64 oracle's restatement: the time posterior on
the 1/16384 s grid (the upsampled event's transform), the distance and
phase posteriors (inner products of the companion plain likelihood), and a
whole reconstruction chain with the same random draws; the reference's own
check (marginalization_test.py:455-475) that every reconstructed value
moves off its fixed value."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [dict(time=True, distance=False, phase=False), dict(time=True, distance=False, phase=True),
         dict(time=True, distance=True, phase=True), dict(time=False, distance=True, phase=False),
         dict(time=False, distance=False, phase=True)]


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("case", CASES)
def test_reconstruction_matches_oracle(case):
    from bilby_amd.configs import make_taylorf2_marginalised
    from bilby_amd.likelihood import GravitationalWaveTransient, interped_sample
    from tests.helpers import draws_as_dicts, fixed_marginalised, oracle_from_setup
    say("reconstruction", case)
    s = make_taylorf2_marginalised(n_draws=8, seed=21, duration=4.0, sampling_frequency=1024.0,
                                   detectors=("H1", "L1"), **case)
    priors = s["priors"].copy()
    lk = GravitationalWaveTransient(s["interferometers"], s["waveform_generator"],
                                    time_marginalization=case["time"],
                                    distance_marginalization=case["distance"],
                                    phase_marginalization=case["phase"], priors=priors,
                                    precision="fp64")
    olk, _ = oracle_from_setup(s)
    inj = dict(s["injection"])
    p = dict(inj)
    fixed = fixed_marginalised({k: np.array([v]) for k, v in inj.items()}, priors)
    p.update({k: float(v[0]) for k, v in fixed.items()})
    if case["time"]:
        p["time_jitter"] = 1.7e-4
        t_g, post_g = lk._time_posterior(p)
        t_o, post_o = olk.time_posterior(p)
        say("time grid", len(t_g), "max |dpost|", float(np.max(np.abs(post_g - post_o))))
        assert len(t_g) == len(t_o) and np.max(np.abs(t_g - t_o)) < 1e-9
        assert np.max(np.abs(post_g - post_o)) < 1e-6 * np.max(post_o)
    q = dict(p, geocent_time=inj["geocent_time"])
    if case["distance"]:
        d_g, post_g = lk._distance_posterior(q)
        d_o, post_o = olk.distance_posterior(q)
        assert np.array_equal(d_g, d_o)
        assert np.max(np.abs(post_g - post_o)) < 1e-8 * np.max(post_o)
    if case["phase"]:
        f_g, post_g = lk._phase_posterior(q)
        f_o, post_o = olk.phase_posterior(q)
        assert np.max(np.abs(post_g - post_o)) < 1e-8
    # the whole chain with the same uniform draws, replayed on the oracle's arrays
    out = lk.generate_posterior_sample_from_marginalized_likelihood(dict(p),
                                                                    rng=np.random.default_rng(3))
    rng = np.random.default_rng(3)
    ref = dict(p)
    if case["time"]:
        t_o, post_o = olk.time_posterior(ref)
        keep = post_o > np.max(post_o) / 1000
        if np.sum(keep) < 3:
            keep[1:-1] = keep[1:-1] | keep[2:] | keep[:-2]
        ref["geocent_time"] = interped_sample(t_o[keep], post_o[keep], rng)
    if case["distance"]:
        ref["luminosity_distance"] = interped_sample(*olk.distance_posterior(ref), rng)
    if case["phase"]:
        ref["phase"] = interped_sample(*olk.phase_posterior(ref), rng)
    say({k: (out[k], ref[k]) for k in ("geocent_time", "luminosity_distance", "phase")})
    assert abs(out["geocent_time"] - ref["geocent_time"]) < 1e-7
    assert abs(out["luminosity_distance"] - ref["luminosity_distance"]) < 1e-6 * ref[
        "luminosity_distance"]
    assert abs(out["phase"] - ref["phase"]) < 1e-6
    for key, on in (("geocent_time", case["time"]), ("luminosity_distance", case["distance"]),
                    ("phase", case["phase"])):
        if on:
            assert out[key] != p[key]
    pr = s["priors"]
    if case["time"]:
        assert pr["geocent_time"].minimum <= out["geocent_time"] <= pr["geocent_time"].maximum
