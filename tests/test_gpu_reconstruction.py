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


def test_relative_binning_reconstructs_distance_and_phase_from_its_inner_products():
    """The relative-binning likelihood reconstructs distance and phase from
    its own <d|h>, <h|h> (relative.py calculate_snrs), compared with the
    relative-binning oracle's; its time reconstruction is not provided."""
    from bilby_amd.configs import make_taylorf2_marginalised
    from bilby_amd.relative import RelativeBinningGravitationalWaveTransient
    from oracle.likelihood import OracleDetector, OracleEvent
    from oracle.relative import OracleRelativeBinning
    say("relative binning reconstruction")
    s = make_taylorf2_marginalised(n_draws=8, seed=5, duration=4.0, sampling_frequency=2048.0,
                                   detectors=("H1", "L1", "V1"), time=False)
    ifos = s["interferometers"]
    inj = dict(s["injection"])
    lk = RelativeBinningGravitationalWaveTransient(
        ifos, s["waveform_generator"], fiducial_parameters=inj, distance_marginalization=True,
        phase_marginalization=True, priors=s["priors"].copy(), epsilon=0.1, precision="fp64")
    dets = [OracleDetector(i.name, i.frequency_array, i.frequency_domain_strain,
                           i.power_spectral_density_array, i.frequency_mask) for i in ifos]
    ev = OracleEvent(dets, ifos.duration, ifos.sampling_frequency, ifos.start_time)
    olk = OracleRelativeBinning(ev, "taylorf2", inj,
                                waveform_arguments=dict(s["waveform_generator"].waveform_arguments),
                                epsilon=0.1)
    q = dict(inj, luminosity_distance=lk._ref_dist, phase=0.0)
    dh, hh = 0j, 0.0
    for det in ev.detectors:
        ratio = olk.edge_response(det, q) / olk.fiducial_points[det.name]
        a, b = olk.snrs_from_ratio(det, ratio)
        dh, hh = dh + a, hh + b
    g_dh, g_hh = lk._plain_inner_products(q)
    assert abs(g_dh - dh) < 1e-9 * abs(dh) and abs(g_hh - hh) < 1e-9 * hh
    x, post = lk._distance_posterior(q)
    assert np.isfinite(post).all() and post.max() > 0
    out = lk.generate_posterior_sample_from_marginalized_likelihood(q, rng=np.random.default_rng(1))
    assert out["luminosity_distance"] != q["luminosity_distance"] and out["phase"] != 0.0
