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
        # a geocent_time away from the fixed value: the window it needs wraps past the
        # segment end; the direct window transform (gwl_window_time_series) gives it
        far = dict(p, geocent_time=inj["geocent_time"])
        t_g, post_g = lk._time_posterior(far)
        t_o, post_o = olk.time_posterior(far)
        say("direct window", len(t_g), "max |dpost|", float(np.max(np.abs(post_g - post_o))))
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
    relative-binning oracle's."""
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


def _replay_distance_phase(lk, ref, inner_products, rng):
    """base.py:642-742 on the oracle's inner products: distance (Eq. C31) then
    phase, each from the inner products at the values drawn so far."""
    from oracle import core
    if lk.distance_marginalization:
        dh, hh = inner_products(ref)
        ratio = ref["luminosity_distance"] / lk._distance_array
        d, h = dh * ratio, hh * ratio ** 2
        ll = core.ln_i0(np.abs(d)) - h / 2 if lk.phase_marginalization else d.real - h / 2
        post = np.exp(ll - np.max(ll)) * lk.distance_prior_array
        ref["luminosity_distance"] = interped(lk._distance_array, post, rng)
    if lk.phase_marginalization:
        dh, hh = inner_products(ref)
        phases = np.linspace(0, 2 * np.pi, 101)
        lp = (dh * np.exp(-2j * phases)).real - hh / 2
        ref["phase"] = interped(phases, np.exp(lp - np.max(lp)), rng)
    return ref


def interped(x, y, rng):
    from bilby_amd.likelihood import interped_sample
    return interped_sample(x, y, rng)


@pytest.mark.parametrize("phase", [False, True])
def test_roq_reconstruction_matches_oracle(phase):
    """ROQ (H1+L1, 400/120 nodes, 1000 ROQ time samples): the time posterior
    on the likelihood's own grid shifted by the jitter (roq.py:1238-1281)
    and the phase posterior from the inner products at the drawn time,
    against the ROQ oracle; the whole chain with the same draws."""
    from bilby_amd.configs import make_roq
    from bilby_amd.roq import ROQGravitationalWaveTransient
    from oracle.roq import OracleROQ
    from tests.helpers import draws_as_dicts, fixed_marginalised
    say("roq reconstruction, phase", phase)
    s = make_roq(n_draws=4, seed=78)
    w = s["weights"]
    o = OracleROQ(["H1", "L1"], s["interferometers"].start_time, w, w["frequency_nodes_linear"],
                  w["frequency_nodes_quadratic"], dict(s["priors"]), time_marginalization=True,
                  phase_marginalization=phase, delta_tc=s["delta_tc"],
                  waveform_arguments=s["waveform_generator"].waveform_arguments)
    priors = s["priors"].copy()
    lk = ROQGravitationalWaveTransient(s["interferometers"], s["waveform_generator"], priors,
                                       weights=w, time_marginalization=True,
                                       phase_marginalization=phase, delta_tc=s["delta_tc"],
                                       precision="fp64")
    p = draws_as_dicts(fixed_marginalised(s["draws"], priors), [1])[0]
    t_g, post_g = lk._time_posterior(p)
    t_o, post_o = o.time_posterior_on_grid(p)
    say("grid", len(t_g), "max |dpost|", float(np.max(np.abs(post_g - post_o))), "peak",
        float(post_o.max()))
    assert np.max(np.abs(t_g - t_o)) < 1e-9
    assert np.max(np.abs(post_g - post_o)) < 1e-9 * np.max(post_o)
    out = lk.generate_posterior_sample_from_marginalized_likelihood(dict(p),
                                                                    rng=np.random.default_rng(3))
    rng = np.random.default_rng(3)
    ref = dict(p)
    ref["geocent_time"] = float(rng.choice(t_o, p=post_o))
    ref = _replay_distance_phase(lk, ref, lambda q: o.inner_products(dict(q, time_jitter=0.0)),
                                 rng)
    say({k: (out[k], ref[k]) for k in ("geocent_time", "phase")})
    assert out["geocent_time"] == ref["geocent_time"]
    assert abs(out["phase"] - ref["phase"]) < 1e-6
    if phase:
        assert out["phase"] != p["phase"]


def test_roq_reconstructs_distance_from_its_inner_products():
    """ROQ with distance and phase marginalised: the inner products at the
    sample's time (roq.py:467-549) against the oracle's, and the distance
    and phase draws of the chain."""
    from bilby_amd.configs import make_roq
    from bilby_amd.roq import ROQGravitationalWaveTransient
    from oracle.roq import OracleROQ
    from tests.helpers import draws_as_dicts, fixed_marginalised
    say("roq reconstruction, distance + phase")
    s = make_roq(n_draws=4, seed=79)
    w = s["weights"]
    o = OracleROQ(["H1", "L1"], s["interferometers"].start_time, w, w["frequency_nodes_linear"],
                  w["frequency_nodes_quadratic"], dict(s["priors"]), time_marginalization=False,
                  phase_marginalization=True,
                  waveform_arguments=s["waveform_generator"].waveform_arguments)
    priors = s["priors"].copy()
    lk = ROQGravitationalWaveTransient(s["interferometers"], s["waveform_generator"], priors,
                                       weights=w, distance_marginalization=True,
                                       phase_marginalization=True, precision="fp64")
    p = draws_as_dicts(fixed_marginalised(s["draws"], priors), [2])[0]
    dh_g, hh_g = lk._plain_inner_products(p)
    dh_o, hh_o = o.inner_products(p)
    say("dh", dh_g, dh_o, "hh", hh_g, hh_o)
    assert abs(dh_g - dh_o) < 1e-9 * max(1.0, abs(dh_o)) and abs(hh_g - hh_o) < 1e-9 * hh_o
    out = lk.generate_posterior_sample_from_marginalized_likelihood(dict(p),
                                                                    rng=np.random.default_rng(4))
    ref = _replay_distance_phase(lk, dict(p), o.inner_products, np.random.default_rng(4))
    say({k: (out[k], ref[k]) for k in ("luminosity_distance", "phase")})
    assert abs(out["luminosity_distance"] - ref["luminosity_distance"]) < 1e-6 * ref[
        "luminosity_distance"]
    assert abs(out["phase"] - ref["phase"]) < 1e-6


@pytest.mark.parametrize("distance", [False, True])
def test_multiband_reconstruction_matches_oracle(distance):
    """Multibanding (16 s light binary, H1+L1+V1), time + phase (+ distance)
    marginalised: the time posterior on the likelihood's own grid from
    d_inner_h_array of the waveform at the sample's time (multiband.py:
    813-848), against the multibanding oracle over the whole grid (zero
    outside the window); then the chain with the same draws."""
    from bilby_amd.multiband import MBGravitationalWaveTransient
    from oracle.multiband import OracleMultiband
    from tests.test_gpu_binned_time import _mb_event, _oracle_event
    say("multiband reconstruction, distance", distance)
    ifos, wfg, inj, priors = _mb_event()
    mc = (inj["mass_1"] * inj["mass_2"]) ** 0.6 / (inj["mass_1"] + inj["mass_2"]) ** 0.2
    lk = MBGravitationalWaveTransient(ifos, wfg, reference_chirp_mass=mc, priors=priors.copy(),
                                      time_marginalization=True,
                                      distance_marginalization=distance,
                                      phase_marginalization=True, precision="fp64")
    olk = OracleMultiband(_oracle_event(ifos), "taylorf2", lk.setup,
                          waveform_arguments=dict(wfg.waveform_arguments),
                          time_marginalization=True, distance_marginalization=distance,
                          phase_marginalization=True, priors=dict(priors),
                          distance_table=getattr(lk, "_dist_margd_loglikelihood_array", None))
    assert lk.accuracy_factor == 5 and lk.linear_interpolation and lk.highest_mode == 2
    assert lk.reference_chirp_mass == lk.setup.reference_chirp_mass
    p = dict(inj, geocent_time=float(ifos.start_time), phase=0.0, time_jitter=3.1e-5)
    if distance:
        p["luminosity_distance"] = lk._ref_dist
    t_g, post_g = lk._time_posterior(p)
    t_o, post_o = olk.time_posterior_on_grid(p)
    N, j_first, L = lk._time_window
    outside = np.ones(len(t_o), dtype=bool)
    outside[j_first:j_first + L] = False
    say("window", L, "of", N, "max |dpost|",
        float(np.max(np.abs(post_g - post_o[j_first:j_first + L]))), "peak", float(post_o.max()))
    assert np.all(post_o[outside] == 0.0)
    assert np.max(np.abs(t_g - t_o[j_first:j_first + L])) < 1e-9
    assert np.max(np.abs(post_g - post_o[j_first:j_first + L])) < 1e-8 * np.max(post_o)
    out = lk.generate_posterior_sample_from_marginalized_likelihood(dict(p),
                                                                    rng=np.random.default_rng(5))
    rng = np.random.default_rng(5)
    ref = dict(p)
    ref["geocent_time"] = float(rng.choice(t_o, p=post_o))
    ref = _replay_distance_phase(lk, ref, lambda q: olk.inner_products(dict(q, time_jitter=0.0)),
                                 rng)
    say({k: (out[k], ref[k]) for k in ("geocent_time", "luminosity_distance", "phase")})
    assert out["geocent_time"] == ref["geocent_time"]
    assert abs(out["luminosity_distance"] - ref["luminosity_distance"]) < 1e-6 * ref[
        "luminosity_distance"]
    assert abs(out["phase"] - ref["phase"]) < 1e-6
    assert abs(out["geocent_time"] - inj["geocent_time"]) < 2e-3


@pytest.mark.parametrize("distance", [False, True])
def test_relative_binning_time_reconstruction_matches_oracle(distance):
    """Relative binning (4 s @ 2048 Hz, H1+L1+V1, epsilon 0.1), time + phase
    (+ distance) marginalised: the time posterior on the 1/16384 s grid from
    the rebuilt waveform (base.py:580-640 with relative.py:380-393), against
    the relative-binning oracle's, for a draw off the fiducial point and
    with jitter; the whole chain with the same draws.  Also the transform on
    the likelihood's own window (gwl_relbin_time_series) combined as
    base.py:794-820 gives the evaluation's time-marginalised lnL."""
    import ctypes
    import torch
    from bilby_amd import _lib
    from bilby_amd.configs import make_taylorf2_marginalised
    from bilby_amd.likelihood import ln_i0_host
    from bilby_amd.relative import RelativeBinningGravitationalWaveTransient
    from oracle.relative import OracleRelativeBinning
    from tests.test_gpu_binned_time import _oracle_event
    say("relative binning time reconstruction, distance", distance)
    s = make_taylorf2_marginalised(n_draws=8, seed=5, duration=4.0, sampling_frequency=2048.0,
                                   detectors=("H1", "L1", "V1"))
    ifos = s["interferometers"]
    inj = dict(s["injection"])
    eps = 0.1
    lk = RelativeBinningGravitationalWaveTransient(
        ifos, s["waveform_generator"], fiducial_parameters=inj, time_marginalization=True,
        distance_marginalization=distance, phase_marginalization=True,
        priors=s["priors"].copy(), epsilon=eps, precision="fp64")
    olk = OracleRelativeBinning(
        _oracle_event(ifos), "taylorf2", inj,
        waveform_arguments=dict(s["waveform_generator"].waveform_arguments),
        time_marginalization=True, distance_marginalization=distance, phase_marginalization=True,
        priors=dict(s["priors"]), epsilon=eps,
        distance_table=getattr(lk, "_dist_margd_loglikelihood_array", None))
    p = dict(inj, geocent_time=float(ifos.start_time), phase=0.0, time_jitter=2.3e-4,
             mass_1=inj["mass_1"] * 1.001)
    if distance:
        p["luminosity_distance"] = lk._ref_dist
    t_g, post_g = lk._time_posterior(p)
    t_o, post_o = olk.time_posterior(p)
    say("grid", len(t_g), "max |dpost|", float(np.max(np.abs(post_g - post_o))), "peak",
        float(post_o.max()))
    assert len(t_g) == len(t_o) and np.max(np.abs(t_g - t_o)) < 1e-9
    assert np.max(np.abs(post_g - post_o)) < 1e-6 * np.max(post_o)
    out = lk.generate_posterior_sample_from_marginalized_likelihood(dict(p),
                                                                    rng=np.random.default_rng(6))
    rng = np.random.default_rng(6)
    ref = dict(p)
    keep = post_o > np.max(post_o) / 1000
    if np.sum(keep) < 3:
        keep[1:-1] = keep[1:-1] | keep[2:] | keep[:-2]
    ref["geocent_time"] = interped(t_o[keep], post_o[keep], rng)

    def binned_inner(q):
        q = dict(q, time_jitter=0.0)
        dh, hh = 0j, 0.0
        for det in olk.event.detectors:
            a, b = olk.snrs_from_ratio(det, olk.edge_response(det, q) / olk.fiducial_points[det.name])
            dh, hh = dh + a, hh + b
        return dh, hh
    ref = _replay_distance_phase(lk, ref, binned_inner, rng)
    say({k: (out[k], ref[k]) for k in ("geocent_time", "luminosity_distance", "phase")})
    assert abs(out["geocent_time"] - ref["geocent_time"]) < 1e-7
    assert abs(out["luminosity_distance"] - ref["luminosity_distance"]) < 1e-6 * ref[
        "luminosity_distance"]
    assert abs(out["phase"] - ref["phase"]) < 1e-6
    assert abs(out["geocent_time"] - inj["geocent_time"]) < 5e-3
    if distance:
        return
    # the transform on the likelihood's own window, combined as base.py:794-820
    rows = torch.from_numpy(lk.rows_from_parameters(dict(p))[:1]).to("cuda:0")
    ev = lk._event
    j_first, n_window, t_step = ctypes.c_int64(), ctypes.c_int32(), ctypes.c_double()
    _lib.check(_lib.LIB.gwl_event_time_window(ev, ctypes.byref(j_first), ctypes.byref(n_window),
                                              ctypes.byref(t_step)), "gwl_event_time_window")
    L = n_window.value
    dh = torch.empty((1, L, 2), dtype=torch.float64, device="cuda:0")
    hh = torch.empty(1, dtype=torch.float64, device="cuda:0")
    _lib.check(_lib.LIB.gwl_relbin_time_series(
        lk._relbin, ctypes.c_void_p(rows.data_ptr()), 1, _lib.NPARAM, _lib.FP64,
        ctypes.c_void_p(dh.data_ptr()), ctypes.c_void_p(hh.data_ptr()),
        ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)), "gwl_relbin_time_series")
    dh = dh.cpu().numpy()[0]
    x = dh[:, 0] + 1j * dh[:, 1]
    times = (ifos.start_time + (j_first.value + np.arange(L) + 1) * t_step.value
             + p["time_jitter"])
    prior = s["priors"]["geocent_time"]
    inside = (times >= prior.minimum) & (times <= prior.maximum)
    ll = ln_i0_host(np.abs(x[inside])) - float(hh.cpu().numpy()[0]) / 2
    m = ll.max()
    lnl = m + np.log(np.sum(np.exp(ll - m)) * lk._delta_tc / (prior.maximum - prior.minimum))
    ev_lnl = lk.log_likelihood_ratio_batch({k: np.array([v]) for k, v in p.items()})[0]
    say("own-window series lnL", lnl, "evaluation", ev_lnl)
    assert abs(lnl - ev_lnl) < 1e-8 * max(1.0, abs(ev_lnl))


@pytest.mark.parametrize("marginalised", [False, True])
def test_per_detector_log_likelihoods(marginalised):
    """compute_per_detector_log_likelihood (base.py:479-500): each
    detector's own lnL ratio against the oracle on that detector alone;
    without marginalisation they add up to the network lnL (the inner
    products are sums over detectors)."""
    from bilby_amd.configs import make_taylorf2_marginalised
    from bilby_amd.likelihood import GravitationalWaveTransient
    from oracle.likelihood import OracleDetector, OracleEvent, OracleGWTransient
    from tests.helpers import fixed_marginalised
    say("per-detector lnL, marginalised", marginalised)
    s = make_taylorf2_marginalised(n_draws=4, seed=23, duration=4.0, sampling_frequency=1024.0,
                                   detectors=("H1", "L1", "V1"), time=marginalised,
                                   distance=marginalised, phase=marginalised)
    opri = dict(s["priors"])
    priors = s["priors"].copy()
    flags = dict(time_marginalization=marginalised, distance_marginalization=marginalised,
                 phase_marginalization=marginalised)
    lk = GravitationalWaveTransient(s["interferometers"], s["waveform_generator"], priors=priors,
                                    precision="fp64", **flags)
    inj = dict(s["injection"])
    p = dict(inj)
    fixed = fixed_marginalised({k: np.array([v]) for k, v in inj.items()}, priors)
    p.update({k: float(v[0]) for k, v in fixed.items()})
    if marginalised:
        p["time_jitter"] = 1.1e-4
    out = lk.compute_per_detector_log_likelihood(p)
    ifos = s["interferometers"]
    for ifo in ifos:
        det = OracleDetector(ifo.name, ifo.frequency_array, ifo.frequency_domain_strain,
                             ifo.power_spectral_density_array, ifo.frequency_mask)
        ev = OracleEvent([det], ifos.duration, ifos.sampling_frequency, ifos.start_time)
        o = OracleGWTransient(ev, "taylorf2",
                              waveform_arguments=dict(s["waveform_generator"].waveform_arguments),
                              priors=dict(opri), **flags)
        ref = o.log_likelihood_ratio(dict(p))
        got = out["%s_log_likelihood" % ifo.name]
        say(ifo.name, got, ref)
        assert abs(got - ref) < 1e-7 * max(1.0, abs(ref))
        assert p["%s_log_likelihood" % ifo.name] == got
    if not marginalised:
        total = sum(out["%s_log_likelihood" % i.name] for i in ifos)
        assert abs(total - lk.log_likelihood_ratio(dict(inj))) < 1e-9 * abs(total)
    assert lk.lal_version == "N/A"


def test_reconstruction_at_the_headline_shape():
    """The reconstruction chain (time on the 1/16384 s grid: transform
    length 131,072; distance; phase) on the 8 s @ 4096 Hz H1+L1+V1 event
    with time, distance and phase marginalised, against the oracle with the
    same draws; and the direct window transform at the injected time."""
    from bilby_amd.configs import make_taylorf2_marginalised
    from bilby_amd.likelihood import GravitationalWaveTransient, interped_sample
    from tests.helpers import fixed_marginalised, oracle_from_setup
    say("reconstruction, 8 s @ 4096 Hz, H1+L1+V1")
    s = make_taylorf2_marginalised(n_draws=4, seed=27, duration=8.0, sampling_frequency=4096.0)
    priors = s["priors"].copy()
    lk = GravitationalWaveTransient(s["interferometers"], s["waveform_generator"],
                                    time_marginalization=True, distance_marginalization=True,
                                    phase_marginalization=True, priors=priors, precision="fp64")
    olk, _ = oracle_from_setup(s)
    inj = dict(s["injection"])
    p = dict(inj)
    fixed = fixed_marginalised({k: np.array([v]) for k, v in inj.items()}, priors)
    p.update({k: float(v[0]) for k, v in fixed.items()})
    p["time_jitter"] = -2.2e-4
    t_g, post_g = lk._time_posterior(p)
    t_o, post_o = olk.time_posterior(p)
    say("grid", len(t_g), "max |dpost|", float(np.max(np.abs(post_g - post_o))))
    assert len(t_g) == len(t_o) and np.max(np.abs(t_g - t_o)) < 1e-9
    assert np.max(np.abs(post_g - post_o)) < 1e-6 * np.max(post_o)
    far = dict(p, geocent_time=inj["geocent_time"] + 0.03)
    t_g, post_g = lk._time_posterior(far)
    t_o, post_o = olk.time_posterior(far)
    assert np.max(np.abs(post_g - post_o)) < 1e-6 * np.max(post_o)
    out = lk.generate_posterior_sample_from_marginalized_likelihood(dict(p),
                                                                    rng=np.random.default_rng(9))
    rng = np.random.default_rng(9)
    ref = dict(p)
    t_o, post_o = olk.time_posterior(ref)
    keep = post_o > np.max(post_o) / 1000
    if np.sum(keep) < 3:
        keep[1:-1] = keep[1:-1] | keep[2:] | keep[:-2]
    ref["geocent_time"] = interped_sample(t_o[keep], post_o[keep], rng)
    ref["luminosity_distance"] = interped_sample(*olk.distance_posterior(ref), rng)
    ref["phase"] = interped_sample(*olk.phase_posterior(ref), rng)
    say({k: (out[k], ref[k]) for k in ("geocent_time", "luminosity_distance", "phase")})
    assert abs(out["geocent_time"] - ref["geocent_time"]) < 1e-7
    assert abs(out["luminosity_distance"] - ref["luminosity_distance"]) < 1e-6 * ref[
        "luminosity_distance"]
    assert abs(out["phase"] - ref["phase"]) < 1e-6
    assert abs(out["geocent_time"] - inj["geocent_time"]) < 2e-3


def test_sine_gaussian_time_reconstruction():
    """Time reconstruction for a sine-Gaussian event (4 s @ 2048 Hz, H1+L1,
    time marginalised): the upsampled companion event on the tiled
    transform, and the direct window transform for a time away from the
    fixed value, against the oracle; the draw with the same random numbers."""
    from bilby_amd.configs import make_sine_gaussian_plain
    from bilby_amd.likelihood import GravitationalWaveTransient, interped_sample
    from bilby_amd.prior import PriorDict, Uniform
    from tests.helpers import oracle_from_setup
    say("sine-Gaussian time reconstruction")
    s = make_sine_gaussian_plain(n_draws=4, seed=41)
    inj = dict(s["injection"])
    tc = inj["geocent_time"]
    s["priors"] = PriorDict(geocent_time=Uniform(tc - 0.1, tc + 0.1))
    s["time_marginalization"] = True
    olk, _ = oracle_from_setup(s)
    priors = s["priors"].copy()
    lk = GravitationalWaveTransient(s["interferometers"], s["waveform_generator"],
                                    time_marginalization=True, priors=priors, precision="fp64")
    p = dict(inj, geocent_time=float(s["interferometers"].start_time), time_jitter=3.3e-4)
    for q in (p, dict(p, geocent_time=tc - 0.05)):
        t_g, post_g = lk._time_posterior(q)
        t_o, post_o = olk.time_posterior(q)
        say("grid", len(t_g), "max |dpost|", float(np.max(np.abs(post_g - post_o))), "peak",
            float(post_o.max()))
        assert len(t_g) == len(t_o) and np.max(np.abs(t_g - t_o)) < 1e-9
        assert np.max(np.abs(post_g - post_o)) < 1e-6 * np.max(post_o)
    out = lk.generate_posterior_sample_from_marginalized_likelihood(dict(p),
                                                                    rng=np.random.default_rng(2))
    t_o, post_o = olk.time_posterior(p)
    keep = post_o > np.max(post_o) / 1000
    if np.sum(keep) < 3:
        keep[1:-1] = keep[1:-1] | keep[2:] | keep[:-2]
    ref = interped_sample(t_o[keep], post_o[keep], np.random.default_rng(2))
    assert abs(out["geocent_time"] - ref) < 1e-7 and abs(ref - tc) < 2e-3
