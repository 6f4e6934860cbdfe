"""GPU checks of calibration marginalisation (gwl_calib: prep kernel, MFMA
contraction with the curves, logsumexp epilogue) against the FP64 oracle
with the same response curves: float64 contraction < 1e-7, float32 and
split-bfloat16 < 1e-3, for plain, phase and distance + phase
marginalisation; per-curve values; invalid rows."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def _setup(tmp_path, n_curves=100):
    from bilby_amd import calibration as cal
    from bilby_amd.configs import make_taylorf2_marginalised
    s = make_taylorf2_marginalised(n_draws=32, seed=8, duration=4.0, sampling_frequency=1024.0,
                                   detectors=("H1", "L1"))
    priors = s["priors"]
    for ifo in s["interferometers"]:
        ifo.calibration_model = cal.CubicSpline("recalib_%s_" % ifo.name, 20.0, 512.0, 5)
        priors.update(cal.constant_uncertainty_spline_priors(0.1, 0.1, 20.0, 512.0, 5, ifo.name))
    files = {ifo.name: str(tmp_path / ("%s_%d.npz" % (ifo.name, n_curves)))
             for ifo in s["interferometers"]}
    return s, files


@pytest.mark.parametrize("flags", [dict(distance=False, phase=False),
                                   dict(distance=False, phase=True),
                                   dict(distance=True, phase=True)])
def test_calibration_marginalised_matches_oracle(flags, tmp_path):
    from bilby_amd.likelihood import GravitationalWaveTransient
    from oracle.calibration import OracleCalibrated
    from oracle.likelihood import OracleDetector, OracleEvent
    from tests.helpers import draws_as_dicts, fixed_marginalised
    say("calibration", flags)
    s, files = _setup(tmp_path)
    priors = s["priors"].copy()
    lk = GravitationalWaveTransient(
        s["interferometers"], s["waveform_generator"], calibration_marginalization=True,
        distance_marginalization=flags["distance"], phase_marginalization=flags["phase"],
        priors=priors, number_of_response_curves=100, calibration_lookup_table=files,
        precision="fp64")
    assert "recalib_index" in lk.marginalized_parameters
    # base.py:1170, 234-236: meta_data and repr report the calibration flag
    assert lk.meta_data["calibration_marginalization"] is True
    assert "calibration_marginalization=True" in repr(lk)
    assert all(float(priors[k].peak) == 0.0 for k in priors if "recalib" in k and "frequency" not in k)
    ifos = s["interferometers"]
    dets = [OracleDetector(i.name, i.frequency_array, i.frequency_domain_strain,
                           i.power_spectral_density_array, i.frequency_mask) for i in ifos]
    ev = OracleEvent(dets, ifos.duration, ifos.sampling_frequency, ifos.start_time)
    opri = dict(s["priors"])
    olk = OracleCalibrated(ev, "taylorf2", lk.calibration_draws,
                           waveform_arguments=dict(s["waveform_generator"].waveform_arguments),
                           distance_marginalization=flags["distance"],
                           phase_marginalization=flags["phase"], priors=opri,
                           distance_table=getattr(lk, "_dist_margd_loglikelihood_array", None))
    draws = fixed_marginalised(s["draws"], priors)
    n = 24
    ref = np.array([olk.log_likelihood_ratio(p) for p in draws_as_dicts(draws, range(n))])
    got = {m: lk.log_likelihood_ratio_batch({k: v[:n] for k, v in draws.items()}, precision=m)
           for m in ("fp64", "fp32", "bf16x3")}
    say({m: float(np.max(np.abs(v - ref))) for m, v in got.items()}, "lnL range", ref.min(),
        ref.max())
    assert np.all(np.isfinite(ref))
    assert np.max(np.abs(got["fp64"] - ref)) < 1e-7
    assert np.max(np.abs(got["fp32"] - ref)) < 1e-3
    assert np.max(np.abs(got["bf16x3"] - ref)) < 1e-3
    p0 = draws_as_dicts(draws, [3])[0]
    per = lk.get_calibration_log_likelihoods(p0)
    assert np.max(np.abs(per - olk.per_curve_log_likelihoods(p0))) < 1e-7
    assert 0 <= lk.generate_calibration_sample_from_marginalized_likelihood(p0) < 100
    bad = {k: v[:n].copy() for k, v in draws.items()}
    bad["mass_1"][5] = -1.0
    out = lk.log_likelihood_ratio_batch(bad)
    assert out[5] == np.nan_to_num(-np.inf)
    keep = [i for i in range(n) if i != 5]
    assert np.max(np.abs(out[keep] - got["fp64"][keep])) < 1e-12


@pytest.mark.parametrize("case", [
    dict(distance=False, phase=False, T=4.0, fs=1024.0, dets=("H1", "L1"), curves=100),
    dict(distance=False, phase=True, T=4.0, fs=1024.0, dets=("H1", "L1"), curves=100),
    dict(distance=True, phase=True, T=4.0, fs=1024.0, dets=("H1", "L1"), curves=100),
    dict(distance=True, phase=True, T=8.0, fs=4096.0, dets=("H1", "L1", "V1"), curves=21),
])
def test_calibration_and_time_marginalised_matches_oracle(case, tmp_path):
    """Calibration together with time marginalisation (k_cal_time_w: one
    transform per curve, base.py:305-324, 862-877) against the oracle's
    tiled per-curve FFTs: float32 transforms < 1e-3, float64 (k_cal_time_w64,
    curves to 48 bits) within 1e-8 max(1, |lnL|), per-curve values in both,
    jittered rows, invalid rows; the split-bfloat16 mode is refused."""
    from bilby_amd import calibration as cal
    from bilby_amd.configs import make_taylorf2_marginalised
    from bilby_amd.likelihood import GravitationalWaveTransient
    from oracle.calibration import OracleCalibrated
    from oracle.likelihood import OracleDetector, OracleEvent
    from tests.helpers import draws_as_dicts, fixed_marginalised
    say("calibration + time", case)
    s = make_taylorf2_marginalised(n_draws=16, seed=9, duration=case["T"],
                                   sampling_frequency=case["fs"], detectors=case["dets"])
    priors = s["priors"]
    fmax = min(512.0, case["fs"] / 2)
    for ifo in s["interferometers"]:
        ifo.calibration_model = cal.CubicSpline("recalib_%s_" % ifo.name, 20.0, fmax, 5)
        priors.update(cal.constant_uncertainty_spline_priors(0.1, 0.1, 20.0, fmax, 5, ifo.name))
    files = {ifo.name: str(tmp_path / ("%s_t.npz" % ifo.name)) for ifo in s["interferometers"]}
    priors = s["priors"].copy()
    lk = GravitationalWaveTransient(
        s["interferometers"], s["waveform_generator"], calibration_marginalization=True,
        time_marginalization=True, distance_marginalization=case["distance"],
        phase_marginalization=case["phase"], priors=priors,
        number_of_response_curves=case["curves"], calibration_lookup_table=files,
        precision="fp32")
    ifos = s["interferometers"]
    dets = [OracleDetector(i.name, i.frequency_array, i.frequency_domain_strain,
                           i.power_spectral_density_array, i.frequency_mask) for i in ifos]
    ev = OracleEvent(dets, ifos.duration, ifos.sampling_frequency, ifos.start_time)
    olk = OracleCalibrated(ev, "taylorf2", lk.calibration_draws,
                           waveform_arguments=dict(s["waveform_generator"].waveform_arguments),
                           time_marginalization=True, distance_marginalization=case["distance"],
                           phase_marginalization=case["phase"], priors=dict(s["priors"]),
                           distance_table=getattr(lk, "_dist_margd_loglikelihood_array", None))
    n = 8 if case["T"] > 4 else 16
    draws = {k: v[:n] for k, v in fixed_marginalised(s["draws"], priors).items()}
    draws["time_jitter"] = np.random.default_rng(3).uniform(-1, 1, n) / case["fs"]
    ref = np.array([olk.log_likelihood_ratio(p) for p in draws_as_dicts(draws, range(n))])
    got = lk.log_likelihood_ratio_batch(draws)
    got64 = lk.log_likelihood_ratio_batch(draws, precision="fp64")
    scale = np.maximum(1.0, np.abs(ref))
    say("fp32", float(np.max(np.abs(got - ref))), "fp64 relative",
        float(np.max(np.abs(got64 - ref) / scale)), "lnL range", ref.min(), ref.max())
    assert np.all(np.isfinite(ref))
    assert np.max(np.abs(got - ref)) < 1e-3
    assert np.max(np.abs(got64 - ref) / scale) < 1e-8
    p0 = draws_as_dicts(draws, [2])[0]
    per = lk.get_calibration_log_likelihoods(p0)
    assert per.shape == (case["curves"],)
    oper = olk.per_curve_log_likelihoods(p0)
    assert np.max(np.abs(per - oper)) < 1e-3
    lk.calibration_mode = 0   # float64 per-curve values
    per64 = lk.get_calibration_log_likelihoods(p0)
    lk.calibration_mode = 1
    assert np.max(np.abs(per64 - oper) / np.maximum(1.0, np.abs(oper))) < 1e-8
    bad = {k: v.copy() for k, v in draws.items()}
    bad["mass_1"][5] = -1.0
    out = lk.log_likelihood_ratio_batch(bad)
    assert out[5] == np.nan_to_num(-np.inf)
    keep = [i for i in range(n) if i != 5]
    assert np.array_equal(out[keep], got[keep])
    out64 = lk.log_likelihood_ratio_batch(bad, precision="fp64")
    assert out64[5] == np.nan_to_num(-np.inf)
    assert np.array_equal(out64[keep], got64[keep])
    with pytest.raises(NotImplementedError):
        lk.log_likelihood_ratio_batch(draws, precision="bf16x3")


@pytest.mark.parametrize("time", [False, True])
def test_reconstruction_with_calibration_and_other_marginalisations(time, tmp_path):
    """generate_posterior_sample_from_marginalized_likelihood (base.py:502-542)
    with calibration, (time,) distance and phase marginalised: the curve drawn
    from the per-curve likelihoods, the time from the uncalibrated waveform
    (base.py:580-640 leaves the curve out), distance and phase from the
    inner products with the drawn curve applied (base.py:286-294,
    gwl_weighted_inner_products), against the oracle with the same draws."""
    from bilby_amd import calibration as cal
    from bilby_amd.configs import make_taylorf2_marginalised
    from bilby_amd.likelihood import GravitationalWaveTransient, interped_sample
    from oracle.calibration import OracleCalibrated
    from oracle.likelihood import OracleDetector, OracleEvent
    from tests.helpers import draws_as_dicts, fixed_marginalised
    say("reconstruction with calibration, time", time)
    s = make_taylorf2_marginalised(n_draws=4, seed=19, duration=4.0, sampling_frequency=1024.0,
                                   detectors=("H1", "L1"))
    priors = s["priors"]
    for ifo in s["interferometers"]:
        ifo.calibration_model = cal.CubicSpline("recalib_%s_" % ifo.name, 20.0, 512.0, 5)
        priors.update(cal.constant_uncertainty_spline_priors(0.1, 0.1, 20.0, 512.0, 5, ifo.name))
    files = {ifo.name: str(tmp_path / ("%s_r.npz" % ifo.name)) for ifo in s["interferometers"]}
    priors = s["priors"].copy()
    lk = GravitationalWaveTransient(
        s["interferometers"], s["waveform_generator"], calibration_marginalization=True,
        time_marginalization=time, distance_marginalization=True, phase_marginalization=True,
        priors=priors, number_of_response_curves=50, calibration_lookup_table=files,
        precision="fp32" if time else "fp64")
    ifos = s["interferometers"]
    dets = [OracleDetector(i.name, i.frequency_array, i.frequency_domain_strain,
                           i.power_spectral_density_array, i.frequency_mask) for i in ifos]
    ev = OracleEvent(dets, ifos.duration, ifos.sampling_frequency, ifos.start_time)
    olk = OracleCalibrated(ev, "taylorf2", lk.calibration_draws,
                           waveform_arguments=dict(s["waveform_generator"].waveform_arguments),
                           time_marginalization=time, distance_marginalization=True,
                           phase_marginalization=True, priors=dict(s["priors"]),
                           distance_table=lk._dist_margd_loglikelihood_array)
    inj = dict(s["injection"])
    p = dict(inj)
    fixed = fixed_marginalised({k: np.array([v]) for k, v in inj.items()}, priors)
    p.update({k: float(v[0]) for k, v in fixed.items()})
    if time:
        p["time_jitter"] = 4.1e-4
    q = dict(p, recalib_index=7)
    dh_g, hh_g = lk._plain_inner_products(q)
    dh_o, hh_o = olk.plain_inner_products(q)
    say("curve 7 <d|h>", dh_g, dh_o, "<h|h>", hh_g, hh_o)
    assert abs(dh_g - dh_o) < 1e-9 * abs(dh_o) and abs(hh_g - hh_o) < 1e-9 * hh_o
    out = lk.generate_posterior_sample_from_marginalized_likelihood(dict(p),
                                                                    rng=np.random.default_rng(8))
    rng = np.random.default_rng(8)
    ref = dict(p)
    ll = olk.per_curve_log_likelihoods(dict(p))
    post = np.exp(ll - np.max(ll))
    ref["recalib_index"] = int(rng.choice(50, p=post / np.sum(post)))
    if time:
        t_o, post_o = olk.time_posterior(ref)
        keep = post_o > np.max(post_o) / 1000
        if np.sum(keep) < 3:
            keep[1:-1] = keep[1:-1] | keep[2:] | keep[:-2]
        ref["geocent_time"] = interped_sample(t_o[keep], post_o[keep], rng)
    ref["luminosity_distance"] = interped_sample(*olk.distance_posterior(ref), rng)
    ref["phase"] = interped_sample(*olk.phase_posterior(ref), rng)
    keys = ("recalib_index", "geocent_time", "luminosity_distance", "phase")
    say({k: (out[k], ref[k]) for k in keys})
    assert out["recalib_index"] == ref["recalib_index"]
    assert abs(out["geocent_time"] - ref["geocent_time"]) < 1e-7
    assert abs(out["luminosity_distance"] - ref["luminosity_distance"]) < 1e-6 * ref[
        "luminosity_distance"]
    assert abs(out["phase"] - ref["phase"]) < 1e-6


@pytest.mark.parametrize("time", [False, True])
def test_per_detector_log_likelihoods_with_calibration(time, tmp_path):
    """base.py:479-500 under calibration marginalisation: each detector's
    lnL marginalised over its own response curves (calculate_snrs per
    interferometer) equals the calibrated FP64 oracle on that detector alone
    with the same curves (float64 1e-7; with time marginalisation float32
    transforms, 1e-3)."""
    from bilby_amd.likelihood import GravitationalWaveTransient
    from oracle.calibration import OracleCalibrated
    from oracle.likelihood import OracleDetector, OracleEvent
    from tests.helpers import draws_as_dicts, fixed_marginalised
    say("per-detector calibration, time", time)
    s, files = _setup(tmp_path, n_curves=40)
    priors = s["priors"].copy()
    lk = GravitationalWaveTransient(
        s["interferometers"], s["waveform_generator"], calibration_marginalization=True,
        distance_marginalization=True, phase_marginalization=True, time_marginalization=time,
        priors=priors, number_of_response_curves=40, calibration_lookup_table=files,
        precision="fp32" if time else "fp64")
    draws = fixed_marginalised(s["draws"], priors)
    table = getattr(lk, "_dist_margd_loglikelihood_array", None)
    tol = 1e-3 if time else 1e-7
    for p in draws_as_dicts(draws, [0, 7]):
        out = lk.compute_per_detector_log_likelihood(dict(p))
        for ifo in s["interferometers"]:
            det = OracleDetector(ifo.name, ifo.frequency_array, ifo.frequency_domain_strain,
                                 ifo.power_spectral_density_array, ifo.frequency_mask)
            ev = OracleEvent([det], s["interferometers"].duration,
                             s["interferometers"].sampling_frequency,
                             s["interferometers"].start_time)
            olk = OracleCalibrated(ev, "taylorf2", {ifo.name: lk.calibration_draws[ifo.name]},
                                   waveform_arguments=dict(s["waveform_generator"].waveform_arguments),
                                   time_marginalization=time, distance_marginalization=True,
                                   phase_marginalization=True, priors=dict(s["priors"]),
                                   distance_table=table)
            ref = olk.log_likelihood_ratio(dict(p))
            got = out["%s_log_likelihood" % ifo.name]
            say(ifo.name, got, ref)
            assert abs(got - ref) < tol
