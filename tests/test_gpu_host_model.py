"""Source models the caller evaluates (GWL_HOST_POLARIZATIONS): any
``frequency_domain_source_model`` is called on the host per sample, as
bilby's WaveformGenerator calls it (waveform_generator.py:113-141), and its
polarizations go through the same projection, inner products and
marginalisations on the GPU.  Checked against the built-in device model of
the same physics and against the FP64 oracle."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def _host_sine_gaussian(frequency_array, hrss, Q, frequency, **kwargs):
    """A user model: bilby's sine-Gaussian (source.py:1153-1209), in numpy."""
    from oracle import core
    return core.sinegaussian(frequency_array, hrss, Q, frequency)


def _host_taylorf2(frequency_array, mass_1, mass_2, luminosity_distance, theta_jn, phase,
                   chi_1=0.0, chi_2=0.0, **kwargs):
    """A user model: TaylorF2 in numpy, ended at maximum_frequency (the
    waveform arguments reach it as keywords)."""
    from oracle import core
    return core.taylorf2(frequency_array, mass_1, mass_2, luminosity_distance, theta_jn, phase,
                         chi_1=chi_1, chi_2=chi_2,
                         reference_frequency=kwargs.get("reference_frequency", 50.0),
                         minimum_frequency=kwargs.get("minimum_frequency", 20.0),
                         maximum_frequency=kwargs.get("maximum_frequency"))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_host_sine_gaussian_matches_device_model_and_oracle(precision):
    """Plain likelihood, H1+L1 4 s @ 2048 Hz: the host-evaluated model
    against the device sine-Gaussian (fp64 1e-9 relative; fp32 within the
    float32 escalation bound 1e-3) and the FP64 oracle (fp64 1e-8
    relative); rows as a dict and as a parameter array agree; a model
    returning None gives bilby's -inf sentinel for that row."""
    from bilby_amd.configs import make_sine_gaussian_plain
    from bilby_amd.likelihood import GravitationalWaveTransient
    from bilby_amd.waveform_generator import WaveformGenerator
    from tests.helpers import draws_as_dicts, oracle_from_setup
    say("host sine-Gaussian", precision)
    s = make_sine_gaussian_plain(n_draws=300, seed=11)
    ifos, dev_wfg = s["interferometers"], s["waveform_generator"]
    wfg = WaveformGenerator(duration=ifos.duration, sampling_frequency=ifos.sampling_frequency,
                            frequency_domain_source_model=_host_sine_gaussian)
    host = GravitationalWaveTransient(ifos, wfg, precision=precision)
    device = GravitationalWaveTransient(ifos, dev_wfg, precision="fp64")
    draws = s["draws"]
    got = host.log_likelihood_ratio_batch(draws)
    ref = device.log_likelihood_ratio_batch(draws)
    rel = np.max(np.abs(got - ref) / np.maximum(1.0, np.abs(ref)))
    say("vs device model: max relative", rel, "max |lnL|", np.max(np.abs(ref)))
    if precision == "fp64":
        assert rel < 1e-9
    else:
        assert np.max(np.abs(got - ref)) < 1e-3 + 1e-9 * np.max(np.abs(ref))
    olk, _ = oracle_from_setup(s)
    idx = [0, 1, 2, 3]
    oref = np.array([olk.log_likelihood_ratio(p) for p in draws_as_dicts(draws, idx)])
    oerr = np.max(np.abs(got[idx] - oref) / np.maximum(1.0, np.abs(oref)))
    say("vs oracle", oerr)
    assert oerr < (1e-8 if precision == "fp64" else 1e-3)
    rows = host.rows_from_parameters({k: v[:8] for k, v in draws.items()})
    assert np.array_equal(host.log_likelihood_ratio_batch(rows), got[:8])
    one = draws_as_dicts(draws, [5])[0]
    assert abs(host.log_likelihood_ratio(one) - got[5]) <= 1e-12 * max(1.0, abs(got[5]))

    def flaky(frequency_array, hrss, Q, frequency, **kw):
        return None if frequency > 300 else _host_sine_gaussian(frequency_array, hrss, Q, frequency)

    lk2 = GravitationalWaveTransient(ifos, WaveformGenerator(
        duration=ifos.duration, sampling_frequency=ifos.sampling_frequency,
        frequency_domain_source_model=flaky), precision=precision)
    out2 = lk2.log_likelihood_ratio_batch({k: v[:32] for k, v in draws.items()})
    bad = draws["frequency"][:32] > 300
    assert np.all(out2[bad] == -np.finfo(float).max)
    assert np.array_equal(out2[~bad], got[:32][~bad])


def test_host_taylorf2_time_distance_phase_marginalised():
    """Time + distance + phase marginalised, H1+L1+V1 4 s @ 2048 Hz: the
    host-evaluated TaylorF2 against the device TaylorF2 (fp64, both through
    the tiled transform: 1e-7 absolute) and the FP64 oracle (1e-6); the
    distance and phase reconstruction from its inner products runs; the
    time reconstruction refuses (it needs a device model)."""
    from bilby_amd.configs import make_taylorf2_marginalised
    from bilby_amd.likelihood import GravitationalWaveTransient
    from bilby_amd.waveform_generator import WaveformGenerator
    from tests.helpers import draws_as_dicts, fixed_marginalised, oracle_from_setup
    say("host TaylorF2 T+D+P")
    s = make_taylorf2_marginalised(n_draws=96, seed=19, duration=4.0, sampling_frequency=2048.0)
    ifos, dev_wfg = s["interferometers"], s["waveform_generator"]
    wfg = WaveformGenerator(duration=ifos.duration, sampling_frequency=ifos.sampling_frequency,
                            frequency_domain_source_model=_host_taylorf2,
                            waveform_arguments=dict(dev_wfg.waveform_arguments,
                                                    maximum_frequency=1024.0))
    kw = dict(time_marginalization=True, distance_marginalization=True,
              phase_marginalization=True, precision="fp64")
    olk, _ = oracle_from_setup(s)
    host = GravitationalWaveTransient(ifos, wfg, priors=s["priors"].copy(), **kw)
    device = GravitationalWaveTransient(ifos, dev_wfg, priors=s["priors"].copy(), **kw)
    draws = fixed_marginalised(s["draws"], host.priors)
    got = host.log_likelihood_ratio_batch(draws)
    ref = device.log_likelihood_ratio_batch(draws)
    say("vs device model: max |diff|", float(np.max(np.abs(got - ref))), "range",
        float(ref.min()), float(ref.max()))
    assert np.max(np.abs(got - ref)) < 1e-7
    idx = [0, 1, 2]
    oref = np.array([olk.log_likelihood_ratio(p) for p in draws_as_dicts(draws, idx)])
    say("vs oracle", float(np.max(np.abs(got[idx] - oref))))
    assert np.max(np.abs(got[idx] - oref)) < 1e-6
    p = draws_as_dicts(draws, [0])[0]
    rng = np.random.default_rng(3)
    d = host.generate_distance_sample_from_marginalized_likelihood(dict(p), rng=rng)
    rng = np.random.default_rng(3)
    d_dev = device.generate_distance_sample_from_marginalized_likelihood(dict(p), rng=rng)
    assert abs(d - d_dev) < 1e-6 * d_dev
    with pytest.raises(NotImplementedError):
        host.generate_time_sample_from_marginalized_likelihood(dict(p))
