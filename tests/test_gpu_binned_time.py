"""GPU checks of the time-marginalised binned likelihoods (the window GEMM of
csrc/gwl_window.hip behind k_relbin and k_mb) against their FP64 oracles on
the same data and setup:
* relative binning with time marginalisation (relative.py:376-391, 410-420)
  alone and with phase / distance marginalisation: float64 within
  1e-8 max(1, |lnL|), float32 < 1e-3, rows with time jitter;
* multibanding with time marginalisation (multiband.py:714-726, 792-800),
  16 s at 2048 Hz, network SNR ~ 50: the same bars;
* invalid rows stay in their row; a batch larger than one GEMM tile."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FLAGS = [dict(distance=False, phase=False), dict(distance=False, phase=True),
         dict(distance=True, phase=True), dict(distance=True, phase=False)]


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def _oracle_event(ifos):
    from oracle.likelihood import OracleDetector, OracleEvent
    dets = [OracleDetector(i.name, i.frequency_array, i.frequency_domain_strain,
                           i.power_spectral_density_array, i.frequency_mask) for i in ifos]
    return OracleEvent(dets, ifos.duration, ifos.sampling_frequency, ifos.start_time)


def _time_draws(inj, n, rng, start, fs, mass_sigma, lk, flags):
    d = {k: np.full(n, float(v)) for k, v in inj.items()}
    d["mass_1"] = d["mass_1"] * (1 + rng.normal(0, mass_sigma, n))
    d["mass_2"] = d["mass_2"] * (1 + rng.normal(0, mass_sigma, n))
    d["ra"] = d["ra"] + rng.normal(0, 0.05, n)
    d["dec"] = d["dec"] + rng.normal(0, 0.05, n)
    d["geocent_time"] = np.full(n, float(start))
    d["time_jitter"] = rng.uniform(-1, 1, n) / fs
    d["phase"] = np.zeros(n) if flags["phase"] else rng.uniform(0, 2 * np.pi, n)
    if flags["distance"]:
        d["luminosity_distance"] = np.full(n, lk._ref_dist)
    else:
        d["luminosity_distance"] = d["luminosity_distance"] * rng.uniform(0.8, 1.2, n)
    return d


def _check(f64, f32, ref):
    assert np.all(np.isfinite(ref))
    scale = np.maximum(1.0, np.abs(ref))
    assert np.max(np.abs(f64 - ref) / scale) < 1e-8, np.max(np.abs(f64 - ref) / scale)
    assert np.max(np.abs(f32 - ref)) < 1e-3, np.max(np.abs(f32 - ref))


@pytest.mark.parametrize("flags", FLAGS)
def test_relative_binning_time_marginalised_matches_oracle(flags):
    from bilby_amd.configs import make_taylorf2_marginalised
    from bilby_amd.relative import RelativeBinningGravitationalWaveTransient
    from oracle.relative import OracleRelativeBinning
    from tests.helpers import draws_as_dicts
    say("relative binning, time marginalised", flags)
    s = make_taylorf2_marginalised(n_draws=8, seed=5, duration=4.0, sampling_frequency=2048.0,
                                   detectors=("H1", "L1", "V1"))
    ifos = s["interferometers"]
    inj = dict(s["injection"])
    eps = 0.1
    lk = RelativeBinningGravitationalWaveTransient(
        ifos, s["waveform_generator"], fiducial_parameters=inj, time_marginalization=True,
        distance_marginalization=flags["distance"], phase_marginalization=flags["phase"],
        priors=s["priors"].copy(), epsilon=eps, precision="fp64")
    olk = OracleRelativeBinning(
        _oracle_event(ifos), "taylorf2", inj,
        waveform_arguments=dict(s["waveform_generator"].waveform_arguments),
        time_marginalization=True, distance_marginalization=flags["distance"],
        phase_marginalization=flags["phase"], priors=dict(s["priors"]), epsilon=eps,
        distance_table=getattr(lk, "_dist_margd_loglikelihood_array", None))
    assert np.array_equal(lk.bin_freqs, olk.bin_freqs)
    n = 40
    draws = _time_draws(inj, n, np.random.default_rng(7), ifos.start_time,
                        ifos.sampling_frequency, 1e-3, lk, flags)
    ref = np.array([olk.log_likelihood_ratio(p) for p in draws_as_dicts(draws, range(n))])
    f64 = lk.log_likelihood_ratio_batch(draws)
    f32 = lk.log_likelihood_ratio_batch(draws, precision="fp32")
    say("bins", lk.number_of_bins, "fp64", np.max(np.abs(f64 - ref)), "fp32",
        np.max(np.abs(f32 - ref)), "lnL range", ref.min(), ref.max())
    _check(f64, f32, ref)
    bad = {k: v.copy() for k, v in draws.items()}
    bad["mass_2"][3] = -1.0
    out = lk.log_likelihood_ratio_batch(bad)
    assert out[3] == np.nan_to_num(-np.inf)
    keep = [i for i in range(n) if i != 3]
    assert np.array_equal(out[keep], f64[keep])


def test_relative_binning_time_marginalised_large_batch():
    """More rows than one GEMM tile (8192): every tile equals the rows
    evaluated on their own."""
    from bilby_amd.configs import make_taylorf2_marginalised
    from bilby_amd.relative import RelativeBinningGravitationalWaveTransient
    say("relative binning, time marginalised, 9000 rows")
    s = make_taylorf2_marginalised(n_draws=8, seed=5, duration=4.0, sampling_frequency=2048.0,
                                   detectors=("H1", "L1", "V1"))
    ifos = s["interferometers"]
    inj = dict(s["injection"])
    flags = dict(distance=True, phase=True)
    lk = RelativeBinningGravitationalWaveTransient(
        ifos, s["waveform_generator"], fiducial_parameters=inj, time_marginalization=True,
        distance_marginalization=True, phase_marginalization=True, priors=s["priors"].copy(),
        epsilon=0.5, precision="fp32")
    draws = _time_draws(inj, 9000, np.random.default_rng(8), ifos.start_time,
                        ifos.sampling_frequency, 1e-3, lk, flags)
    full = lk.log_likelihood_ratio_batch(draws)
    tail = lk.log_likelihood_ratio_batch({k: v[8190:] for k, v in draws.items()})
    assert np.all(np.isfinite(full))
    assert np.max(np.abs(full[8190:] - tail)) < 1e-9


T, FS = 16.0, 2048.0


def _mb_event():
    from bilby_amd import InterferometerList, WaveformGenerator, taylorf2
    from bilby_amd.likelihood import inject_signal
    from bilby_amd.prior import PriorDict, Uniform, UniformComovingVolume
    tc = 1187008882.0
    start = tc - T + 2.0
    ifos = InterferometerList(["H1", "L1", "V1"])
    ifos.set_strain_data_from_power_spectral_densities(FS, T, start,
                                                       random_state=np.random.default_rng(70817))
    wfg = WaveformGenerator(duration=T, sampling_frequency=FS,
                            frequency_domain_source_model=taylorf2,
                            waveform_arguments=dict(reference_frequency=20.0,
                                                    minimum_frequency=20.0))
    inj = dict(mass_1=1.5, mass_2=1.3, chi_1=0.0, chi_2=0.0, luminosity_distance=60.0,
               theta_jn=0.4, phase=1.3, ra=1.3, dec=-1.2, psi=0.659, geocent_time=tc)
    inject_signal(ifos, wfg, inj)
    priors = PriorDict(geocent_time=Uniform(tc - 0.1, tc + 0.1),
                       luminosity_distance=UniformComovingVolume(10.0, 100.0),
                       phase=Uniform(0.0, 2 * np.pi, boundary="periodic"))
    return ifos, wfg, inj, priors


@pytest.mark.parametrize("flags", FLAGS)
def test_multiband_time_marginalised_matches_oracle(flags):
    from bilby_amd.multiband import MBGravitationalWaveTransient
    from oracle.multiband import OracleMultiband
    from tests.helpers import draws_as_dicts
    say("multiband, time marginalised", flags)
    ifos, wfg, inj, priors = _mb_event()
    mc = (inj["mass_1"] * inj["mass_2"]) ** 0.6 / (inj["mass_1"] + inj["mass_2"]) ** 0.2
    lk = MBGravitationalWaveTransient(ifos, wfg, reference_chirp_mass=mc, priors=priors.copy(),
                                      time_marginalization=True,
                                      distance_marginalization=flags["distance"],
                                      phase_marginalization=flags["phase"], precision="fp64")
    olk = OracleMultiband(_oracle_event(ifos), "taylorf2", lk.setup,
                          waveform_arguments=dict(wfg.waveform_arguments),
                          time_marginalization=True, distance_marginalization=flags["distance"],
                          phase_marginalization=flags["phase"], priors=dict(priors),
                          distance_table=getattr(lk, "_dist_margd_loglikelihood_array", None))
    n = 24
    draws = _time_draws(inj, n, np.random.default_rng(4), ifos.start_time, FS, 2e-5, lk, flags)
    ref = np.array([olk.log_likelihood_ratio(p) for p in draws_as_dicts(draws, range(n))])
    f64 = lk.log_likelihood_ratio_batch(draws)
    f32 = lk.log_likelihood_ratio_batch(draws, precision="fp32")
    N, j_first, n_window = lk._time_window
    say("points", len(lk.setup.banded_frequency_points), "N", N, "window", n_window,
        "fp64", np.max(np.abs(f64 - ref)), "fp32", np.max(np.abs(f32 - ref)),
        "lnL range", ref.min(), ref.max())
    _check(f64, f32, ref)
    bad = {k: v.copy() for k, v in draws.items()}
    bad["mass_2"][3] = -1.0
    out = lk.log_likelihood_ratio_batch(bad)
    assert out[3] == np.nan_to_num(-np.inf)
    keep = [i for i in range(n) if i != 3]
    assert np.array_equal(out[keep], f64[keep])
