"""GPU checks of the multibanded likelihood (k_mb, bilby_amd.multiband)
against the FP64 oracle evaluation on the same weights: float64 within
1e-10 of |lnL|, float32 < 1e-3, for plain, phase and distance + phase marginalisation; the
multibanded value against the full GPU likelihood within the reference's
1.5e-2 (likelihood_test.py:1380-1428, network SNR ~ 50 as there); invalid
rows."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

T, FS = 16.0, 2048.0


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def _event():
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


def _draws(inj, n, rng):
    d = {k: np.full(n, float(v)) for k, v in inj.items()}
    d["mass_1"] = d["mass_1"] * (1 + rng.normal(0, 2e-5, n))
    d["mass_2"] = d["mass_2"] * (1 + rng.normal(0, 2e-5, n))
    d["ra"] = d["ra"] + rng.normal(0, 0.05, n)
    d["dec"] = d["dec"] + rng.normal(0, 0.05, n)
    d["geocent_time"] = d["geocent_time"] + rng.normal(0, 2e-4, n)
    d["luminosity_distance"] = d["luminosity_distance"] * rng.uniform(0.8, 1.2, n)
    return d


@pytest.mark.parametrize("flags", [dict(distance=False, phase=False),
                                   dict(distance=False, phase=True),
                                   dict(distance=True, phase=True)])
def test_multiband_matches_oracle_and_full(flags):
    from bilby_amd.likelihood import GravitationalWaveTransient
    from bilby_amd.multiband import MBGravitationalWaveTransient
    from oracle.likelihood import OracleDetector, OracleEvent
    from oracle.multiband import OracleMultiband
    from tests.helpers import draws_as_dicts
    say("multiband", flags)
    ifos, wfg, inj, priors = _event()
    mc = (inj["mass_1"] * inj["mass_2"]) ** 0.6 / (inj["mass_1"] + inj["mass_2"]) ** 0.2
    lk = MBGravitationalWaveTransient(ifos, wfg, reference_chirp_mass=mc, priors=priors.copy(),
                                      distance_marginalization=flags["distance"],
                                      phase_marginalization=flags["phase"], precision="fp64")
    full = GravitationalWaveTransient(ifos, wfg, priors=priors.copy(),
                                      distance_marginalization=flags["distance"],
                                      phase_marginalization=flags["phase"], precision="fp64")
    dets = [OracleDetector(i.name, i.frequency_array, i.frequency_domain_strain,
                           i.power_spectral_density_array, i.frequency_mask) for i in ifos]
    ev = OracleEvent(dets, ifos.duration, ifos.sampling_frequency, ifos.start_time)
    olk = OracleMultiband(ev, "taylorf2", lk.setup, waveform_arguments=dict(wfg.waveform_arguments),
                          distance_marginalization=flags["distance"],
                          phase_marginalization=flags["phase"], priors=dict(priors),
                          distance_table=getattr(lk, "_dist_margd_loglikelihood_array", None))
    n = 32
    draws = _draws(inj, n, np.random.default_rng(4))
    if flags["phase"]:
        draws["phase"][:] = 0.0
    if flags["distance"]:
        draws["luminosity_distance"][:] = lk._ref_dist
    ref = np.array([olk.log_likelihood_ratio(p) for p in draws_as_dicts(draws, range(n))])
    f64 = lk.log_likelihood_ratio_batch(draws)
    f32 = lk.log_likelihood_ratio_batch(draws, precision="fp32")
    fl = full.log_likelihood_ratio_batch(draws)
    say("points", len(lk.setup.banded_frequency_points), "bands", lk.setup.number_of_bands,
        "fp64", np.max(np.abs(f64 - ref)), "fp32", np.max(np.abs(f32 - ref)),
        "vs full", np.max(np.abs(f64 - fl)), "lnL range", ref.min(), ref.max())
    assert np.all(np.isfinite(ref))
    scale = max(1.0, float(np.max(np.abs(ref))))
    assert np.max(np.abs(f64 - ref)) < 1e-10 * scale     # float64 sums of ~15k terms
    assert np.max(np.abs(f32 - ref)) < 1e-3
    assert np.max(np.abs(f64 - fl)) < 1.5e-2
    bad = {k: v.copy() for k, v in draws.items()}
    bad["mass_2"][3] = -1.0
    out = lk.log_likelihood_ratio_batch(bad)
    assert out[3] == np.nan_to_num(-np.inf)
    keep = [i for i in range(n) if i != 3]
    assert np.array_equal(out[keep], f64[keep])
