"""GPU checks of the multibanded likelihood with the IFFT-FFT <h|h>
(linear_interpolation=False; multiband.py:613-646, 766-783): the diagonal
coefficients in k_mb plus the odd-bin transforms of k_mb_fft_lds (bands of
M_b / 2 <= 8192 points, in LDS) and k_mb_fft_split (longer bands, four-step
over global scratch), against the FP64 oracle on the same weights:
* 16 s at 2048 Hz, network SNR ~ 50 (the reference test's configuration):
  float64 within 1e-10 max(1, |lnL|), float32 < 1e-3, and within the
  reference's 5e-3 of the full likelihood at the injection
  (likelihood_test.py:1380-1386; 1.5e-2 over the draws around it);
* 8 s at 4096 Hz, whose second band has M_b = 32768 (split transform, more
  (row, detector) problems than workgroups);
* with time marginalisation (T + D + P) on the 16 s event;
* invalid rows stay in their row."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def _event(T, FS, distance):
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
    inj = dict(mass_1=1.5, mass_2=1.3, chi_1=0.0, chi_2=0.0, luminosity_distance=distance,
               theta_jn=0.4, phase=1.3, ra=1.3, dec=-1.2, psi=0.659, geocent_time=tc)
    inject_signal(ifos, wfg, inj)
    priors = PriorDict(geocent_time=Uniform(tc - 0.1, tc + 0.1),
                       luminosity_distance=UniformComovingVolume(10.0, 100.0),
                       phase=Uniform(0.0, 2 * np.pi, boundary="periodic"))
    return ifos, wfg, inj, priors


def _oracle_event(ifos):
    from oracle.likelihood import OracleDetector, OracleEvent
    dets = [OracleDetector(i.name, i.frequency_array, i.frequency_domain_strain,
                           i.power_spectral_density_array, i.frequency_mask) for i in ifos]
    return OracleEvent(dets, ifos.duration, ifos.sampling_frequency, ifos.start_time)


def _draws(inj, n, rng, lk, flags, time_marg=False, start=None, fs=None):
    d = {k: np.full(n, float(v)) for k, v in inj.items()}
    d["mass_1"] = d["mass_1"] * (1 + rng.normal(0, 2e-5, n))
    d["mass_2"] = d["mass_2"] * (1 + rng.normal(0, 2e-5, n))
    d["ra"] = d["ra"] + rng.normal(0, 0.05, n)
    d["dec"] = d["dec"] + rng.normal(0, 0.05, n)
    if time_marg:
        d["geocent_time"] = np.full(n, float(start))
        d["time_jitter"] = rng.uniform(-1, 1, n) / fs
    else:
        d["geocent_time"] = d["geocent_time"] + rng.normal(0, 2e-4, n)
    d["luminosity_distance"] = d["luminosity_distance"] * rng.uniform(0.8, 1.2, n)
    if flags.get("phase"):
        d["phase"][:] = 0.0
    if flags.get("distance"):
        d["luminosity_distance"][:] = lk._ref_dist
    return d


def _pair(ifos, wfg, inj, priors, flags, time_marg=False):
    from bilby_amd.multiband import MBGravitationalWaveTransient
    from oracle.multiband import OracleMultiband
    mc = (inj["mass_1"] * inj["mass_2"]) ** 0.6 / (inj["mass_1"] + inj["mass_2"]) ** 0.2
    lk = MBGravitationalWaveTransient(ifos, wfg, reference_chirp_mass=mc, priors=priors.copy(),
                                      linear_interpolation=False, time_marginalization=time_marg,
                                      distance_marginalization=flags["distance"],
                                      phase_marginalization=flags["phase"], precision="fp64")
    olk = OracleMultiband(_oracle_event(ifos), "taylorf2", lk.setup,
                          waveform_arguments=dict(wfg.waveform_arguments),
                          time_marginalization=time_marg,
                          distance_marginalization=flags["distance"],
                          phase_marginalization=flags["phase"], priors=dict(priors),
                          distance_table=getattr(lk, "_dist_margd_loglikelihood_array", None))
    return lk, olk


@pytest.mark.parametrize("flags", [dict(distance=False, phase=False),
                                   dict(distance=True, phase=True)])
def test_ifft_fft_matches_oracle_and_full(flags):
    from bilby_amd.likelihood import GravitationalWaveTransient
    from tests.helpers import draws_as_dicts
    say("multiband IFFT-FFT, 16 s @ 2048 Hz", flags)
    ifos, wfg, inj, priors = _event(16.0, 2048.0, 60.0)
    lk, olk = _pair(ifos, wfg, inj, priors, flags)
    assert lk.linear_interpolation is False
    say("bands", lk.setup.number_of_bands, "Mbs", list(lk.setup.Mbs))
    full = GravitationalWaveTransient(ifos, wfg, priors=priors.copy(),
                                      distance_marginalization=flags["distance"],
                                      phase_marginalization=flags["phase"], precision="fp64")
    n = 32
    draws = _draws(inj, n, np.random.default_rng(4), lk, flags)
    for k in ("mass_1", "mass_2", "ra", "dec", "geocent_time"):   # row 0: the injection
        draws[k][0] = inj[k]
    ref = np.array([olk.log_likelihood_ratio(p) for p in draws_as_dicts(draws, range(n))])
    f64 = lk.log_likelihood_ratio_batch(draws)
    f32 = lk.log_likelihood_ratio_batch(draws, precision="fp32")
    fl = full.log_likelihood_ratio_batch(draws)
    say("fp64", np.max(np.abs(f64 - ref)), "fp32", np.max(np.abs(f32 - ref)),
        "vs full at the injection", abs(f64[0] - fl[0]), "over the draws",
        np.max(np.abs(f64 - fl)), "lnL range", ref.min(), ref.max())
    assert np.all(np.isfinite(ref))
    scale = max(1.0, float(np.max(np.abs(ref))))
    assert np.max(np.abs(f64 - ref)) < 1e-10 * scale
    assert np.max(np.abs(f32 - ref)) < 1e-3
    # the reference compares the two likelihoods at the injection
    assert abs(f64[0] - fl[0]) < 5e-3
    assert np.max(np.abs(f64 - fl)) < 1.5e-2
    bad = {k: v.copy() for k, v in draws.items()}
    bad["mass_2"][3] = -1.0
    out = lk.log_likelihood_ratio_batch(bad)
    assert out[3] == np.nan_to_num(-np.inf)
    keep = [i for i in range(n) if i != 3]
    assert np.array_equal(out[keep], f64[keep])


def test_ifft_fft_split_transform_matches_oracle():
    from tests.helpers import draws_as_dicts
    say("multiband IFFT-FFT, 8 s @ 4096 Hz (split transform)")
    flags = dict(distance=False, phase=True)
    ifos, wfg, inj, priors = _event(8.0, 4096.0, 60.0)
    lk, olk = _pair(ifos, wfg, inj, priors, flags)
    mh = max(int(m) // 2 for m in lk.setup.Mbs[1:])
    say("bands", lk.setup.number_of_bands, "Mbs", list(lk.setup.Mbs))
    assert mh > 8192
    n = 200                         # 600 (row, detector) problems: more than the workgroups
    draws = _draws(inj, n, np.random.default_rng(5), lk, flags)
    f64 = lk.log_likelihood_ratio_batch(draws)
    f32 = lk.log_likelihood_ratio_batch(draws, precision="fp32")
    sel = list(range(0, n, 5)) + [n - 1]
    ref = np.array([olk.log_likelihood_ratio(p) for p in draws_as_dicts(draws, sel)])
    say("fp64", np.max(np.abs(f64[sel] - ref)), "fp32", np.max(np.abs(f32[sel] - ref)),
        "lnL range", ref.min(), ref.max())
    scale = max(1.0, float(np.max(np.abs(ref))))
    assert np.max(np.abs(f64[sel] - ref)) < 1e-10 * scale
    assert np.max(np.abs(f32[sel] - ref)) < 1e-3
    # the first rows alone give the same values (other workgroup assignment)
    head = {k: v[:7].copy() for k, v in draws.items()}
    assert np.max(np.abs(lk.log_likelihood_ratio_batch(head) - f64[:7])) < 1e-12 * scale


def test_ifft_fft_time_marginalised_matches_oracle():
    from tests.helpers import draws_as_dicts
    flags = dict(distance=True, phase=True)
    say("multiband IFFT-FFT, time marginalised", flags)
    ifos, wfg, inj, priors = _event(16.0, 2048.0, 60.0)
    lk, olk = _pair(ifos, wfg, inj, priors, flags, time_marg=True)
    n = 24
    draws = _draws(inj, n, np.random.default_rng(6), lk, flags, time_marg=True,
                   start=ifos.start_time, fs=2048.0)
    ref = np.array([olk.log_likelihood_ratio(p) for p in draws_as_dicts(draws, range(n))])
    f64 = lk.log_likelihood_ratio_batch(draws)
    f32 = lk.log_likelihood_ratio_batch(draws, precision="fp32")
    say("fp64", np.max(np.abs(f64 - ref)), "fp32", np.max(np.abs(f32 - ref)),
        "lnL range", ref.min(), ref.max())
    assert np.all(np.isfinite(ref))
    scale = np.maximum(1.0, np.abs(ref))
    assert np.max(np.abs(f64 - ref) / scale) < 1e-8
    assert np.max(np.abs(f32 - ref)) < 1e-3


def test_ifft_fft_weights_round_trip_on_device(tmp_path):
    """A likelihood built from saved IFFT-FFT weights (multiband.py:689-712,
    .npz here) holds the same bands and tables and gives the same values, bit
    for bit, as the one that wrote them."""
    from bilby_amd.multiband import MBGravitationalWaveTransient
    flags = dict(distance=False, phase=True)
    say("multiband IFFT-FFT weights round trip")
    ifos, wfg, inj, priors = _event(16.0, 2048.0, 60.0)
    lk, _ = _pair(ifos, wfg, inj, priors, flags)
    fn = str(tmp_path / "mb_ifft_weights.npz")
    lk.save_weights(fn)
    lk2 = MBGravitationalWaveTransient(ifos, wfg, weights=fn, priors=priors.copy(),
                                       phase_marginalization=True, precision="fp64")
    assert lk2.linear_interpolation is False
    draws = _draws(inj, 16, np.random.default_rng(7), lk, flags)
    a = lk.log_likelihood_ratio_batch(draws)
    b = lk2.log_likelihood_ratio_batch(draws)
    say("max |a - b|", float(np.max(np.abs(a - b))))
    assert np.array_equal(a, b)
