"""GPU checks of the relative-binning likelihood (k_relbin,
bilby_amd.relative) against the FP64 relative-binning oracle on the same
data: float64 < 1e-7, float32 < 1e-3 for plain, phase and distance + phase
marginalisation; invalid rows; serving; the fiducial update."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def _setup(seed=5):
    from bilby_amd.configs import make_taylorf2_marginalised
    return make_taylorf2_marginalised(n_draws=8, seed=seed, duration=4.0,
                                      sampling_frequency=2048.0, detectors=("H1", "L1", "V1"))


def _near(inj, n, rng):
    d = {k: np.full(n, float(v)) for k, v in inj.items()}
    d["mass_1"] = d["mass_1"] * (1 + rng.normal(0, 0.003, n))
    d["mass_2"] = d["mass_2"] * (1 + rng.normal(0, 0.003, n))
    d["ra"] = d["ra"] + rng.normal(0, 0.1, n)
    d["dec"] = d["dec"] + rng.normal(0, 0.1, n)
    d["psi"] = d["psi"] + rng.normal(0, 0.2, n)
    d["theta_jn"] = d["theta_jn"] + rng.normal(0, 0.1, n)
    d["geocent_time"] = d["geocent_time"] + rng.normal(0, 5e-4, n)
    d["luminosity_distance"] = d["luminosity_distance"] * rng.uniform(0.7, 1.3, n)
    d["phase"] = rng.uniform(0, 2 * np.pi, n)
    return d


def _oracle(s, inj, flags, epsilon):
    from oracle.likelihood import OracleDetector, OracleEvent
    from oracle.relative import OracleRelativeBinning
    ifos = s["interferometers"]
    dets = [OracleDetector(ifo.name, ifo.frequency_array, ifo.frequency_domain_strain,
                           ifo.power_spectral_density_array, ifo.frequency_mask) for ifo in ifos]
    ev = OracleEvent(dets, ifos.duration, ifos.sampling_frequency, ifos.start_time)
    priors = dict(s["priors"])
    return OracleRelativeBinning(ev, "taylorf2", inj,
                                 waveform_arguments=dict(s["waveform_generator"].waveform_arguments),
                                 distance_marginalization=flags["distance"],
                                 phase_marginalization=flags["phase"], priors=priors,
                                 epsilon=epsilon)


@pytest.mark.parametrize("flags", [dict(distance=False, phase=False),
                                   dict(distance=False, phase=True),
                                   dict(distance=True, phase=True)])
def test_relative_binning_matches_oracle(flags):
    from bilby_amd.relative import RelativeBinningGravitationalWaveTransient
    from tests.helpers import draws_as_dicts
    say("relative binning", flags)
    s = _setup()
    inj = dict(s["injection"])
    eps = 0.1
    lk = RelativeBinningGravitationalWaveTransient(
        s["interferometers"], s["waveform_generator"], fiducial_parameters=inj,
        distance_marginalization=flags["distance"], phase_marginalization=flags["phase"],
        priors=s["priors"].copy(), epsilon=eps, precision="fp64")
    olk = _oracle(s, inj, flags, eps)
    assert np.array_equal(lk.bin_freqs, olk.bin_freqs)
    rng = np.random.default_rng(3)
    draws = _near(inj, 48, rng)
    if flags["phase"]:
        draws["phase"][:] = 0.0
    if flags["distance"]:
        draws["luminosity_distance"][:] = lk._ref_dist
    ref = np.array([olk.log_likelihood_ratio(p) for p in draws_as_dicts(draws, range(48))])
    f64 = lk.log_likelihood_ratio_batch(draws)
    f32 = lk.log_likelihood_ratio_batch(draws, precision="fp32")
    one = lk.log_likelihood_ratio(draws_as_dicts(draws, [5])[0])
    say("bins", lk.number_of_bins, "fp64", np.max(np.abs(f64 - ref)), "fp32",
        np.max(np.abs(f32 - ref)), "lnL range", ref.min(), ref.max())
    assert np.all(np.isfinite(ref))
    assert np.max(np.abs(f64 - ref)) < 1e-7
    assert np.max(np.abs(f32 - ref)) < 1e-3
    assert abs(one - f64[5]) < 1e-9
    # invalid rows get nan_to_num(-inf), neighbours untouched
    bad = {k: v.copy() for k, v in draws.items()}
    bad["mass_1"][7] = -3.0
    out = lk.log_likelihood_ratio_batch(bad)
    assert out[7] == np.nan_to_num(-np.inf)
    keep = [i for i in range(48) if i != 7]
    assert np.array_equal(out[keep], f64[keep])


def test_relative_binning_served():
    from bilby_amd.relative import RelativeBinningGravitationalWaveTransient
    from bilby_amd.serving import LikelihoodServer
    say("relative binning: serving")
    s = _setup(seed=6)
    inj = dict(s["injection"])
    lk = RelativeBinningGravitationalWaveTransient(
        s["interferometers"], s["waveform_generator"], fiducial_parameters=inj, epsilon=0.1,
        precision="fp32")
    rng = np.random.default_rng(4)
    draws = _near(inj, 100, rng)
    direct = lk.log_likelihood_ratio_batch(draws)
    server = LikelihoodServer(hbm_budget=20e9, max_batch_rows=64)
    try:
        server.register("rb", lk)
        served = np.concatenate([server.submit("rb", lk.rows_from_parameters(
            {k: v[a:a + 50] for k, v in draws.items()})).result() for a in (0, 50)])
    finally:
        server.close()
    assert np.array_equal(served, direct)


def _bns_event():
    from bilby_amd import InterferometerList, WaveformGenerator, taylorf2
    from bilby_amd.likelihood import inject_signal
    T, fs = 4.0, 2048.0
    start = 1126259642.413 - T / 2
    ifos = InterferometerList(["H1", "L1", "V1"])
    ifos.set_strain_data_from_power_spectral_densities(fs, T, start,
                                                       random_state=np.random.default_rng(12))
    wfg = WaveformGenerator(duration=T, sampling_frequency=fs,
                            frequency_domain_source_model=taylorf2,
                            waveform_arguments=dict(reference_frequency=50.0,
                                                    minimum_frequency=20.0))
    inj = dict(mass_1=1.5, mass_2=1.3, chi_1=0.0, chi_2=0.0, luminosity_distance=60.0,
               theta_jn=0.5, phase=1.1, ra=1.2, dec=-0.4, psi=0.8, geocent_time=start + T / 2)
    inject_signal(ifos, wfg, inj)
    return ifos, wfg, inj


def test_fiducial_update_raises_the_full_likelihood():
    """update_fiducial_parameters (relative.py:158-171, 260-289) from a point
    2e-4 off in both masses of a light binary (the binned surface is valid
    across the bounds): the full likelihood at the
    updated fiducial point is higher than at the starting one."""
    from bilby_amd.likelihood import GravitationalWaveTransient
    from bilby_amd.prior import PriorDict, Uniform
    from bilby_amd.relative import RelativeBinningGravitationalWaveTransient
    say("relative binning: fiducial update")
    ifos, wfg, inj = _bns_event()
    full = GravitationalWaveTransient(ifos, wfg, precision="fp64")
    fid = dict(inj)
    fid["mass_1"] *= 1.0002
    fid["mass_2"] *= 1.0002
    priors = PriorDict(mass_1=Uniform(inj["mass_1"] * 0.999, inj["mass_1"] * 1.001),
                       mass_2=Uniform(inj["mass_2"] * 0.999, inj["mass_2"] * 1.001))
    moved = RelativeBinningGravitationalWaveTransient(
        ifos, wfg, fiducial_parameters=fid, priors=priors, epsilon=0.1,
        update_fiducial_parameters=True,
        maximization_kwargs=dict(maxiter=40, popsize=16, polish=False, tol=1e-10))
    before = full.log_likelihood_ratio(fid)
    after = full.log_likelihood_ratio(moved.fiducial_parameters)
    binned = moved.log_likelihood_ratio(moved.fiducial_parameters)
    say("bins", moved.number_of_bins, "full lnL before", before, "after", after,
        "binned at the new fiducial", binned)
    assert after > before
    # at its own fiducial point the binned value is the full one, less the
    # last grid point, which relative.py:325-329 leaves out of the last bin
    # when the last edge is the last masked frequency
    assert abs(binned - after) < 2e-3
