"""The benchmark workloads as seeded, reproducible setups.

Each ``make_*`` returns the detectors (with seeded coloured noise plus an
injected signal), the waveform generator, the priors and a seeded batch of
prior draws.  The same host arrays feed the GPU likelihood and the FP64
oracle used by the tests.

  sine_gaussian_plain  : H1+L1, 4 s at 2048 Hz, sine-Gaussian, no
                         marginalisation, 10^4 draws.
  taylorf2_marginalised: H1+L1+V1, 8 s at 4096 Hz, non-spinning TaylorF2,
                         time + distance + phase marginalisation, 4096 draws.
"""
import numpy as np

from .detector import InterferometerList
from .prior import PriorDict, Uniform, UniformComovingVolume
from .waveform_generator import WaveformGenerator, sinegaussian, taylorf2

GPS_REFERENCE = 1126259642.413


def _sky_draws(rng, n, tc_centre, tc_half_width):
    return dict(
        ra=rng.uniform(0, 2 * np.pi, n),
        dec=np.arcsin(rng.uniform(-1, 1, n)),
        psi=rng.uniform(0, np.pi, n),
        geocent_time=tc_centre + rng.uniform(-tc_half_width, tc_half_width, n),
    )


def make_sine_gaussian_plain(n_draws=10000, seed=1234, duration=4.0, sampling_frequency=2048.0,
                             detectors=("H1", "L1"), inject=True, device=0):
    rng = np.random.default_rng(seed)
    start = GPS_REFERENCE - duration / 2
    centre = start + duration / 2
    ifos = InterferometerList(list(detectors))
    ifos.set_strain_data_from_power_spectral_densities(sampling_frequency, duration, start,
                                                       random_state=rng)
    wfg = WaveformGenerator(duration=duration, sampling_frequency=sampling_frequency,
                            frequency_domain_source_model=sinegaussian)
    injection = dict(hrss=1e-22, Q=9.0, frequency=200.0, ra=1.375, dec=-1.2108, psi=2.659,
                     geocent_time=centre)
    if inject:
        from .likelihood import inject_signal
        inject_signal(ifos, wfg, injection, device=device)
    draws = _sky_draws(rng, n_draws, centre, 0.1)
    draws.update(
        frequency=rng.uniform(50, 500, n_draws),
        Q=rng.uniform(3, 20, n_draws),
        hrss=np.exp(rng.uniform(np.log(1e-23), np.log(1e-21), n_draws)),
    )
    priors = PriorDict()
    return dict(interferometers=ifos, waveform_generator=wfg, priors=priors, draws=draws,
                injection=injection, time_marginalization=False, distance_marginalization=False,
                phase_marginalization=False)


def make_taylorf2_marginalised(n_draws=4096, seed=4321, duration=8.0, sampling_frequency=4096.0,
                               detectors=("H1", "L1", "V1"), chi_range=0.0, inject=True,
                               distance=True, phase=True, time=True, device=0):
    rng = np.random.default_rng(seed)
    start = GPS_REFERENCE - duration / 2
    centre = start + duration / 2
    ifos = InterferometerList(list(detectors))
    ifos.set_strain_data_from_power_spectral_densities(sampling_frequency, duration, start,
                                                       random_state=rng)
    wfg = WaveformGenerator(duration=duration, sampling_frequency=sampling_frequency,
                            frequency_domain_source_model=taylorf2,
                            waveform_arguments=dict(reference_frequency=50.0,
                                                    minimum_frequency=20.0))
    injection = dict(mass_1=36.0, mass_2=29.0, chi_1=0.0, chi_2=0.0, luminosity_distance=800.0,
                     theta_jn=0.4, phase=1.3, ra=1.375, dec=-1.2108, psi=2.659,
                     geocent_time=centre + 0.013)
    if inject:
        from .likelihood import inject_signal
        inject_signal(ifos, wfg, injection, device=device)
    priors = PriorDict(
        geocent_time=Uniform(centre - 0.1, centre + 0.1, name="geocent_time"),
        luminosity_distance=UniformComovingVolume(100.0, 2000.0, name="luminosity_distance"),
        phase=Uniform(0, 2 * np.pi, name="phase", boundary="periodic"),
    )
    draws = _sky_draws(rng, n_draws, centre, 0.1)
    draws.update(
        mass_1=rng.uniform(5, 50, n_draws),
        mass_2=rng.uniform(5, 50, n_draws),
        theta_jn=np.arccos(rng.uniform(-1, 1, n_draws)),
        phase=rng.uniform(0, 2 * np.pi, n_draws),
        luminosity_distance=priors["luminosity_distance"].sample(n_draws, rng),
        chi_1=rng.uniform(-chi_range, chi_range, n_draws) if chi_range else np.zeros(n_draws),
        chi_2=rng.uniform(-chi_range, chi_range, n_draws) if chi_range else np.zeros(n_draws),
        time_jitter=rng.uniform(-1 / sampling_frequency, 1 / sampling_frequency, n_draws),
    )
    return dict(interferometers=ifos, waveform_generator=wfg, priors=priors, draws=draws,
                injection=injection, time_marginalization=time,
                distance_marginalization=distance, phase_marginalization=phase)


def marginalised_draws(setup, likelihood_priors):
    """Draws as a sampler would send them once the likelihood has fixed the
    marginalised keys (base.py:186-222): those keys take the prior's fixed
    value."""
    d = dict(setup["draws"])
    n = len(next(iter(d.values())))
    for key in ("geocent_time", "phase", "luminosity_distance"):
        v = likelihood_priors.get(key)
        if isinstance(v, float):
            d[key] = np.full(n, v)
    return d
