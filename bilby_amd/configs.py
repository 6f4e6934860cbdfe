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
                               distance=True, phase=True, time=True, device=0,
                               taylorf2_termination=None):
    rng = np.random.default_rng(seed)
    start = GPS_REFERENCE - duration / 2
    centre = start + duration / 2
    ifos = InterferometerList(list(detectors))
    ifos.set_strain_data_from_power_spectral_densities(sampling_frequency, duration, start,
                                                       random_state=rng)
    wfa = dict(reference_frequency=50.0, minimum_frequency=20.0)
    if taylorf2_termination is not None:
        wfa["taylorf2_termination"] = taylorf2_termination
    wfg = WaveformGenerator(duration=duration, sampling_frequency=sampling_frequency,
                            frequency_domain_source_model=taylorf2, waveform_arguments=wfa)
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


def gram_schmidt(G):
    """Orthonormalise the columns of G (modified Gram-Schmidt)."""
    Q = np.array(G, dtype=complex)
    for k in range(Q.shape[1]):
        for j in range(k):
            Q[:, k] -= np.vdot(Q[:, j], Q[:, k]) * Q[:, j]
        Q[:, k] /= np.linalg.norm(Q[:, k])
    return Q


def _tf2_amplitude(f, m1, m2, dl_mpc):
    """Leading-order TaylorF2 amplitude (the model's A(f), ignoring the
    inclination factor), used only to give the synthetic weights a
    realistic scale."""
    from .constants import gravitational_constant, parsec, solar_mass, speed_of_light
    mtsun = gravitational_constant * solar_mass / speed_of_light ** 3
    m = m1 + m2
    eta = m1 * m2 / m ** 2
    mc = eta ** 0.6 * m * mtsun
    return (np.sqrt(5 / 24) * np.pi ** (-2 / 3) * speed_of_light / (dl_mpc * 1e6 * parsec)
            * mc ** (5 / 6) * np.asarray(f) ** (-7 / 6))


def _nodes(n, f_lo, f_hi, duration):
    """n distinct frequencies on the 1/duration grid, geometrically spaced."""
    k = np.unique(np.round(np.geomspace(f_lo, f_hi, n) * duration))
    m = n
    while len(k) < n:
        m += 1
        k = np.unique(np.round(np.geomspace(f_lo, f_hi, m) * duration))
    return k[:n] / duration


def make_roq(n_draws=65536, seed=2468, duration=128.0, sampling_frequency=4096.0,
             detectors=("H1", "L1"), n_linear=400, n_quadratic=120, n_time_samples=1000,
             time_half_span=0.1, prior_half_width=0.07, delta_tc=2e-4, chi_range=0.0):
    """ROQ workload: linear basis of n_linear nodes and quadratic basis of
    n_quadratic nodes, complex64 weights built by Gram-Schmidt on seeded
    complex Gaussian vectors over a duration x sampling_frequency band; the
    linear weights are given on n_time_samples times spanning
    +-time_half_span s about the segment centre."""
    rng = np.random.default_rng(seed)
    start = GPS_REFERENCE - duration / 2
    centre = start + duration / 2
    ifos = InterferometerList(list(detectors))
    ifos.set_strain_data_from_zero_noise(sampling_frequency, duration, start)
    fl = _nodes(n_linear, 20.0, 1024.0, duration)
    fq = _nodes(n_quadratic, 20.0, 1024.0, duration)
    ts = np.linspace(duration / 2 - time_half_span, duration / 2 + time_half_span, n_time_samples)
    a_ref_l = _tf2_amplitude(fl, 30.0, 30.0, 1000.0)
    a_ref_q = _tf2_amplitude(fq, 30.0, 30.0, 1000.0)
    # physical scales exceed the float32 range: keep the complex64 mantissas of
    # unit-scale vectors and carry the scale as an exact power of two
    s_lin = 2.0 ** np.round(np.log2(20.0 * np.sqrt(n_time_samples) / np.linalg.norm(a_ref_l)))
    s_quad = 2.0 ** np.round(np.log2(100.0 / np.sum(a_ref_q ** 2)))
    weights = dict(time_samples=ts, frequency_nodes_linear=fl, frequency_nodes_quadratic=fq)
    for name in detectors:
        G = (rng.normal(size=(n_time_samples, n_linear))
             + 1j * rng.normal(size=(n_time_samples, n_linear)))
        W = gram_schmidt(G).astype(np.complex64)
        q = ((0.5 + rng.uniform(size=n_quadratic)) + 0.01j * rng.normal(size=n_quadratic))
        weights[name + "_linear"] = s_lin * W.astype(np.complex128)
        weights[name + "_quadratic"] = s_quad * q.astype(np.complex64).astype(np.complex128)
    wfg = WaveformGenerator(duration=duration, sampling_frequency=sampling_frequency,
                            frequency_domain_source_model=taylorf2,
                            waveform_arguments=dict(reference_frequency=20.0, minimum_frequency=20.0))
    priors = PriorDict(
        geocent_time=Uniform(centre - prior_half_width, centre + prior_half_width, name="geocent_time"),
        phase=Uniform(0, 2 * np.pi, name="phase", boundary="periodic"),
        luminosity_distance=UniformComovingVolume(100.0, 2000.0, name="luminosity_distance"),
    )
    n_t = int(np.ceil(2 * prior_half_width / delta_tc))
    dtc = 2 * prior_half_width / n_t
    draws = _sky_draws(rng, n_draws, centre, prior_half_width)
    draws.update(
        mass_1=rng.uniform(5, 50, n_draws),
        mass_2=rng.uniform(5, 50, n_draws),
        theta_jn=np.arccos(rng.uniform(-1, 1, n_draws)),
        phase=rng.uniform(0, 2 * np.pi, n_draws),
        luminosity_distance=rng.uniform(300.0, 2000.0, n_draws),
        chi_1=rng.uniform(-chi_range, chi_range, n_draws) if chi_range else np.zeros(n_draws),
        chi_2=rng.uniform(-chi_range, chi_range, n_draws) if chi_range else np.zeros(n_draws),
        time_jitter=rng.uniform(-dtc / 2, dtc / 2, n_draws),
    )
    return dict(interferometers=ifos, waveform_generator=wfg, priors=priors, draws=draws,
                weights=weights, delta_tc=delta_tc)


def make_uneven_events(n_events=64, seed=97531, rows_range=(256, 4096), chi_range=0.8,
                       durations=(4, 8, 16, 32, 64, 128), sampling_frequencies=(2048.0, 4096.0),
                       device=0):
    """Uneven workload: events with segment durations log-uniform over
    ``durations`` s, sampling rates from ``sampling_frequencies``, 2 or 3
    detectors, aligned-spin TaylorF2, time + distance + phase marginalised,
    per-event batches log-uniform in ``rows_range``."""
    rng = np.random.default_rng(seed)
    events = []
    for e in range(n_events):
        duration = float(durations[rng.integers(len(durations))])
        fs = float(sampling_frequencies[rng.integers(len(sampling_frequencies))])
        dets = ("H1", "L1", "V1") if rng.uniform() < 0.5 else ("H1", "L1")
        n_rows = int(np.exp(rng.uniform(np.log(rows_range[0]), np.log(rows_range[1]))))
        sub_seed = int(rng.integers(2 ** 31))
        setup = make_taylorf2_marginalised(n_draws=n_rows, seed=sub_seed, duration=duration,
                                           sampling_frequency=fs, detectors=dets,
                                           chi_range=chi_range, device=device)
        setup["n_rows"] = n_rows
        events.append(setup)
    return events
