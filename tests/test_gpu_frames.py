"""The reference's checks of the detector frame and the detector time
reference (likelihood_test.py:288-331) on the GPU path:
* reference_frame="H1L1" with zenith / azimuth equals the sky-frame
  likelihood of the converted ra / dec (assertEqual there; bit for bit here),
  plain and time + distance + phase marginalised;
* time_reference="H1" with H1_time = geocent + delay equals the geocentric
  likelihood to 8 places; time marginalisation is switched off for it
  (base.py:171-178)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("marginalised", [False, True])
def test_reference_frame_agrees_with_default(marginalised):
    from bilby_amd import frames
    from bilby_amd.configs import make_taylorf2_marginalised
    from bilby_amd.likelihood import GravitationalWaveTransient
    from tests.helpers import fixed_marginalised
    say("reference frame H1L1, marginalised", marginalised)
    s = make_taylorf2_marginalised(n_draws=32, seed=12, duration=4.0, sampling_frequency=2048.0,
                                   time=marginalised, distance=marginalised, phase=marginalised)
    flags = dict(time_marginalization=marginalised, distance_marginalization=marginalised,
                 phase_marginalization=marginalised)
    pri_a, pri_b = s["priors"].copy(), s["priors"].copy()
    sky = GravitationalWaveTransient(s["interferometers"], s["waveform_generator"], priors=pri_a,
                                     precision="fp64", **flags)
    det = GravitationalWaveTransient(s["interferometers"], s["waveform_generator"], priors=pri_b,
                                     precision="fp64", reference_frame="H1L1", **flags)
    draws = fixed_marginalised(s["draws"], pri_a) if marginalised else dict(s["draws"])
    n = len(draws["mass_1"])
    rng = np.random.default_rng(5)
    zen = np.arccos(rng.uniform(-1, 1, n))
    azi = rng.uniform(0, 2 * np.pi, n)
    from bilby_amd.detector import get_empty_interferometer
    delta = (get_empty_interferometer("H1").geometry.vertex
             - get_empty_interferometer("L1").geometry.vertex)
    ra, dec = frames.zenith_azimuth_to_ra_dec(zen, azi, draws["geocent_time"], delta)
    with_sky = dict(draws, ra=ra, dec=dec)
    with_frame = {k: v for k, v in draws.items() if k not in ("ra", "dec")}
    with_frame.update(zenith=zen, azimuth=azi)
    a = sky.log_likelihood_ratio_batch(with_sky)
    b = det.log_likelihood_ratio_batch(with_frame)
    say("max |diff|", float(np.max(np.abs(a - b))), "lnL range", a.min(), a.max())
    assert np.array_equal(a, b)
    one = {k: float(v[3]) for k, v in with_frame.items()}
    assert abs(det.log_likelihood_ratio(one) - a[3]) < 1e-9 * max(1.0, abs(a[3]))
    assert det.meta_data["reference_frame"] == "H1L1"


def test_time_reference_agrees_with_default():
    from bilby_amd import frames
    from bilby_amd.configs import make_taylorf2_marginalised
    from bilby_amd.detector import get_empty_interferometer
    from bilby_amd.likelihood import GravitationalWaveTransient
    say("time reference H1")
    s = make_taylorf2_marginalised(n_draws=32, seed=13, duration=4.0, sampling_frequency=2048.0,
                                   time=False, distance=False, phase=False)
    geo = GravitationalWaveTransient(s["interferometers"], s["waveform_generator"],
                                     priors=s["priors"].copy(), precision="fp64")
    h1 = GravitationalWaveTransient(s["interferometers"], s["waveform_generator"],
                                    priors=s["priors"].copy(), precision="fp64",
                                    time_reference="H1")
    draws = dict(s["draws"])
    v = get_empty_interferometer("H1").geometry.vertex
    with_h1 = dict(draws)
    with_h1["H1_time"] = draws["geocent_time"] + frames.time_delay_from_geocenter(
        v, draws["ra"], draws["dec"], draws["geocent_time"])
    with_h1["geocent_time"] = np.zeros_like(draws["geocent_time"])
    b = h1.log_likelihood_ratio_batch(with_h1)
    # every row: the geocentric likelihood at geocent = H1_time - delay(H1_time)
    conv = frames.sky_frame("sky", "H1")(dict(with_h1))
    a = geo.log_likelihood_ratio_batch(dict(draws, geocent_time=conv["geocent_time"]))
    assert np.array_equal(a, b)
    # the reference's check at the injection (8 places): the delay evaluated
    # at H1_time instead of geocent_time moves geocent_time by ~1e-8 s, which
    # only matters away from the likelihood's peak in time
    inj = dict(s["injection"])
    one = dict(inj)
    one["H1_time"] = inj["geocent_time"] + float(frames.time_delay_from_geocenter(
        v, inj["ra"], inj["dec"], inj["geocent_time"]))
    d8 = abs(h1.log_likelihood_ratio(one) - geo.log_likelihood_ratio(inj))
    say("injection |diff|", d8, "lnL", geo.log_likelihood_ratio(inj))
    assert d8 < 1e-8
    tm = GravitationalWaveTransient(s["interferometers"], s["waveform_generator"],
                                    priors=make_taylorf2_marginalised(
                                        n_draws=4, seed=13, duration=4.0,
                                        sampling_frequency=2048.0)["priors"].copy(),
                                    time_marginalization=True, time_reference="H1")
    assert not tm.time_marginalization and "geocent_time" not in tm.marginalized_parameters
