"""Detector-frame sky parameters and detector time references
(bilby_amd.frames, base.py:1091-1137) on the CPU:
* the rotation of gw/geometry.py:216-282 takes the z axis onto the unit
  separation and is orthonormal;
* a separation along z leaves the direction unrotated: ra = azimuth + GMST,
  dec = pi/2 - zenith;
* zenith 0 points along the separation, so the delay between the two
  detectors is the light travel time |x_1 - x_2| / c (the largest it can be);
* the conversion and GMST equal the oracle's restatements (applied
  rotation by rotation, scalar GMST) across leap seconds;
* a detector time reference: <ifo>_time = geocent + delay(geocent) maps back
  to geocent within the delay's change over the delay (< 1e-9 s);
* RowBuilder with a frame gives the rows of the converted sky parameters."""
import numpy as np
import pytest

from bilby_amd import frames
from bilby_amd.detector import get_empty_interferometer
from oracle import core


def _separation():
    return (get_empty_interferometer("H1").geometry.vertex
            - get_empty_interferometer("L1").geometry.vertex)


def test_rotation_takes_z_onto_the_separation():
    rng = np.random.default_rng(1)
    for d in [_separation(), *rng.normal(size=(5, 3))]:
        r = frames.rotation_matrix_from_delta(d)
        assert np.allclose(r @ np.array([0.0, 0.0, 1.0]), d / np.linalg.norm(d), atol=1e-15)
        assert np.allclose(r @ r.T, np.eye(3), atol=1e-15)
        assert np.linalg.det(r) == pytest.approx(1.0, abs=1e-14)


def test_separation_along_z_leaves_the_direction():
    t = 1187008882.43
    z, a = np.array([0.3, 1.1, 2.5]), np.array([0.2, 4.0, 6.0])
    ra, dec = frames.zenith_azimuth_to_ra_dec(z, a, t, [0.0, 0.0, 5.0])
    gmst = frames.greenwich_mean_sidereal_time(t)
    assert np.allclose(ra, (a + gmst) % (2 * np.pi), atol=1e-14)
    assert np.allclose(dec, np.pi / 2 - z, atol=1e-14)


def test_zenith_zero_gives_the_light_travel_time_between_the_pair():
    h1 = get_empty_interferometer("H1").geometry.vertex
    l1 = get_empty_interferometer("L1").geometry.vertex
    t = 1126259462.4
    ra, dec = frames.zenith_azimuth_to_ra_dec(0.0, 0.7, t, h1 - l1)
    dt = (frames.time_delay_from_geocenter(h1, ra, dec, t)
          - frames.time_delay_from_geocenter(l1, ra, dec, t))
    assert dt == pytest.approx(-np.linalg.norm(h1 - l1) / frames.SPEED_OF_LIGHT, abs=1e-15)


def test_conversion_and_gmst_equal_the_oracle():
    rng = np.random.default_rng(2)
    d = _separation()
    # times around the 2015-07-01 and 2017-01-01 leap seconds and later
    times = [1119744015.7, 1119744016.2, 1119744017.9, 1167264016.5, 1167264018.1,
             1187008882.43, 1400000000.999]
    z = np.arccos(rng.uniform(-1, 1, len(times)))
    a = rng.uniform(0, 2 * np.pi, len(times))
    for t, zz, aa in zip(times, z, a):
        assert frames.greenwich_mean_sidereal_time(t) == core.greenwich_mean_sidereal_time(t)
        ra, dec = frames.zenith_azimuth_to_ra_dec(zz, aa, t, d)
        ora, odec = core.zenith_azimuth_to_ra_dec(zz, aa, t, d)
        assert abs(ra - ora) < 1e-12 and abs(dec - odec) < 1e-12
        for det in ("H1", "L1", "V1"):
            v = get_empty_interferometer(det).geometry.vertex
            assert frames.time_delay_from_geocenter(v, ra, dec, t) == pytest.approx(
                core.time_delay_from_geocenter(v, ra, dec, t), abs=1e-16)
    vec = frames.zenith_azimuth_to_ra_dec(z, a, np.array(times), d)
    assert np.allclose(vec[0], [core.zenith_azimuth_to_ra_dec(zz, aa, t, d)[0]
                                for t, zz, aa in zip(times, z, a)], atol=1e-12)


def test_detector_time_reference_round_trip():
    frame = frames.sky_frame("sky", "H1")
    v = get_empty_interferometer("H1").geometry.vertex
    geocent = np.array([1126259462.4, 1126259462.41])
    ra, dec = np.array([1.3, 4.0]), np.array([-1.2, 0.4])
    h1_time = geocent + frames.time_delay_from_geocenter(v, ra, dec, geocent)
    out = frame({"ra": ra, "dec": dec, "H1_time": h1_time, "geocent_time": np.zeros(2)})
    assert np.max(np.abs(out["geocent_time"] - geocent)) < 1e-9
    # without <ifo>_time the geocent time is kept (base.py:1131-1135)
    out = frame({"ra": ra, "dec": dec, "geocent_time": geocent})
    assert np.array_equal(out["geocent_time"], geocent)


def test_row_builder_applies_the_frame():
    from bilby_amd.abi import RowBuilder
    frame = frames.sky_frame("H1L1", "L1")
    rng = np.random.default_rng(3)
    n = 7
    p = dict(zenith=np.arccos(rng.uniform(-1, 1, n)), azimuth=rng.uniform(0, 2 * np.pi, n),
             L1_time=1126259462.4 + rng.uniform(-0.01, 0.01, n), psi=rng.uniform(0, np.pi, n),
             mass_1=np.full(n, 30.0), mass_2=np.full(n, 25.0), luminosity_distance=np.full(n, 400.))
    rows = RowBuilder(lambda x: x, {}, frame=frame)(dict(p))
    conv = frame(dict(p))
    plain = RowBuilder(lambda x: x, {})({k: v for k, v in conv.items()
                                         if k not in ("zenith", "azimuth", "L1_time")})
    assert np.array_equal(rows, plain)
    ra, dec = frames.zenith_azimuth_to_ra_dec(p["zenith"], p["azimuth"], p["L1_time"],
                                              _separation())
    assert np.allclose(rows[:, 0], ra) and np.allclose(rows[:, 1], dec)
    with pytest.raises(ValueError):
        frames.sky_frame("H1", "geocent")
