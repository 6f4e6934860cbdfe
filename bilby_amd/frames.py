"""Detector-frame sky parameters and non-geocentre time references, as
bilby's GravitationalWaveTransient.get_sky_frame_parameters
(bilby/gw/likelihood/base.py:1091-1137) applies them before every
evaluation:

* ``reference_frame`` other than "sky" (two detectors, e.g. "H1L1"): the
  parameters carry ``zenith`` and ``azimuth`` in the frame whose z axis is
  the separation of the two detectors' vertices; ra / dec follow from the
  rotation of that axis onto the Earth frame and the Greenwich mean
  sidereal time (gw/utils.py:232-256, gw/geometry.py:216-282, 347-380,
  core/utils/conversion.py:27-30, gw/detector/networks.py:488-490);
* ``time_reference`` a detector (e.g. "H1"): the parameters carry
  ``<ifo>_time`` and geocent_time = <ifo>_time - (delay from the geocentre
  to that detector at <ifo>_time) (gw/geometry.py:283-345).

The conversion is vectorised over rows on the host, as part of building
the parameter rows (bilby_amd.abi.RowBuilder), so that every caller path
(batched calls, sampler pools, worker-process clients) applies it the same
way; the device evaluates the resulting sky-frame rows.
"""
import numpy as np

SPEED_OF_LIGHT = 299792458.0

# gw/time.py:74-93: GPS times at which leap seconds were added
LEAP_SECONDS = np.array([46828800, 78364801, 109900802, 173059203, 252028804, 315187205,
                         346723206, 393984007, 425520008, 457056009, 504489610, 551750411,
                         599184012, 820108813, 914803214, 1025136015, 1119744016, 1167264017],
                        dtype=float)


def greenwich_mean_sidereal_time(gps_time):
    """gw/time.py:114-165 with a zero equation of the equinoxes, in radians:
    the UTC day of the whole GPS second (leap seconds removed) as a Julian
    day counted from the GPS epoch 1980-01-06 (367 y - 7 (y + (m + 9) // 12)
    // 4 + 275 m // 9 + d = 723231 for that date, gw/time.py:53-66), the
    fractional second kept apart, and the IAU 1982 polynomial."""
    gps = np.asarray(gps_time, dtype=float)
    whole = np.floor(gps)
    n_leap = np.sum(whole[..., None] > LEAP_SECONDS, axis=-1)
    julian_day = (723231.0 + (whole - n_leap) / 86400.0) + 1721013.5
    t_hi = (julian_day - 2451545.0) / 36525.0
    t_lo = np.mod(gps, 1.0) / (36525.0 * 86400.0)
    t = t_hi + t_lo
    sidereal = (-6.2e-6 * t + 0.093104) * t ** 2 + 67310.54841
    sidereal = sidereal + 8640184.812866 * t_lo
    sidereal = sidereal + 3155760000.0 * t_lo
    sidereal = sidereal + 8640184.812866 * t_hi
    sidereal = sidereal + 3155760000.0 * t_hi
    return sidereal * 2 * np.pi / 86400.0


def rotation_matrix_from_delta(delta_x):
    """gw/geometry.py:216-282: R3(gamma) R2(beta) R1(alpha), which takes the z
    axis onto the unit vector of ``delta_x``."""
    d = np.asarray(delta_x, dtype=float)
    d = d / np.sqrt(np.sum(d ** 2))
    alpha = np.arctan2(-d[1] * d[2], d[0])
    beta = np.arccos(d[2])
    gamma = np.arctan2(d[1], d[0])
    r1 = np.array([[np.cos(alpha), -np.sin(alpha), 0.0], [np.sin(alpha), np.cos(alpha), 0.0],
                   [0.0, 0.0, 1.0]])
    r2 = np.array([[np.cos(beta), 0.0, np.sin(beta)], [0.0, 1.0, 0.0],
                   [-np.sin(beta), 0.0, np.cos(beta)]])
    r3 = np.array([[np.cos(gamma), -np.sin(gamma), 0.0], [np.sin(gamma), np.cos(gamma), 0.0],
                   [0.0, 0.0, 1.0]])
    return r3 @ r2 @ r1


def zenith_azimuth_to_ra_dec(zenith, azimuth, geocent_time, delta_x):
    """gw/utils.py:232-256 with gw/geometry.py:347-380: the direction in the
    detector-pair frame rotated to the Earth frame, theta / phi there, then
    ra = (phi + GMST) mod 2 pi, dec = pi / 2 - theta."""
    zenith = np.asarray(zenith, dtype=float)
    azimuth = np.asarray(azimuth, dtype=float)
    omega_prime = np.stack([np.sin(zenith) * np.cos(azimuth), np.sin(zenith) * np.sin(azimuth),
                            np.cos(zenith)])
    omega = np.tensordot(rotation_matrix_from_delta(delta_x), omega_prime, axes=1)
    theta = np.arccos(omega[2])
    phi = np.arctan2(omega[1], omega[0]) % (2 * np.pi)
    ra = (phi + greenwich_mean_sidereal_time(geocent_time)) % (2 * np.pi)
    dec = np.pi / 2 - theta
    return ra, dec


def time_delay_from_geocenter(vertex, ra, dec, time):
    """gw/geometry.py:283-345: (omega . (0 - vertex)) / c, omega the unit
    vector to the source at GMST(time)."""
    gmst = greenwich_mean_sidereal_time(time) % (2 * np.pi)
    phi = np.asarray(ra, dtype=float) - gmst
    theta = np.pi / 2 - np.asarray(dec, dtype=float)
    v = -np.asarray(vertex, dtype=float)
    return (np.sin(theta) * np.cos(phi) * v[0] + np.sin(theta) * np.sin(phi) * v[1]
            + np.cos(theta) * v[2]) / SPEED_OF_LIGHT


class SkyFrame:
    """The conversion of get_sky_frame_parameters for one likelihood
    (base.py:1091-1137): ``delta_x`` the vertex separation of the reference
    frame's two detectors (None for the sky frame), ``time_reference`` the
    name whose ``<name>_time`` parameter is the reference time and
    ``time_vertex`` that detector's vertex (None for the geocentre).
    Picklable; applied to a parameter dict of scalars or 1-D arrays."""

    def __init__(self, delta_x=None, time_reference="geocent", time_vertex=None):
        self.delta_x = None if delta_x is None else np.asarray(delta_x, dtype=float)
        self.time_reference = time_reference
        self.time_vertex = None if time_vertex is None else np.asarray(time_vertex, dtype=float)

    @property
    def is_identity(self):
        return self.delta_x is None and self.time_vertex is None

    def __call__(self, parameters):
        p = parameters
        key = "%s_time" % self.time_reference
        time = p.get(key)
        if time is None:
            time = p.get("geocent_time")
        if self.delta_x is not None:
            if "zenith" in p and "azimuth" in p:
                ra, dec = zenith_azimuth_to_ra_dec(p["zenith"], p["azimuth"], time, self.delta_x)
            elif "ra" in p and "dec" in p:   # base.py:1118-1125: fall back to ra / dec
                ra, dec = p["ra"], p["dec"]
            else:
                raise KeyError("zenith")
        else:
            ra, dec = p["ra"], p["dec"]
        if self.time_vertex is not None and key in p:
            geocent_time = np.asarray(p[key], dtype=float) - time_delay_from_geocenter(
                self.time_vertex, ra, dec, p[key])
        else:
            geocent_time = p["geocent_time"]
        p["ra"], p["dec"], p["geocent_time"] = ra, dec, geocent_time
        return p


def sky_frame(reference_frame="sky", time_reference="geocenter"):
    """The SkyFrame of a likelihood's ``reference_frame`` and
    ``time_reference`` arguments (base.py:1066-1089, 171-181): "sky", a
    string of two detector names ("H1L1"), or a list / InterferometerList
    whose first two entries are used; "geocent..." or a detector name."""
    from .detector import InterferometerList, get_empty_interferometer
    delta_x = None
    if not (isinstance(reference_frame, str) and reference_frame == "sky"):
        if isinstance(reference_frame, str):
            if len(reference_frame) < 4:
                raise ValueError("Unable to parse reference frame {}".format(reference_frame))
            pair = InterferometerList([reference_frame[:2], reference_frame[2:4]])
        elif isinstance(reference_frame, (list, tuple)):
            pair = InterferometerList(list(reference_frame)[:2])
        else:
            raise ValueError("Unable to parse reference frame {}".format(reference_frame))
        delta_x = pair[0].geometry.vertex - pair[1].geometry.vertex
    if "geocent" in time_reference:
        return SkyFrame(delta_x, "geocent", None)
    return SkyFrame(delta_x, time_reference,
                    get_empty_interferometer(time_reference).geometry.vertex)
