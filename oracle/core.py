"""FP64 restatement of bilby's grids, noise, geometry, sources and inner products.

Test infrastructure only (see oracle/__init__.py).
"""
import numpy as np

# bilby/core/utils/constants.py:3-7
SPEED_OF_LIGHT = 299792458.0
PARSEC = 3.085677581491367e16
SOLAR_MASS = 1.988409870698050731911960804878414216e30
RADIUS_OF_EARTH = 6378136.6
GRAVITATIONAL_CONSTANT = 6.6743e-11
MTSUN_SI = GRAVITATIONAL_CONSTANT * SOLAR_MASS / SPEED_OF_LIGHT ** 3


# ---------------------------------------------------------------- grids / noise
def frequency_series(sampling_frequency, duration):
    """bilby/core/utils/series.py:115-134 (create_frequency_series)."""
    n_samples = np.round(duration * sampling_frequency)
    n_freq = int(np.round(n_samples / 2) + 1)
    return np.linspace(0, sampling_frequency / 2, num=n_freq)


def white_noise(number_of_samples, duration, rng):
    """bilby/core/utils/series.py:161-198 (safe_white_noise)."""
    frequencies = np.arange(number_of_samples) / duration
    norm1 = 0.5 * duration ** 0.5
    re1, im1 = rng.normal(0, norm1, (2, len(frequencies)))
    noise = re1 + 1j * im1
    noise[0] = 0
    noise[-1] = 0
    return noise, frequencies


def noise_realisation(psd_of_f, psd_fmin, psd_fmax, number_of_samples, duration, rng):
    """bilby/gw/detector/psd.py:350-376 (get_noise_realisation).

    ``psd_of_f`` is the PSD interpolant (bilby: interp1d fill=+inf,
    psd.py:236-244); bins outside [psd_fmin, psd_fmax] are set to 0.
    """
    noise, frequencies = white_noise(number_of_samples, duration, rng)
    with np.errstate(invalid="ignore"):
        strain = psd_of_f(frequencies) ** 0.5 * noise
    oob = (frequencies < psd_fmin) | (frequencies > psd_fmax)
    strain[oob] = 0j
    return np.nan_to_num(strain), frequencies


def psd_table_interpolant(freq, psd):
    """bilby/gw/detector/psd.py:236-244: linear interp1d, fill +inf outside."""
    freq = np.asarray(freq, float)
    psd = np.asarray(psd, float)

    def f(x):
        x = np.asarray(x, float)
        out = np.interp(x, freq, psd)
        out[(x < freq[0]) | (x > freq[-1])] = np.inf
        return out
    return f


def analytic_aligo_psd(f, scale=1.0):
    """Analytic aLIGO fit used by the benchmark configurations:
    S(f) = 1e-49 [x^-4.14 - 5 x^-2 + 111 (1 - x^2 + x^4/2) / (1 + x^2/2)],
    x = f / 215 Hz; S(0) = +inf."""
    f = np.asarray(f, float)
    out = np.full(f.shape, np.inf)
    nz = f > 0
    x = f[nz] / 215.0
    out[nz] = scale * 1e-49 * (x ** -4.14 - 5.0 * x ** -2
                               + 111.0 * (1 - x ** 2 + x ** 4 / 2) / (1 + x ** 2 / 2))
    return out


def frequency_mask(frequency_array, minimum_frequency, maximum_frequency, sampling_frequency):
    """bilby/gw/detector/strain_data.py:102-110, 142-164."""
    fmax = min(maximum_frequency, sampling_frequency / 2)
    return (frequency_array >= minimum_frequency) & (frequency_array <= fmax)


# ---------------------------------------------------------------- geometry
# bilby/gw/detector/detectors/{H1,L1,V1,K1}.interferometer
DETECTORS = {
    "H1": dict(latitude=46 + 27. / 60 + 18.528 / 3600,
               longitude=-(119 + 24. / 60 + 27.5657 / 3600), elevation=142.554,
               xarm_azimuth=125.9994, yarm_azimuth=215.9994,
               xarm_tilt=-6.195e-4, yarm_tilt=1.25e-5,
               minimum_frequency=20, maximum_frequency=2048),
    "L1": dict(latitude=30 + 33. / 60 + 46.4196 / 3600,
               longitude=-(90 + 46. / 60 + 27.2654 / 3600), elevation=-6.574,
               xarm_azimuth=197.7165, yarm_azimuth=287.7165,
               xarm_tilt=-3.121e-4, yarm_tilt=-6.107e-4,
               minimum_frequency=20, maximum_frequency=2048),
    "V1": dict(latitude=43 + 37. / 60 + 53.0921 / 3600,
               longitude=10 + 30. / 60 + 16.1878 / 3600, elevation=51.884,
               xarm_azimuth=70.5674, yarm_azimuth=160.5674,
               xarm_tilt=0.0, yarm_tilt=0.0,
               minimum_frequency=20, maximum_frequency=2048),
    "K1": dict(latitude=36 + 24 / 60 + 42.69722 / 3600,
               longitude=137 + 18 / 60 + 21.44171 / 3600, elevation=414.181,
               xarm_azimuth=90 - 60.39623489157727, yarm_azimuth=90 + 29.60357629670688,
               xarm_tilt=0.0031414, yarm_tilt=-0.0036270,
               minimum_frequency=20, maximum_frequency=2048),
}


def calculate_arm(arm_tilt, arm_azimuth, longitude, latitude):
    """bilby/gw/geometry.py:51-93 (angles in radians)."""
    e_long = np.array([-np.sin(longitude), np.cos(longitude), longitude * 0])
    e_lat = np.array([-np.sin(latitude) * np.cos(longitude),
                      -np.sin(latitude) * np.sin(longitude), np.cos(latitude)])
    e_h = np.array([np.cos(latitude) * np.cos(longitude),
                    np.cos(latitude) * np.sin(longitude), np.sin(latitude)])
    return (np.cos(arm_tilt) * np.cos(arm_azimuth) * e_long
            + np.cos(arm_tilt) * np.sin(arm_azimuth) * e_lat
            + np.sin(arm_tilt) * e_h)


def detector_geometry(name):
    """Detector tensor (bilby/gw/geometry.py:96-115, detector/geometry.py:250-306)
    and vertex (bilby/gw/utils.py:59-88, detector/geometry.py:200-214)."""
    d = DETECTORS[name]
    lat = d["latitude"] * np.pi / 180
    lon = d["longitude"] * np.pi / 180
    x = calculate_arm(d["xarm_tilt"], d["xarm_azimuth"] * np.pi / 180, lon, lat)
    y = calculate_arm(d["yarm_tilt"], d["yarm_azimuth"] * np.pi / 180, lon, lat)
    tensor = (np.outer(x, x) - np.outer(y, y)) / 2
    a, b = 6378137, 6356752.314
    radius = a ** 2 * (a ** 2 * np.cos(lat) ** 2 + b ** 2 * np.sin(lat) ** 2) ** (-0.5)
    h = d["elevation"]
    vertex = np.array([(radius + h) * np.cos(lat) * np.cos(lon),
                       (radius + h) * np.cos(lat) * np.sin(lon),
                       ((b / a) ** 2 * radius + h) * np.sin(lat)])
    return tensor, vertex


# bilby/gw/time.py:74-93
LEAP_SECONDS = np.array([
    46828800, 78364801, 109900802, 173059203, 252028804, 315187205, 346723206,
    393984007, 425520008, 457056009, 504489610, 551750411, 599184012,
    820108813, 914803214, 1025136015, 1119744016, 1167264017])


def greenwich_mean_sidereal_time(gps_time):
    """bilby/gw/time.py:114-192 (GPS -> UTC -> Julian day -> GMST, radians)."""
    gps_int = gps_time // 1
    n_leap = np.sum(gps_int > LEAP_SECONDS)
    second = gps_int - n_leap
    # datetime(1980, 1, 6) + seconds; julian_day time.py:54-63
    year, month, day = 1980, 1, 6
    julian_day = (367 * year - 7 * (year + (month + 9) // 12) // 4
                  + 275 * month // 9 + day + second / 86400.0 + 1721013.5)
    t_hi = (julian_day - 2451545.0) / 36525.0
    t_lo = (gps_time % 1) / (36525.0 * 86400.0)
    t = t_hi + t_lo
    st = 0.0 + (-6.2e-6 * t + 0.093104) * t ** 2 + 67310.54841
    st += 8640184.812866 * t_lo
    st += 3155760000.0 * t_lo
    st += 8640184.812866 * t_hi
    st += 3155760000.0 * t_hi
    return st * 2 * np.pi / 86400.0


def polarization_tensors(ra, dec, time, psi):
    """bilby/gw/geometry.py:118-186 (plus and cross)."""
    gmst = greenwich_mean_sidereal_time(time) % (2 * np.pi)
    phi = ra - gmst
    theta = np.pi / 2 - dec
    u = np.array([np.cos(phi) * np.cos(theta), np.cos(theta) * np.sin(phi), -np.sin(theta)])
    v = np.array([-np.sin(phi), np.cos(phi), 0.0])
    m = -u * np.sin(psi) - v * np.cos(psi)
    n = -u * np.cos(psi) + v * np.sin(psi)
    return np.outer(m, m) - np.outer(n, n), np.outer(m, n) + np.outer(n, m)


def antenna_response(tensor, ra, dec, time, psi):
    """bilby/gw/detector/interferometer.py:267-301 + geometry.py:261-279."""
    pp, pc = polarization_tensors(ra, dec, time, psi)
    return np.einsum("ij,ij->", tensor, pp), np.einsum("ij,ij->", tensor, pc)


def time_delay_from_geocenter(vertex, ra, dec, time):
    """bilby/gw/geometry.py:282-343."""
    gmst = greenwich_mean_sidereal_time(time) % (2 * np.pi)
    phi = ra - gmst
    theta = np.pi / 2 - dec
    omega = np.array([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)])
    return omega @ (np.zeros(3) - vertex) / SPEED_OF_LIGHT


def rotate_z_to_delta(vector, delta_x):
    """bilby/gw/geometry.py:216-282 applied to ``vector``: the rotations
    R1(alpha) about z, R2(beta) about y, R3(gamma) about z, one after the
    other (alpha = atan2(-d_y d_z, d_x), beta = acos(d_z), gamma =
    atan2(d_y, d_x) of the unit separation d)."""
    d = np.asarray(delta_x, float)
    d = d / np.linalg.norm(d)
    alpha = np.arctan2(-d[1] * d[2], d[0])
    beta = np.arccos(d[2])
    gamma = np.arctan2(d[1], d[0])
    x, y, z = (np.asarray(c, float) for c in vector)
    x, y = np.cos(alpha) * x - np.sin(alpha) * y, np.sin(alpha) * x + np.cos(alpha) * y
    x, z = np.cos(beta) * x + np.sin(beta) * z, -np.sin(beta) * x + np.cos(beta) * z
    x, y = np.cos(gamma) * x - np.sin(gamma) * y, np.sin(gamma) * x + np.cos(gamma) * y
    return x, y, z


def zenith_azimuth_to_ra_dec(zenith, azimuth, time, delta_x):
    """bilby/gw/utils.py:232-256 with geometry.py:347-380 and
    core/utils/conversion.py:27-30."""
    wx, wy, wz = rotate_z_to_delta((np.sin(zenith) * np.cos(azimuth),
                                    np.sin(zenith) * np.sin(azimuth), np.cos(zenith)), delta_x)
    theta = np.arccos(wz)
    phi = np.arctan2(wy, wx) % (2 * np.pi)
    ra = (phi + greenwich_mean_sidereal_time(time)) % (2 * np.pi)
    return ra, np.pi / 2 - theta


# ---------------------------------------------------------------- sources
def sinegaussian(frequency_array, hrss, Q, frequency, **kwargs):
    """bilby/gw/source.py:1153-1209."""
    tau = Q / (2.0 ** 0.5 * np.pi * frequency)
    temp = Q / (4.0 * np.pi ** 0.5 * frequency)
    fm = frequency_array - frequency
    fp = frequency_array + frequency
    negative_term = np.exp(-fm ** 2 * np.pi ** 2 * tau ** 2)
    positive_term = np.exp(-fp ** 2 * np.pi ** 2 * tau ** 2)
    h_plus = hrss * np.pi ** 0.5 * tau / 2 * (negative_term + positive_term) / (
        temp * (1 + np.exp(-Q ** 2))) ** 0.5
    h_cross = -1j * hrss * np.pi ** 0.5 * tau / 2 * (negative_term - positive_term) / (
        temp * (1 - np.exp(-Q ** 2))) ** 0.5
    return {"plus": h_plus + 0j, "cross": h_cross}


EULER_GAMMA = 0.5772156649015329


def taylorf2_phase_coefficients(m1, m2, chi1, chi2):
    """TaylorF2 3.5PN phase coefficients phi_k, and the ln-v coefficients
    phi_k^(l), so that

        Psi_PN(f) = 3/(128 eta) sum_k (phi_k + phi_k^(l) ln v) v^(k-5),
        v = (pi M f)^(1/3).

    Non-spinning terms: Buonanno, Iyer, Ochsner, Pan & Sathyaprakash 2009
    (arXiv:0907.0700, Eq. 3.18).  Aligned-spin terms (1.5PN, 2.5PN, 3PN,
    3.5PN spin-orbit; 2PN spin-spin) as collected in Khan et al. 2016
    (arXiv:1508.07253, Appendix B).  lalsimulation is the reference's
    implementation (bilby/gw/source.py:633-643); it is not available here,
    so parity with it is unpinned.
    """
    M = m1 + m2
    eta = m1 * m2 / M ** 2
    delta = (m1 - m2) / M
    chis = 0.5 * (chi1 + chi2)
    chia = 0.5 * (chi1 - chi2)
    pi = np.pi
    phi = np.zeros(8)
    phil = np.zeros(8)
    phi[0] = 1.0
    phi[1] = 0.0
    phi[2] = 3715.0 / 756.0 + 55.0 * eta / 9.0
    phi[3] = (-16.0 * pi + 113.0 * delta * chia / 3.0
              + (113.0 / 3.0 - 76.0 * eta / 3.0) * chis)
    phi[4] = (15293365.0 / 508032.0 + 27145.0 * eta / 504.0 + 3085.0 * eta ** 2 / 72.0
              + (-405.0 / 8.0 + 200.0 * eta) * chia ** 2
              - 405.0 / 4.0 * delta * chia * chis
              + (-405.0 / 8.0 + 5.0 * eta / 2.0) * chis ** 2)
    p5 = (38645.0 * pi / 756.0 - 65.0 * pi * eta / 9.0
          + delta * (-732985.0 / 2268.0 - 140.0 * eta / 9.0) * chia
          + (-732985.0 / 2268.0 + 24260.0 * eta / 81.0 + 340.0 * eta ** 2 / 9.0) * chis)
    phi[5] = p5
    phil[5] = 3.0 * p5
    phi[6] = (11583231236531.0 / 4694215680.0 - 6848.0 * EULER_GAMMA / 21.0
              - 640.0 * pi ** 2 / 3.0
              + (-15737765635.0 / 3048192.0 + 2255.0 * pi ** 2 / 12.0) * eta
              + 76055.0 * eta ** 2 / 1728.0 - 127825.0 * eta ** 3 / 1296.0
              - 6848.0 / 21.0 * np.log(4.0)
              + 2270.0 * pi * delta * chia / 3.0
              + (2270.0 * pi / 3.0 - 520.0 * pi * eta) * chis)
    phil[6] = -6848.0 / 21.0
    phi[7] = (77096675.0 * pi / 254016.0 + 378515.0 * pi * eta / 1512.0
              - 74045.0 * pi * eta ** 2 / 756.0
              + delta * (-25150083775.0 / 3048192.0 + 26804935.0 * eta / 6048.0
                         - 1985.0 * eta ** 2 / 48.0) * chia
              + (-25150083775.0 / 3048192.0 + 10566655595.0 * eta / 762048.0
                 - 1042165.0 * eta ** 2 / 3024.0 + 5345.0 * eta ** 3 / 36.0) * chis)
    return phi, phil, eta


def taylorf2_psi_pn(f, m1, m2, chi1, chi2):
    """Psi_PN(f) in radians (see taylorf2_phase_coefficients)."""
    phi, phil, eta = taylorf2_phase_coefficients(m1, m2, chi1, chi2)
    msec = (m1 + m2) * MTSUN_SI
    v = (np.pi * msec * np.asarray(f, float)) ** (1.0 / 3.0)
    lnv = np.log(v)
    s = np.zeros_like(v)
    for k in range(8):
        s = s + (phi[k] + phil[k] * lnv) * v ** (k - 5)
    return 3.0 / (128.0 * eta) * s


def taylorf2(frequency_array, mass_1, mass_2, luminosity_distance, theta_jn, phase,
             chi_1=0.0, chi_2=0.0, reference_frequency=50.0, minimum_frequency=20.0,
             maximum_frequency=None, taylorf2_termination="maximum_frequency", **kwargs):
    """Frequency-domain TaylorF2 (stationary-phase, leading-order amplitude,
    3.5PN phase, aligned spins) as lalsimulation's XLALSimInspiralTaylorF2
    computes it under the bilby wrapper bilby/gw/source.py:552-690 (masses
    in solar masses, distance in Mpc, aligned spins -> iota = theta_jn per
    conversion.py:146-153).

    Band: bins from minimum_frequency up to maximum_frequency (default
    frequency_array[-1], source.py:338-343).  bilby hands maximum_frequency to
    lalsimulation as fEnd (source.py:633-643), which fills every bin up to
    fEnd and ends at the Schwarzschild ISCO only for fEnd = 0; bilby then
    zeroes the bins outside [minimum_frequency, maximum_frequency]
    (source.py:618-619, 678-679).  taylorf2_termination='isco' (an opt-in
    of this build, no counterpart in the reference) also cuts at
    f_isco = 1/(6^(3/2) pi M).  maximum_frequency = np.inf with no
    minimum_frequency cut is the frequency-sequence form bilby's ROQ model
    uses (source.py:1125-1128: every node evaluated).

        h+ = A(f) (1 + cos^2 iota)/2 exp(-i Psi(f))
        hx = -i A(f) cos(iota) exp(-i Psi(f))
        A(f) = -sqrt(5/24) pi^(-2/3) (c/D) (G Mc/c^3)^(5/6) f^(-7/6)
        Psi(f) = Psi_PN(f) - Psi_PN(f_ref) - 2 phase - pi/4   (t_c = 0)

    The negative sign of A is lalsimulation's (amp0 = -4 m1 m2 / r MRSUN
    MTSUN sqrt(pi/12) times the Newtonian sqrt(-E'(v)/F(v)) v).
    """
    f = np.asarray(frequency_array, float)
    if maximum_frequency is None:
        maximum_frequency = f[-1]
    m_tot = mass_1 + mass_2
    eta = mass_1 * mass_2 / m_tot ** 2
    mchirp_sec = eta ** 0.6 * m_tot * MTSUN_SI
    dist_m = luminosity_distance * 1e6 * PARSEC
    band = (f >= minimum_frequency) & (f <= maximum_frequency)
    if taylorf2_termination == "isco":
        band &= f <= 1.0 / (6.0 ** 1.5 * np.pi * m_tot * MTSUN_SI)
    elif taylorf2_termination != "maximum_frequency":
        raise ValueError("taylorf2_termination %r" % (taylorf2_termination,))
    hp = np.zeros(f.shape, complex)
    hc = np.zeros(f.shape, complex)
    fb = f[band]
    amp = (-np.sqrt(5.0 / 24.0) * np.pi ** (-2.0 / 3.0) * SPEED_OF_LIGHT / dist_m
           * mchirp_sec ** (5.0 / 6.0) * fb ** (-7.0 / 6.0))
    psi = (taylorf2_psi_pn(fb, mass_1, mass_2, chi_1, chi_2)
           - taylorf2_psi_pn(np.array([reference_frequency]), mass_1, mass_2, chi_1, chi_2)[0]
           - 2.0 * phase - np.pi / 4)
    base = amp * np.exp(-1j * psi)
    ci = np.cos(theta_jn)
    hp[band] = base * (1 + ci ** 2) / 2
    hc[band] = -1j * base * ci
    return {"plus": hp, "cross": hc}


# ---------------------------------------------------------------- inner products
def noise_weighted_inner_product(aa, bb, power_spectral_density, duration):
    """bilby/gw/utils.py:118-138."""
    integrand = aa.conj() * bb / power_spectral_density
    return 4 / duration * integrand.sum()


def matched_filter_snr(signal, frequency_domain_strain, power_spectral_density, duration):
    """bilby/gw/utils.py:141-168."""
    rho = noise_weighted_inner_product(signal, frequency_domain_strain, power_spectral_density, duration)
    return rho / noise_weighted_inner_product(signal, signal, power_spectral_density, duration) ** 0.5


def ln_i0(value):
    """bilby/gw/utils.py:1006-1022: log(i0e(x)) + |x| (scipy i0e, as bilby)."""
    from scipy.special import i0e
    return np.log(i0e(value)) + np.abs(value)
