"""Detectors, strain data and PSDs on the host: the setup side of the
likelihood, mirroring bilby.gw.detector (Interferometer, InterferometerList,
PowerSpectralDensity).  These objects hold host arrays; the likelihood
uploads them once and keeps them resident on the GPU.
"""
import numpy as np

from . import constants as C

# bilby/gw/detector/detectors/{H1,L1,V1,K1}.interferometer
DETECTOR_CONSTANTS = {
    "H1": dict(length=4, latitude=46 + 27. / 60 + 18.528 / 3600,
               longitude=-(119 + 24. / 60 + 27.5657 / 3600), elevation=142.554,
               xarm_azimuth=125.9994, yarm_azimuth=215.9994,
               xarm_tilt=-6.195e-4, yarm_tilt=1.25e-5,
               minimum_frequency=20, maximum_frequency=2048),
    "L1": dict(length=4, latitude=30 + 33. / 60 + 46.4196 / 3600,
               longitude=-(90 + 46. / 60 + 27.2654 / 3600), elevation=-6.574,
               xarm_azimuth=197.7165, yarm_azimuth=287.7165,
               xarm_tilt=-3.121e-4, yarm_tilt=-6.107e-4,
               minimum_frequency=20, maximum_frequency=2048),
    "V1": dict(length=3, latitude=43 + 37. / 60 + 53.0921 / 3600,
               longitude=10 + 30. / 60 + 16.1878 / 3600, elevation=51.884,
               xarm_azimuth=70.5674, yarm_azimuth=160.5674,
               xarm_tilt=0.0, yarm_tilt=0.0,
               minimum_frequency=20, maximum_frequency=2048),
    "K1": dict(length=3, latitude=36 + 24 / 60 + 42.69722 / 3600,
               longitude=137 + 18 / 60 + 21.44171 / 3600, elevation=414.181,
               xarm_azimuth=90 - 60.39623489157727, yarm_azimuth=90 + 29.60357629670688,
               xarm_tilt=0.0031414, yarm_tilt=-0.0036270,
               minimum_frequency=20, maximum_frequency=2048),
}


def create_frequency_series(sampling_frequency, duration):
    """bilby/core/utils/series.py:115-134."""
    n_samples = np.round(duration * sampling_frequency)
    n_freq = int(np.round(n_samples / 2) + 1)
    return np.linspace(0, sampling_frequency / 2, num=n_freq)


def _arm(tilt, azimuth, longitude, latitude):
    """Unit vector along an arm (bilby/gw/geometry.py:51-93)."""
    e_long = np.array([-np.sin(longitude), np.cos(longitude), 0.0])
    e_lat = np.array([-np.sin(latitude) * np.cos(longitude),
                      -np.sin(latitude) * np.sin(longitude), np.cos(latitude)])
    e_h = np.array([np.cos(latitude) * np.cos(longitude),
                    np.cos(latitude) * np.sin(longitude), np.sin(latitude)])
    return (np.cos(tilt) * np.cos(azimuth) * e_long + np.cos(tilt) * np.sin(azimuth) * e_lat
            + np.sin(tilt) * e_h)


class InterferometerGeometry:
    """bilby/gw/detector/geometry.py:7-316 (angles given in degrees, tilts in
    radians, elevation in metres)."""

    def __init__(self, length, latitude, longitude, elevation, xarm_azimuth, yarm_azimuth,
                 xarm_tilt=0.0, yarm_tilt=0.0):
        self.length = length
        self.latitude = latitude
        self.longitude = longitude
        self.elevation = elevation
        self.xarm_azimuth = xarm_azimuth
        self.yarm_azimuth = yarm_azimuth
        self.xarm_tilt = xarm_tilt
        self.yarm_tilt = yarm_tilt

    @property
    def x(self):
        return _arm(self.xarm_tilt, np.radians(self.xarm_azimuth), np.radians(self.longitude),
                    np.radians(self.latitude))

    @property
    def y(self):
        return _arm(self.yarm_tilt, np.radians(self.yarm_azimuth), np.radians(self.longitude),
                    np.radians(self.latitude))

    @property
    def detector_tensor(self):
        """(x x - y y) / 2 (bilby/gw/geometry.py:96-115)."""
        x, y = self.x, self.y
        return (np.outer(x, x) - np.outer(y, y)) / 2

    @property
    def vertex(self):
        """bilby/gw/utils.py:59-88 (ellipsoidal Earth)."""
        lat = np.radians(self.latitude)
        lon = np.radians(self.longitude)
        a, b = 6378137, 6356752.314
        radius = a ** 2 * (a ** 2 * np.cos(lat) ** 2 + b ** 2 * np.sin(lat) ** 2) ** (-0.5)
        h = self.elevation
        return np.array([(radius + h) * np.cos(lat) * np.cos(lon),
                         (radius + h) * np.cos(lat) * np.sin(lon),
                         ((b / a) ** 2 * radius + h) * np.sin(lat)])


class PowerSpectralDensity:
    """bilby/gw/detector/psd.py: a tabulated PSD interpolated linearly with
    +inf outside its range (psd.py:236-244), or an analytic function
    evaluated directly on the grid (``psd_function``)."""

    def __init__(self, frequency_array=None, psd_array=None, asd_array=None, psd_function=None):
        self.psd_function = psd_function
        self.frequency_array = None if frequency_array is None else np.asarray(frequency_array, float)
        if asd_array is not None:
            psd_array = np.asarray(asd_array, float) ** 2
        self.psd_array = None if psd_array is None else np.asarray(psd_array, float)

    @classmethod
    def from_amplitude_spectral_density_file(cls, path):
        f, asd = np.genfromtxt(path).T
        return cls(frequency_array=f, asd_array=asd)

    @classmethod
    def from_power_spectral_density_file(cls, path):
        f, psd = np.genfromtxt(path).T
        return cls(frequency_array=f, psd_array=psd)

    def power_spectral_density_interpolated(self, f):
        f = np.asarray(f, float)
        if self.psd_function is not None:
            return self.psd_function(f)
        out = np.interp(f, self.frequency_array, self.psd_array)
        out[(f < self.frequency_array[0]) | (f > self.frequency_array[-1])] = np.inf
        return out

    @property
    def _range(self):
        if self.psd_function is not None:
            return 0.0, np.inf
        return self.frequency_array.min(), self.frequency_array.max()

    def get_noise_realisation(self, number_of_samples, duration, random_state):
        """psd.py:350-376 with series.py:161-198 (safe_white_noise)."""
        from . import random as _random
        rng = _random.resolve_random_state(random_state)
        frequencies = np.arange(number_of_samples) / duration
        norm1 = 0.5 * duration ** 0.5
        re1, im1 = rng.normal(0, norm1, (2, len(frequencies)))
        white = re1 + 1j * im1
        white[0] = 0
        white[-1] = 0
        with np.errstate(invalid="ignore"):
            strain = self.power_spectral_density_interpolated(frequencies) ** 0.5 * white
        lo, hi = self._range
        strain[(frequencies < lo) | (frequencies > hi)] = 0j
        return np.nan_to_num(strain), frequencies


def aligo_analytic_psd(scale=1.0):
    """Analytic aLIGO fit S(f) = 1e-49 [x^-4.14 - 5 x^-2 + 111 (1 - x^2 +
    x^4/2)/(1 + x^2/2)], x = f/215 Hz (S(0) = +inf), times ``scale``."""
    def fn(f):
        f = np.asarray(f, float)
        out = np.full(f.shape, np.inf)
        nz = f > 0
        x = f[nz] / 215.0
        out[nz] = scale * 1e-49 * (x ** -4.14 - 5.0 * x ** -2
                                   + 111.0 * (1 - x ** 2 + x ** 4 / 2) / (1 + x ** 2 / 2))
        return out
    return PowerSpectralDensity(psd_function=fn)


class Interferometer:
    """bilby.gw.detector.Interferometer (interferometer.py) restricted to what
    the likelihood reads: geometry, frequency mask, masked frequency-domain
    strain and the PSD on the grid."""

    def __init__(self, name, power_spectral_density, minimum_frequency, maximum_frequency,
                 length, latitude, longitude, elevation, xarm_azimuth, yarm_azimuth,
                 xarm_tilt=0.0, yarm_tilt=0.0):
        self.name = name
        self.power_spectral_density = power_spectral_density
        self.minimum_frequency = minimum_frequency
        self._maximum_frequency = maximum_frequency
        self.geometry = InterferometerGeometry(length, latitude, longitude, elevation,
                                               xarm_azimuth, yarm_azimuth, xarm_tilt, yarm_tilt)
        self.sampling_frequency = None
        self.duration = None
        self.start_time = 0.0
        self.frequency_array = None
        self._strain = None
        # calibration model (bilby_amd.calibration); None = no recalibration
        self.calibration_model = None

    @property
    def maximum_frequency(self):
        """strain_data.py:102-110: clamped to the Nyquist frequency."""
        if self.sampling_frequency is None:
            return self._maximum_frequency
        return min(self._maximum_frequency, self.sampling_frequency / 2)

    @maximum_frequency.setter
    def maximum_frequency(self, value):
        self._maximum_frequency = value

    @property
    def frequency_mask(self):
        """strain_data.py:142-164."""
        f = self.frequency_array
        return (f >= self.minimum_frequency) & (f <= self.maximum_frequency)

    @property
    def mask_range(self):
        idx = np.nonzero(self.frequency_mask)[0]
        return int(idx[0]), int(idx[-1])

    @property
    def frequency_domain_strain(self):
        """strain_data.py:212-233: the data times the mask."""
        return self._strain * self.frequency_mask

    @property
    def power_spectral_density_array(self):
        return self.power_spectral_density.power_spectral_density_interpolated(self.frequency_array)

    def _set_grid(self, sampling_frequency, duration, start_time):
        self.sampling_frequency = float(sampling_frequency)
        self.duration = float(duration)
        self.start_time = float(start_time)
        self.frequency_array = create_frequency_series(self.sampling_frequency, self.duration)

    def set_strain_data_from_power_spectral_density(self, sampling_frequency, duration,
                                                    start_time=0, random_state=None):
        """strain_data.py:627-662."""
        self._set_grid(sampling_frequency, duration, start_time)
        strain, _ = self.power_spectral_density.get_noise_realisation(
            len(self.frequency_array), self.duration, random_state)
        self._strain = strain

    def set_strain_data_from_frequency_domain_strain(self, strain, sampling_frequency, duration,
                                                     start_time=0):
        self._set_grid(sampling_frequency, duration, start_time)
        strain = np.asarray(strain, complex)
        if strain.shape != self.frequency_array.shape:
            raise ValueError("Data frequencies do not match frequency_array")
        self._strain = strain.copy()

    def set_strain_data_from_zero_noise(self, sampling_frequency, duration, start_time=0):
        self._set_grid(sampling_frequency, duration, start_time)
        self._strain = np.zeros(len(self.frequency_array), complex)

    def add_to_strain(self, signal):
        self._strain = self._strain + np.asarray(signal, complex)


def get_empty_interferometer(name, power_spectral_density=None):
    """bilby/gw/detector/networks.py:420-456 for the detectors H1, L1, V1, K1
    (bilby reads K1's design PSD from a file; here, without a PSD given, K1
    takes the analytic aLIGO curve scaled by 4 and V1 by 2)."""
    if name not in DETECTOR_CONSTANTS:
        raise ValueError("Interferometer {} not implemented".format(name))
    kw = dict(DETECTOR_CONSTANTS[name])
    if power_spectral_density is None:
        power_spectral_density = aligo_analytic_psd({"V1": 2.0, "K1": 4.0}.get(name, 1.0))
    return Interferometer(name=name, power_spectral_density=power_spectral_density, **kw)


class InterferometerList(list):
    """bilby/gw/detector/networks.py:13-352 (the parts the likelihood uses)."""

    def __init__(self, interferometers):
        super().__init__()
        for ifo in interferometers:
            if isinstance(ifo, str):
                ifo = get_empty_interferometer(ifo)
            self.append(ifo)

    def _common(self, attr):
        vals = {getattr(ifo, attr) for ifo in self}
        if len(vals) != 1:
            raise ValueError("The {} of all interferometers are not the same".format(attr))
        return vals.pop()

    @property
    def duration(self):
        return self._common("duration")

    @property
    def sampling_frequency(self):
        return self._common("sampling_frequency")

    @property
    def start_time(self):
        return self._common("start_time")

    @property
    def frequency_array(self):
        return self[0].frequency_array

    def set_strain_data_from_power_spectral_densities(self, sampling_frequency, duration,
                                                      start_time=0, random_state=None):
        """networks.py:87-112: one realisation per detector, in detector
        order, each resolving ``random_state`` as bilby does."""
        for ifo in self:
            ifo.set_strain_data_from_power_spectral_density(
                sampling_frequency, duration, start_time, random_state=random_state)

    def set_strain_data_from_zero_noise(self, sampling_frequency, duration, start_time=0):
        for ifo in self:
            ifo.set_strain_data_from_zero_noise(sampling_frequency, duration, start_time)
