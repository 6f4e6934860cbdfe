"""FP64 restatement of bilby's GravitationalWaveTransient.log_likelihood_ratio
with time / distance / phase marginalisation.

Test infrastructure only (see oracle/__init__.py).
"""
import copy
import ctypes
import os

import numpy as np
from scipy.interpolate import RectBivariateSpline
from scipy.special import logsumexp

from . import core

_CLIB = []


def _oracle_clib():
    if not _CLIB:
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "liboracle.so")
        lib = None
        if os.path.exists(path):
            lib = ctypes.CDLL(path)
            lib.oracle_distance_table.restype = None
        _CLIB.append(lib)
    return _CLIB[0]


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class OracleDetector:
    """Per-detector data as bilby's Interferometer exposes it:
    masked strain (strain_data.py:212-233), PSD on the grid
    (interferometer.py:551-564), frequency mask (strain_data.py:142-164),
    detector tensor and vertex (detector/geometry.py:200-269)."""

    def __init__(self, name, frequency_array, strain, psd, mask, tensor=None, vertex=None):
        self.name = name
        self.frequency_array = frequency_array
        self.mask = mask
        self.strain = np.asarray(strain, complex) * mask
        self.psd = np.asarray(psd, float)
        if tensor is None:
            tensor, vertex = core.detector_geometry(name)
        self.tensor = tensor
        self.vertex = vertex


class OracleEvent:
    def __init__(self, detectors, duration, sampling_frequency, start_time):
        self.detectors = detectors
        self.duration = float(duration)
        self.sampling_frequency = float(sampling_frequency)
        self.start_time = float(start_time)
        self.frequency_array = detectors[0].frequency_array


def detector_response(det, polarizations, params, start_time, antenna_time=None):
    """bilby/gw/detector/interferometer.py:303-368 (calibration = ones); the
    antenna patterns at ``antenna_time`` when given (reference_time,
    :336-339), else at geocent_time."""
    t_antenna = params["geocent_time"] if antenna_time is None else antenna_time
    fp, fc = core.antenna_response(det.tensor, params["ra"], params["dec"],
                                   t_antenna, params["psi"])
    signal = polarizations["plus"] * det.mask * fp + polarizations["cross"] * det.mask * fc
    time_shift = core.time_delay_from_geocenter(det.vertex, params["ra"], params["dec"],
                                                params["geocent_time"])
    dt_geocent = params["geocent_time"] - start_time
    dt = dt_geocent + time_shift
    return signal * np.exp(-1j * 2 * np.pi * dt * det.frequency_array)


def source_polarizations(source, frequency_array, params, waveform_arguments):
    if source == "sinegaussian":
        return core.sinegaussian(frequency_array, params["hrss"], params["Q"], params["frequency"])
    if source == "taylorf2":
        p = dict(params)
        chi1 = p.get("chi_1", p.get("a_1", 0.0) * np.cos(p.get("tilt_1", 0.0)))
        chi2 = p.get("chi_2", p.get("a_2", 0.0) * np.cos(p.get("tilt_2", 0.0)))
        return core.taylorf2(frequency_array, p["mass_1"], p["mass_2"], p["luminosity_distance"],
                             p["theta_jn"], p["phase"], chi_1=chi1, chi_2=chi2,
                             **waveform_arguments)
    raise ValueError(source)


class OracleGWTransient:
    """bilby/gw/likelihood/base.py:150-229 (constructor side effects on the
    marginalised priors) and :419-446 (log_likelihood_ratio).

    ``priors`` is a dict of prior objects exposing ``minimum``, ``maximum``,
    ``prob`` and ``rescale`` like bilby's; it is copied, as bilby does
    (base.py:392-400), and the caller's dict is updated with the fixed
    values of the marginalised keys (base.py:186-222).
    """

    def __init__(self, event, source, waveform_arguments=None, time_marginalization=False,
                 distance_marginalization=False, phase_marginalization=False, priors=None,
                 jitter_time=True, distance_table=None):
        self.event = event
        self.source = source
        self.waveform_arguments = dict(waveform_arguments or {})
        self.time_marginalization = time_marginalization
        self.distance_marginalization = distance_marginalization
        self.phase_marginalization = phase_marginalization
        self.priors = dict(priors) if priors is not None else None
        self.jitter_time = jitter_time if time_marginalization else False
        # interferometers.reference_time (interferometer.py:336-339): when
        # set, the antenna patterns are taken at this time
        self.antenna_time = None
        self.duration = event.duration
        if time_marginalization:
            self._setup_time_marginalization()
            priors["geocent_time"] = float(event.start_time)
        if phase_marginalization:
            priors["phase"] = 0.0
        if distance_marginalization:
            prior = self.priors["luminosity_distance"]
            self._distance_array = np.linspace(prior.minimum, prior.maximum, int(1e4))
            self.distance_prior_array = np.array([prior.prob(d) for d in self._distance_array])
            self._ref_dist = prior.rescale(0.5)
            if distance_table is None:
                distance_table = self.create_lookup_table()
            self._dist_margd_loglikelihood_array = distance_table
            self._interp = RectBivariateSpline(self._d_inner_h_ref_array,
                                               self._optimal_snr_squared_ref_array,
                                               distance_table.T, kx=3, ky=3, s=0)
            priors["luminosity_distance"] = float(self._ref_dist)

    # ---- time: base.py:1027-1035
    def _setup_time_marginalization(self):
        ev = self.event
        self._delta_tc = 2 / ev.sampling_frequency
        self._times = ev.start_time + np.linspace(
            0, ev.duration, int(ev.duration / 2 * ev.sampling_frequency + 1))[1:]

    # ---- distance axes: base.py:890-914
    @property
    def _optimal_snr_squared_ref_array(self):
        return np.logspace(-5, 10, 400)

    @property
    def _d_inner_h_ref_array(self):
        if self.phase_marginalization:
            return np.logspace(-5, 10, 800)
        return np.hstack((-np.logspace(3, -3, 400), np.logspace(-3, 10, 400)))

    def create_lookup_table(self):
        """base.py:994-1018.  Uses the oracle's C restatement
        (oracle/c/distance_table.c) when built, else numpy."""
        lib = _oracle_clib()
        if lib is not None:
            x = np.ascontiguousarray(self._d_inner_h_ref_array)
            y = np.ascontiguousarray(self._optimal_snr_squared_ref_array)
            d = np.ascontiguousarray(self._distance_array, dtype=float)
            p = np.ascontiguousarray(self.distance_prior_array, dtype=float)
            out = np.zeros((len(y), len(x)))
            lib.oracle_distance_table(len(y), len(x), _dp(y), _dp(x), len(d), _dp(d), _dp(p),
                                      ctypes.c_double(self._ref_dist),
                                      int(self.phase_marginalization), _dp(out))
            return out
        return self.create_lookup_table_numpy()

    def create_lookup_table_numpy(self, rows=None):
        """base.py:994-1018 (vectorised over the d-axis; same sums)."""
        scaling = self._ref_dist / self._distance_array
        d_full = np.outer(self._d_inner_h_ref_array, scaling)
        h_full = np.outer(self._optimal_snr_squared_ref_array, scaling ** 2)
        if self.phase_marginalization:
            d_full = core.ln_i0(np.abs(d_full))
        delta = self._distance_array[1] - self._distance_array[0]
        prior_term = self.distance_prior_array * delta
        rows = range(400) if rows is None else rows
        table = np.zeros((400, 800))
        for ii in rows:
            table[ii] = logsumexp(d_full - h_full[ii][None, :] / 2, b=prior_term, axis=1)
        log_norm = logsumexp(0 / self._distance_array, b=self.distance_prior_array * delta)
        return table - log_norm

    def _interp_dist(self, d, h):
        """bilby/core/utils/calculus.py:221-262 (BoundedRectBivariateSpline,
        fill -inf outside the axes' bounding box)."""
        x = self._d_inner_h_ref_array
        y = self._optimal_snr_squared_ref_array
        d = np.atleast_1d(np.asarray(d, float))
        hh = np.broadcast_to(np.asarray(h, float), d.shape)
        out = self._interp(d, hh, grid=False)
        bad = (d < x.min()) | (d > x.max()) | (hh < y.min()) | (hh > y.max())
        out = np.where(bad, -np.inf, out)
        return out

    # ---- per-detector SNRs: base.py:260-354
    def calculate_snrs(self, polarizations, det, params):
        ev = self.event
        signal = detector_response(det, polarizations, params, ev.start_time,
                                   antenna_time=self.antenna_time)
        m = det.mask
        d_inner_h = core.noise_weighted_inner_product(signal[m], det.strain[m], det.psd[m], ev.duration)
        h_inner_h = core.noise_weighted_inner_product(signal[m], signal[m], det.psd[m], ev.duration).real
        arr = None
        if self.time_marginalization:
            with np.errstate(invalid="ignore", divide="ignore"):
                integrand = signal[0:-1] * det.strain.conj()[0:-1] / det.psd[0:-1]
            integrand = np.where(np.isfinite(det.psd[0:-1]), integrand, 0)
            arr = 4 / ev.duration * np.fft.fft(integrand)
        return d_inner_h, h_inner_h, arr

    def log_likelihood_ratio(self, parameters):
        """base.py:419-446."""
        params = copy.deepcopy(parameters)
        pols = source_polarizations(self.source, self.event.frequency_array, params,
                                    self.waveform_arguments)
        if pols is None:
            return np.nan_to_num(-np.inf)
        if self.time_marginalization and self.jitter_time:
            params["geocent_time"] += params["time_jitter"]
        d_inner_h = 0j
        h_inner_h = 0.0
        arr = None
        for det in self.event.detectors:
            dh, hh, a = self.calculate_snrs(pols, det, params)
            d_inner_h += dh
            h_inner_h += hh
            if a is not None:
                arr = a if arr is None else arr + a
        return float(np.real(self.compute_log_likelihood(d_inner_h, h_inner_h, arr, params)))

    # ---- base.py:448-477, 775-820
    def compute_log_likelihood(self, d_inner_h, h_inner_h, arr, params):
        if self.time_marginalization:
            return self.time_marginalized_likelihood(arr, h_inner_h, params)
        if self.distance_marginalization:
            return self.distance_marginalized_likelihood(d_inner_h, h_inner_h, params)
        if self.phase_marginalization:
            return self.phase_marginalized_likelihood(d_inner_h, h_inner_h)
        return np.real(d_inner_h) - h_inner_h / 2

    def _setup_rho(self, d_inner_h, h_inner_h, params):
        """base.py:879-885."""
        h_ref = h_inner_h * params["luminosity_distance"] ** 2 / self._ref_dist ** 2.
        d_ref = d_inner_h * params["luminosity_distance"] / self._ref_dist
        return d_ref, h_ref

    def distance_marginalized_likelihood(self, d_inner_h, h_inner_h, params):
        d_ref, h_ref = self._setup_rho(d_inner_h, h_inner_h, params)
        d_ref = np.abs(d_ref) if self.phase_marginalization else np.real(d_ref)
        out = self._interp_dist(d_ref, h_ref)
        return out if np.ndim(d_inner_h) else out[0]

    def phase_marginalized_likelihood(self, d_inner_h, h_inner_h):
        return core.ln_i0(np.abs(d_inner_h)) - h_inner_h / 2

    def time_marginalized_likelihood(self, arr, h_inner_h, params):
        times = self._times
        if self.jitter_time:
            times = self._times + params["time_jitter"]
        prior = self.priors["geocent_time"]
        tmask = (times >= prior.minimum) & (times <= prior.maximum)
        times = times[tmask]
        arr = arr[tmask]
        time_prior_array = prior.prob(times) * self._delta_tc
        if self.distance_marginalization:
            ll = self.distance_marginalized_likelihood(arr, h_inner_h, params)
        elif self.phase_marginalization:
            ll = self.phase_marginalized_likelihood(arr, h_inner_h)
        else:
            ll = arr.real - h_inner_h / 2
        return logsumexp(ll, b=time_prior_array)

    # ---- posterior reconstruction: base.py:580-742
    def time_arrays(self, params):
        """base.py:580-626 line by line: times on a 1/16384 s grid from the
        (jittered) geocent time modulo the segment, inside [min, max) of the
        prior; d_inner_h the FFT of the zero-padded h conj(d) / S (no 4/T, as
        there) at those times; <h|h>."""
        params = copy.deepcopy(params)
        ev = self.event
        if self.jitter_time:
            params["geocent_time"] += params["time_jitter"]
        pols = source_polarizations(self.source, ev.frequency_array, params,
                                    self.waveform_arguments)
        T = ev.duration
        n_time_steps = int(T * 16384)
        times = (params["geocent_time"] - ev.start_time) + np.arange(n_time_steps) / 16384.0
        times = times % T + ev.start_time
        prior = self.priors["geocent_time"]
        in_prior = (times >= prior.minimum) & (times < prior.maximum)
        times = times[in_prior]
        d_inner_h = np.zeros(len(times), dtype=complex)
        h_inner_h = 0.0
        for det in ev.detectors:
            n = len(det.strain)
            signal = detector_response(det, pols, params, ev.start_time,
                                       antenna_time=self.antenna_time)
            signal_long = np.zeros(n_time_steps, dtype=complex)
            data = np.zeros(n_time_steps, dtype=complex)
            psd = np.ones(n_time_steps)
            signal_long[:n] = signal
            data[:n] = np.conj(det.strain)
            psd[:n][det.mask] = det.psd[det.mask]
            d_inner_h += np.fft.fft(signal_long * data / psd)[in_prior]
            m = det.mask
            h_inner_h += core.noise_weighted_inner_product(signal[m], signal[m], det.psd[m],
                                                           T).real
        return times, d_inner_h, h_inner_h, params

    def time_posterior(self, params):
        """base.py:627-640 before the 1/1000 cut: the time-conditioned lnL
        (distance / phase marginalised as configured) times the prior."""
        times, d_inner_h, h_inner_h, params = self.time_arrays(params)
        prior = self.priors["geocent_time"]
        if self.distance_marginalization:
            log_like = self.distance_marginalized_likelihood(d_inner_h, h_inner_h, params)
        elif self.phase_marginalization:
            log_like = core.ln_i0(np.abs(d_inner_h)) - h_inner_h / 2
        else:
            log_like = d_inner_h.real - h_inner_h / 2
        return times, np.exp(log_like - np.max(log_like)) * prior.prob(times)

    def plain_inner_products(self, params):
        """base.py:686-694: <d|h>, <h|h> summed over detectors."""
        pols = source_polarizations(self.source, self.event.frequency_array, params,
                                    self.waveform_arguments)
        dh, hh = 0j, 0.0
        for det in self.event.detectors:
            signal = detector_response(det, pols, params, self.event.start_time,
                                       antenna_time=self.antenna_time)
            m = det.mask
            dh += core.noise_weighted_inner_product(signal[m], det.strain[m], det.psd[m],
                                                    self.duration)
            hh += core.noise_weighted_inner_product(signal[m], signal[m], det.psd[m],
                                                    self.duration).real
        return dh, hh

    def distance_posterior(self, params):
        """base.py:642-684."""
        dh, hh = self.plain_inner_products(params)
        d = dh * params["luminosity_distance"] / self._distance_array
        h = hh * params["luminosity_distance"] ** 2 / self._distance_array ** 2
        if self.phase_marginalization:
            log_like = core.ln_i0(np.abs(d)) - h / 2
        else:
            log_like = d.real - h / 2
        return self._distance_array, np.exp(log_like - np.max(log_like)) * self.distance_prior_array

    def phase_posterior(self, params):
        """base.py:712-742."""
        dh, hh = self.plain_inner_products(params)
        phases = np.linspace(0, 2 * np.pi, 101)
        log_post = (dh * np.exp(-2j * phases)).real - hh / 2
        return phases, np.exp(log_post - np.max(log_post))

    def noise_log_likelihood(self):
        """base.py:402-413."""
        log_l = 0.0
        for det in self.event.detectors:
            m = det.mask
            log_l -= abs(core.noise_weighted_inner_product(det.strain[m], det.strain[m], det.psd[m],
                                                           self.duration) / 2)
        return float(np.real(log_l))

    def log_likelihood(self, parameters):
        return self.log_likelihood_ratio(parameters) + self.noise_log_likelihood()
