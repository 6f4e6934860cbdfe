"""FP64 restatement of bilby's calibration marginalisation
(bilby/gw/likelihood/base.py:286-346, 822-877), also with time
marginalisation (one FFT per response curve, :305-324), and of the
cubic-spline response model (bilby/gw/detector/calibration.py:257-383).

Test infrastructure only (see oracle/__init__.py).
"""
import copy

import numpy as np
from scipy.special import logsumexp

from . import core
from .likelihood import OracleGWTransient, detector_response, source_polarizations


def cubic_spline_factor(frequency_array, node_amplitudes, node_phases, minimum_frequency,
                        maximum_frequency):
    """calibration.py:302-383: natural log10-f cubic spline with the end rows
    M0 - 2 M1 + M2 = 0 (LIGO-T2300140 Eq. 9), response
    (1 + dA)(2 + i dphi)/(2 - i dphi)."""
    n = len(node_amplitudes)
    log_nodes = np.linspace(np.log10(minimum_frequency), np.log10(maximum_frequency), n)
    step = log_nodes[1] - log_nodes[0]
    tmp1 = np.zeros((n, n))
    tmp1[0, 0], tmp1[0, 1], tmp1[0, 2] = -1, 2, -1
    tmp1[-1, -3], tmp1[-1, -2], tmp1[-1, -1] = -1, 2, -1
    tmp2 = np.zeros((n, n))
    for i in range(1, n - 1):
        tmp1[i, i - 1], tmp1[i, i], tmp1[i, i + 1] = 1 / 6, 2 / 3, 1 / 6
        tmp2[i, i - 1], tmp2[i, i], tmp2[i, i + 1] = 1, -2, 1
    matrix = np.linalg.solve(tmp1, tmp2)
    f = np.asarray(frequency_array, float)
    with np.errstate(divide="ignore"):
        x = np.nan_to_num(np.log10(f) - log_nodes[0], neginf=0.0) / step
    prev = np.clip(x.astype(int), 0, n - 2)
    b = x - prev
    a = 1 - b
    c = (a ** 3 - a) / 6
    d = (b ** 3 - b) / 6

    def spline(values):
        values = np.asarray(values, float)
        coef = matrix.dot(values)
        return a * values[prev] + b * values[prev + 1] + c * coef[prev] + d * coef[prev + 1]

    amp = spline(node_amplitudes)
    phase = spline(node_phases)
    return np.nan_to_num((1 + amp) * (2 + 1j * phase) / (2 - 1j * phase))


class OracleCalibrated(OracleGWTransient):
    """GravitationalWaveTransient with calibration_marginalization=True.
    ``curves``: detector name -> complex [n_curves, n_masked_bins] on that
    detector's masked frequencies."""

    def __init__(self, event, source, curves, time_marginalization=False, **kwargs):
        super().__init__(event, source, time_marginalization=time_marginalization, **kwargs)
        self.curves = {k: np.asarray(v, complex) for k, v in curves.items()}
        self.number_of_response_curves = len(next(iter(self.curves.values())))

    def calibration_snrs(self, params):
        """base.py:332-346 summed over detectors: d_inner_h_array[c] =
        4/T sum conj(d) h cal_c / S, optimal_snr_squared_array[c] =
        4/T sum |h|^2 |cal_c|^2 / S."""
        pols = source_polarizations(self.source, self.event.frequency_array, params,
                                    self.waveform_arguments)
        dh = np.zeros(self.number_of_response_curves, complex)
        hh = np.zeros(self.number_of_response_curves)
        norm = 4 / self.duration
        for det in self.event.detectors:
            signal = detector_response(det, pols, params, self.event.start_time)
            m = det.mask
            x = norm * np.conj(det.strain[m]) * signal[m] / det.psd[m]
            y = norm * np.abs(signal[m]) ** 2 / det.psd[m]
            cal = self.curves[det.name]
            dh += cal.dot(x)
            hh += (np.abs(cal) ** 2).dot(y)
        return dh, hh

    def calibration_time_arrays(self, params):
        """base.py:305-324 summed over detectors: the integrand
        conj(d) h / S tiled over the curves, multiplied by each curve on the
        masked bins, FFT over frequency of [0:-1] -> d_inner_h_array
        [n_curves, N]; optimal_snr_squared_array as without time."""
        pols = source_polarizations(self.source, self.event.frequency_array, params,
                                    self.waveform_arguments)
        norm = 4 / self.duration
        arr, hh = 0, np.zeros(self.number_of_response_curves)
        for det in self.event.detectors:
            signal = detector_response(det, pols, params, self.event.start_time,
                                       antenna_time=self.antenna_time)
            m = det.mask
            with np.errstate(invalid="ignore", divide="ignore"):
                base = np.conj(det.strain) * signal / det.psd
            base = np.where(signal != 0, base, 0)
            integrand = np.tile(base, (self.number_of_response_curves, 1)).T
            integrand[m] *= self.curves[det.name].T
            arr = arr + norm * np.fft.fft(integrand[0:-1], axis=0).T
            y = norm * np.abs(signal[m]) ** 2 / det.psd[m]
            hh += (np.abs(self.curves[det.name]) ** 2).dot(y)
        return arr, hh

    def per_curve_log_likelihoods(self, parameters):
        """base.py:822-858 (get_calibration_log_likelihoods); with time
        marginalisation each curve's value is time-marginalised
        (base.py:794-820 per curve)."""
        params = copy.deepcopy(parameters)
        if self.time_marginalization:
            if self.jitter_time:
                params["geocent_time"] += params["time_jitter"]
            arr, hh = self.calibration_time_arrays(params)
            return np.array([self.time_marginalized_likelihood(arr[c], hh[c], params)
                             for c in range(self.number_of_response_curves)])
        dh, hh = self.calibration_snrs(params)
        if self.distance_marginalization:
            return self.distance_marginalized_likelihood(dh, hh, params)
        if self.phase_marginalization:
            return core.ln_i0(np.abs(dh)) - hh / 2
        return np.real(dh) - hh / 2

    def log_likelihood_ratio(self, parameters):
        """base.py:419-446 with :860-877."""
        p = parameters
        if not (p.get("mass_1", 1.0) > 0 and p.get("mass_2", 1.0) > 0
                and p.get("luminosity_distance", 1.0) > 0):
            return np.nan_to_num(-np.inf)
        ll = self.per_curve_log_likelihoods(parameters)
        return float(logsumexp(ll) - np.log(self.number_of_response_curves))

    def plain_inner_products(self, params):
        """base.py:286-294 via _calculate_inner_products (:686-694): <d|h>,
        <h|h> summed over detectors, the signal multiplied on the mask by
        response curve params['recalib_index'] when it is given (the
        distance and phase reconstruction after a calibration draw)."""
        if "recalib_index" not in params:
            return super().plain_inner_products(params)
        index = int(params["recalib_index"])
        pols = source_polarizations(self.source, self.event.frequency_array, params,
                                    self.waveform_arguments)
        dh, hh = 0j, 0.0
        for det in self.event.detectors:
            signal = detector_response(det, pols, params, self.event.start_time,
                                       antenna_time=self.antenna_time).copy()
            m = det.mask
            signal[m] *= self.curves[det.name][index]
            dh += core.noise_weighted_inner_product(signal[m], det.strain[m], det.psd[m],
                                                    self.duration)
            hh += core.noise_weighted_inner_product(signal[m], signal[m], det.psd[m],
                                                    self.duration).real
        return dh, hh

    def log_likelihood_ratio_with_curve(self, parameters, index):
        """The likelihood not marginalised over calibration, with response
        curve ``index`` applied to the signal (base.py:286-290,
        'recalib_index'); time-marginalised through the FFT of the calibrated
        signal (base.py:325-330) when time marginalisation is on."""
        params = copy.deepcopy(parameters)
        pols = source_polarizations(self.source, self.event.frequency_array, params,
                                    self.waveform_arguments)
        if self.time_marginalization and self.jitter_time:
            params["geocent_time"] += params["time_jitter"]
        d_inner_h, h_inner_h, arr = 0j, 0.0, None
        for det in self.event.detectors:
            signal = detector_response(det, pols, params, self.event.start_time,
                                       antenna_time=self.antenna_time)
            m = det.mask
            signal = signal.copy()
            signal[m] *= self.curves[det.name][index]
            d_inner_h += core.noise_weighted_inner_product(signal[m], det.strain[m], det.psd[m],
                                                           self.duration)
            h_inner_h += core.noise_weighted_inner_product(signal[m], signal[m], det.psd[m],
                                                           self.duration).real
            if self.time_marginalization:
                with np.errstate(invalid="ignore", divide="ignore"):
                    integrand = signal[0:-1] * det.strain.conj()[0:-1] / det.psd[0:-1]
                integrand = np.where(signal[0:-1] != 0, integrand, 0)
                a = 4 / self.duration * np.fft.fft(integrand)
                arr = a if arr is None else arr + a
        return float(np.real(self.compute_log_likelihood(d_inner_h, h_inner_h, arr, params)))
