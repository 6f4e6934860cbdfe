"""FP64 restatement of bilby's MBGravitationalWaveTransient evaluation
(bilby/gw/likelihood/multiband.py:728-807: <h|h> by linear interpolation or
by the IFFT-FFT algorithm of :766-783; time marginalisation :714-726,
792-800) on given multiband weights, and a literal loop form of the
quadratic coefficients (multiband.py:551-611) to check the product's
prefix-sum form against.

Test infrastructure only (see oracle/__init__.py).
"""
import copy
import math
import types

import numpy as np

from . import core
from .likelihood import OracleGWTransient, detector_response, source_polarizations


class OracleMultiband(OracleGWTransient):
    """``weights``: an object or dict with banded_frequency_points,
    unique_to_original_frequencies, waveform_frequencies, linear_coeffs and
    quadratic_coeffs (detector name -> array); with time marginalisation
    also durations, Nbs and start_end_idxs; with linear_interpolation False
    instead of quadratic_coeffs: durations, Nbs, Mbs, Ks_Ke,
    start_end_idxs, Tbhats, windows, square_root_windows and Ibcs
    (detector name -> {str(band): array})."""

    def __init__(self, event, source, weights, time_marginalization=False, **kwargs):
        super().__init__(event, source, time_marginalization=time_marginalization, **kwargs)
        w = weights if isinstance(weights, dict) else weights.weights
        self.w = w
        self.antenna_time = None
        if time_marginalization:
            self._setup_time_marginalization_multiband()

    def _setup_time_marginalization_multiband(self):
        """multiband.py:714-726."""
        w = self.w
        N = int(w["Nbs"][-1]) // 2
        self._delta_tc = w["durations"][0] / N
        self._times = self.event.start_time + np.arange(N) * self._delta_tc
        self._full_d_h = np.zeros(N, dtype=complex)
        self._full_to_multiband = [int(f * w["durations"][0]) for f in w["banded_frequency_points"]]
        prior = self.priors["geocent_time"]
        self.antenna_time = (prior.minimum + prior.maximum) / 2

    def calculate_snrs_mb(self, pols_unique, det, params):
        """multiband.py:750-765, 792-800."""
        w = self.w
        idx = w["unique_to_original_frequencies"]
        modes = {k: v[idx] for k, v in pols_unique.items()}
        at_points = types.SimpleNamespace(tensor=det.tensor, vertex=det.vertex,
                                          frequency_array=w["banded_frequency_points"],
                                          mask=np.ones(len(idx), dtype=bool))
        strain = detector_response(at_points, modes, params, self.event.start_time,
                                   antenna_time=self.antenna_time)
        lin = w["linear_coeffs"][det.name]
        d_inner_h = np.conj(np.dot(strain, lin))
        if w.get("linear_interpolation", True):
            h_inner_h = np.vdot(np.abs(strain) ** 2, w["quadratic_coeffs"][det.name]).real
        else:
            h_inner_h = self.ifft_fft_optimal_snr_squared(strain, det)
        arr = None
        if self.time_marginalization:
            fm = self._full_to_multiband
            self._full_d_h[fm] *= 0
            for b in range(len(w["durations"])):
                s, e = w["start_end_idxs"][b]
                self._full_d_h[fm[s:e + 1]] += strain[s:e + 1] * lin[s:e + 1]
            arr = np.fft.fft(self._full_d_h)
        return d_inner_h, h_inner_h, arr

    def ifft_fft_optimal_snr_squared(self, strain, det):
        """multiband.py:766-783: the first band summed directly on the
        original grid (4/T mask w / S); band b >= 1: sqrt(w) h placed at bins
        K_s .. K_e of an M_b / 2 + 1 half spectrum, its inverse real
        transform written into the last M_b samples of N^_b zeros,
        transformed back, |.|^2 weighted by Re Ibc_b, times 4 / T^_b."""
        w = self.w
        total = 0.0
        for b in range(len(w["durations"])):
            ks, ke = (int(v) for v in w["Ks_Ke"][b])
            s, e = (int(v) for v in w["start_end_idxs"][b])
            mb = int(w["Mbs"][b])
            seg = strain[s:e + 1]
            if b == 0:
                total += (4. / self.event.duration) * np.vdot(
                    np.abs(seg) ** 2,
                    det.mask[ks:ke + 1] * w["windows"][s:e + 1] / det.psd[ks:ke + 1]).real
            else:
                wth = np.zeros(mb // 2 + 1, dtype=complex)
                wth[ks:ke + 1] = w["square_root_windows"][s:e + 1] * seg
                hbc = np.zeros(min(2 * mb, int(w["Nbs"][b])))
                hbc[-mb:] = np.fft.irfft(wth)
                thbc = np.fft.rfft(hbc)
                ibc = w["Ibcs"][det.name][str(b)]
                total += (4. / w["Tbhats"][b]) * np.vdot(np.abs(thbc) ** 2, ibc.real).real
        return total

    def log_likelihood_ratio(self, parameters):
        params = copy.deepcopy(parameters)
        wa = dict(self.waveform_arguments)
        wa.setdefault("maximum_frequency", self.event.frequency_array[-1])
        pols = source_polarizations(self.source, self.w["waveform_frequencies"], params, wa)
        if self.time_marginalization and self.jitter_time:
            params["geocent_time"] += params["time_jitter"]
        d_inner_h, h_inner_h, arr = 0j, 0.0, None
        for det in self.event.detectors:
            dh, hh, a = self.calculate_snrs_mb(pols, det, params)
            d_inner_h += dh
            h_inner_h += hh
            if a is not None:
                arr = a if arr is None else arr + a
        return float(np.real(self.compute_log_likelihood(d_inner_h, h_inner_h, arr, params)))


    def inner_products(self, parameters):
        """multiband.py:750-807 as base.py:686-694 calls it: <d|h>, <h|h> at
        the parameters' own time, summed over detectors."""
        params = copy.deepcopy(parameters)
        wa = dict(self.waveform_arguments)
        wa.setdefault("maximum_frequency", self.event.frequency_array[-1])
        pols = source_polarizations(self.source, self.w["waveform_frequencies"], params, wa)
        d_inner_h, h_inner_h = 0j, 0.0
        for det in self.event.detectors:
            dh, hh, _ = self.calculate_snrs_mb(pols, det, params)
            d_inner_h += dh
            h_inner_h += hh
        return d_inner_h, h_inner_h

    def time_posterior_on_grid(self, parameters):
        """multiband.py:813-848: d_inner_h_array of the waveform at the
        parameters' geocent_time (no jitter added), the time-conditioned lnL
        (distance / phase marginalised as configured) on _times plus the
        jitter, times the prior, normalised (the reference then draws with
        numpy's choice)."""
        params = copy.deepcopy(parameters)
        wa = dict(self.waveform_arguments)
        wa.setdefault("maximum_frequency", self.event.frequency_array[-1])
        pols = source_polarizations(self.source, self.w["waveform_frequencies"], params, wa)
        h_inner_h, arr = 0.0, None
        for det in self.event.detectors:
            _, hh, a = self.calculate_snrs_mb(pols, det, params)
            h_inner_h += hh
            arr = a.copy() if arr is None else arr + a
        if self.distance_marginalization:
            log_like = self.distance_marginalized_likelihood(arr, h_inner_h, params)
        elif self.phase_marginalization:
            log_like = core.ln_i0(np.abs(arr)) - h_inner_h / 2
        else:
            log_like = arr.real - h_inner_h / 2
        times = self._times
        if self.jitter_time:
            times = times + params["time_jitter"]
        post = np.exp(log_like - np.max(log_like)) * self.priors["geocent_time"].prob(times)
        return times, post / np.sum(post)


def quadratic_coefficients_loop(setup, ifos):
    """multiband.py:551-611 loop for loop: ``setup`` holds durations, fb_dfb,
    start_end_idxs, banded_frequency_points and a _window(delta_f, start,
    length, b) method; ``ifos`` as bilby's interferometers."""
    T = float(setup.durations[0])
    out = {ifo.name: [] for ifo in ifos}
    for b in range(len(setup.durations)):
        s, e = setup.start_end_idxs[b]
        fb = setup.banded_frequency_points[s:e + 1]
        prefactor = 4 * setup.durations[b] / T
        fnow, dfnow = setup.fb_dfb[b]
        fnext, _ = setup.fb_dfb[b + 1]
        lo = math.ceil((fnow - dfnow) * T)
        window = setup._window(1 / T, lo, math.floor(fnext * T) - lo + 1, b)
        for ifo in ifos:
            psd = ifo.power_spectral_density_array
            hi = min(lo + len(window) - 1, len(psd) - 1)
            mask = np.asarray(ifo.frequency_mask[lo:hi + 1])
            wps = np.zeros(hi + 1 - lo)
            wps[mask] = 1. / psd[lo:hi + 1][mask]
            wps *= window[:len(wps)]
            coeffs = np.zeros(len(fb))
            for k in range(len(coeffs) - 1):
                a = lo if k == 0 else max(lo, math.ceil(T * fb[k]))
                z = hi if k == len(coeffs) - 2 else min(hi, math.ceil(T * fb[k + 1]) - 1)
                f = np.arange(a, z + 1) / T
                seg = wps[a - lo:z - lo + 1]
                coeffs[k] += prefactor * np.sum((fb[k + 1] - f) * seg)
                coeffs[k + 1] += prefactor * np.sum((f - fb[k]) * seg)
            out[ifo.name].append(coeffs)
    return {k: np.concatenate(v) for k, v in out.items()}
