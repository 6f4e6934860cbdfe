"""FP64 restatement of bilby's RelativeBinningGravitationalWaveTransient
(bilby/gw/likelihood/relative.py), with time marginalisation through the
rebuilt full waveform and its FFT (:376-391, 410-420).

Test infrastructure only (see oracle/__init__.py).  Written bin by bin, as
the reference loops, so that it stays independent of the vectorised setup in
bilby_amd.relative.
"""
import copy
import types

import numpy as np

from . import core
from .likelihood import OracleGWTransient, detector_response, source_polarizations


class OracleRelativeBinning(OracleGWTransient):
    """relative.py:102-174 (constructor), :177-231 (setup_bins),
    :233-258 (set_fiducial_waveforms), :314-360 (compute_summary_data),
    :362-374 (ratios), :376-391 (full waveform), :393-428 (calculate_snrs)."""

    def __init__(self, event, source, fiducial_parameters, waveform_arguments=None,
                 distance_marginalization=False, phase_marginalization=False, priors=None,
                 chi=1, epsilon=0.5, distance_table=None, time_marginalization=False,
                 jitter_time=True):
        super().__init__(event, source, waveform_arguments,
                         time_marginalization=time_marginalization,
                         distance_marginalization=distance_marginalization,
                         phase_marginalization=phase_marginalization, priors=priors,
                         jitter_time=jitter_time, distance_table=distance_table)
        fid = dict(fiducial_parameters)
        if time_marginalization:
            fid["geocent_time"] = event.start_time
        if distance_marginalization:
            fid["luminosity_distance"] = self._ref_dist
        if phase_marginalization:
            fid["phase"] = 0.0
        self.fiducial_parameters = fid
        self.chi = chi
        self.epsilon = epsilon
        self.gamma = np.array([-5 / 3, -2 / 3, 1, 5 / 3, 7 / 3])
        self.maximum_frequency = event.frequency_array[-1]
        self.set_fiducial_waveforms(fid)
        self.setup_bins()
        self.compute_summary_data()

    def set_fiducial_waveforms(self, parameters):
        """relative.py:233-258."""
        f = self.event.frequency_array
        pols = source_polarizations(self.source, f, parameters, self.waveform_arguments)
        last = np.where(pols["plus"] != 0j)[0][-1]
        self.maximum_frequency = f[last]
        self.fiducial = {}
        for det in self.event.detectors:
            wf = detector_response(det, pols, parameters, self.event.start_time)
            wf = wf * (det.frequency_array <= self.maximum_frequency)
            self.fiducial[det.name] = wf

    def setup_bins(self):
        """relative.py:177-231, loop for loop."""
        f = self.event.frequency_array
        gamma = self.gamma[:, np.newaxis]
        maximum_frequency = f[0]
        minimum_frequency = f[-1]
        for det in self.event.detectors:
            mf = det.frequency_array[det.mask]
            maximum_frequency = max(maximum_frequency, mf[-1])
            minimum_frequency = min(minimum_frequency, mf[0])
        maximum_frequency = min(maximum_frequency, self.maximum_frequency)
        useful = f[(f >= minimum_frequency) & (f <= maximum_frequency)]
        d_alpha = self.chi * 2 * np.pi / np.abs(
            (minimum_frequency ** gamma) * np.heaviside(-gamma, 1)
            - (maximum_frequency ** gamma) * np.heaviside(gamma, 1))
        d_phi = np.sum(np.sign(gamma) * d_alpha * useful ** gamma, axis=0)
        d_phi_from_start = d_phi - d_phi[0]
        number_of_bins = int(d_phi_from_start[-1] // self.epsilon)
        bin_inds, bin_freqs = [], []
        last_index = -1
        for i in range(number_of_bins + 1):
            idx = np.where(d_phi_from_start >= ((i / number_of_bins) * d_phi_from_start[-1]))[0][0]
            if idx == last_index:
                continue
            last_index = idx
            bin_freqs.append(useful[idx])
            bin_inds.append(np.where(f >= useful[idx])[0][0])
        self.bin_inds = np.array(bin_inds, dtype=int)
        self.bin_freqs = np.array(bin_freqs)
        self.number_of_bins = len(self.bin_inds) - 1
        self.bin_widths = self.bin_freqs[1:] - self.bin_freqs[:-1]
        self.fiducial_points = {name: wf[self.bin_inds] for name, wf in self.fiducial.items()}

    def compute_summary_data(self):
        """relative.py:314-360."""
        T = self.event.duration
        self.summary_data = {}
        for det in self.event.detectors:
            m = det.mask
            mf = det.frequency_array[m]
            inds = [int(np.where(mf == edge)[0][0]) for edge in self.bin_freqs]
            if inds[-1] < len(mf) - 1:
                inds[-1] += 1
            strain, h0, psd = det.strain[m], self.fiducial[det.name][m], det.psd[m]
            a0, b0, a1, b1 = np.zeros((4, self.number_of_bins), dtype=complex)
            for i in range(self.number_of_bins):
                s, e = inds[i], inds[i + 1]
                centre = (mf[s] + mf[e]) / 2
                sl = slice(s, e)
                delta = mf[sl] - centre
                a0[i] = core.noise_weighted_inner_product(h0[sl], strain[sl], psd[sl], T)
                b0[i] = core.noise_weighted_inner_product(h0[sl], h0[sl], psd[sl], T)
                a1[i] = core.noise_weighted_inner_product(h0[sl], strain[sl] * delta, psd[sl], T)
                b1[i] = core.noise_weighted_inner_product(h0[sl], h0[sl] * delta, psd[sl], T)
            self.summary_data[det.name] = (a0, a1, b0, b1)

    def edge_response(self, det, params):
        """The detector response at the bin edges (interferometer.py:303-368
        with ``frequencies`` given: no mask)."""
        pols = source_polarizations(self.source, self.bin_freqs, params, self.waveform_arguments)
        at_edges = types.SimpleNamespace(tensor=det.tensor, vertex=det.vertex,
                                         frequency_array=self.bin_freqs,
                                         mask=np.ones(len(self.bin_freqs), dtype=bool))
        return detector_response(at_edges, pols, params, self.event.start_time)

    def snrs_from_ratio(self, det, ratio):
        """relative.py:370-374, 400-404: <d|h>, <h|h> from the edge ratios."""
        r0 = (ratio[1:] + ratio[:-1]) / 2
        r1 = (ratio[1:] - ratio[:-1]) / self.bin_widths
        a0, a1, b0, b1 = self.summary_data[det.name]
        d_inner_h = (a0 * r0.conj() + a1 * r1.conj()).sum()
        h_inner_h = (b0 * abs(r0) ** 2 + 2 * b1 * (r0 * r1.conj()).real).sum()
        return d_inner_h, h_inner_h.real

    def full_waveform(self, det, ratio):
        """relative.py:376-391: the fiducial response times the ratio rebuilt
        bin by bin as r0 + r1 (f - f_centre); bin b covers grid bins
        [bin_inds[b], bin_inds[b + 1]), the last bin also bin_inds[-1]."""
        r0 = (ratio[1:] + ratio[:-1]) / 2
        r1 = (ratio[1:] - ratio[:-1]) / self.bin_widths
        f = det.frequency_array
        full_ratio = np.zeros(len(f), dtype=complex)
        nb = self.number_of_bins
        for b in range(nb):
            lo = self.bin_inds[b]
            hi = self.bin_inds[b + 1] + (1 if b == nb - 1 else 0)
            centre = (self.bin_freqs[b] + self.bin_freqs[b + 1]) / 2
            full_ratio[lo:hi] = r0[b] + r1[b] * (f[lo:hi] - centre)
        return self.fiducial[det.name] * full_ratio

    def time_arrays(self, params):
        """base.py:580-626 with relative.py:380-393 (_compute_full_waveform):
        the time array of the posterior reconstruction from the rebuilt full
        waveform, which relative binning inherits from the base likelihood:
        times on the 1/16384 s grid from the (jittered) geocent time modulo
        the segment inside [min, max) of the prior, the FFT of the
        zero-padded rebuilt h conj(d) / S (no 4/T), and <h|h> of the
        rebuilt waveform on the mask."""
        params = copy.deepcopy(params)
        params.pop("fiducial", None)
        ev = self.event
        if self.jitter_time:
            params["geocent_time"] += params["time_jitter"]
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
            ratio = self.edge_response(det, params) / self.fiducial_points[det.name]
            signal = self.full_waveform(det, ratio)
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

    def log_likelihood_ratio(self, parameters):
        """base.py:419-446 with relative.py:393-428 per detector."""
        params = copy.deepcopy(parameters)
        params.pop("fiducial", None)
        m1, m2 = params["mass_1"], params["mass_2"]
        if not (m1 > 0 and m2 > 0 and params["luminosity_distance"] > 0):
            return np.nan_to_num(-np.inf)
        if self.time_marginalization and self.jitter_time:
            params["geocent_time"] += params["time_jitter"]
        d_inner_h, h_inner_h, arr = 0j, 0.0, None
        for det in self.event.detectors:
            ratio = self.edge_response(det, params) / self.fiducial_points[det.name]
            dh, hh = self.snrs_from_ratio(det, ratio)
            d_inner_h += dh
            h_inner_h += hh
            if self.time_marginalization:
                full = self.full_waveform(det, ratio)
                with np.errstate(invalid="ignore", divide="ignore"):
                    integrand = full[0:-1] * det.strain.conj()[0:-1] / det.psd[0:-1]
                integrand = np.where(full[0:-1] != 0, integrand, 0)
                a = 4 / self.event.duration * np.fft.fft(integrand)
                arr = a if arr is None else arr + a
        return float(np.real(self.compute_log_likelihood(d_inner_h, h_inner_h, arr, params)))
