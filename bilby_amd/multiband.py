"""Multibanded likelihood on the GPU: bilby's MBGravitationalWaveTransient
(bilby/gw/likelihood/multiband.py; Morisaki, arXiv:2104.07813) with either
algorithm for <h|h>: linear interpolation or IFFT-FFT.

``MultibandSetup`` restates the host-side construction (float64 numpy, no
GPU): frequency bands of durations T, T/2, T/4, ... chosen from the 0PN
time-to-merger of a reference chirp mass, the banded frequency points, the
linear coefficients (data filtered, down-sampled and windowed per band) and
the quadratic coefficients.  Each evaluation runs ``k_mb``
(csrc/gwl_relbin.hip): TaylorF2 at every banded point, the detector
response, <d|h> = conj(sum h L) and <h|h> = sum |h|^2 Q, then the
marginalisation.  With time marginalisation (multiband.py:714-726, 792-800)
the transform of the scattered h L terms at the time-prior window is a
fixed linear map of the per-point terms, evaluated as one GEMM per batch
(gwl_mb_set_time).  With linear_interpolation=False (multiband.py:613-646,
766-783) the per-point quadratic table holds the diagonal part of the
IFFT-FFT <h|h> (the first band's direct sum and the even bins of the other
bands' transforms) and gwl_mb_set_fft adds, per band, the odd bins from two
half-length transforms per row and detector on the GPU.
"""
import ctypes
import math

import numpy as np
import torch

from . import _lib
from .constants import gravitational_constant, radius_of_earth, solar_mass, speed_of_light
from ._lib import LIB, check, dptr
from .likelihood import GravitationalWaveTransient, grid_time_posterior


def round_up_to_power_of_two(x):
    """bilby/core/utils/calculus.py:302-313."""
    return 2 ** math.ceil(np.log2(x))


class MultibandSetup:
    """Bands and coefficients (multiband.py:311-647).

    ``interferometers``: objects with ``name``, ``frequency_array``,
    ``frequency_mask``, ``frequency_domain_strain``,
    ``power_spectral_density_array``, ``minimum_frequency`` and
    ``maximum_frequency``; ``duration`` is the data duration.
    """

    def __init__(self, interferometers, duration, reference_chirp_mass, highest_mode=2,
                 accuracy_factor=5, time_offset=2.12, delta_f_end=53.0,
                 maximum_banding_frequency=None, minimum_banding_duration=0.0,
                 linear_interpolation=True):
        if not isinstance(linear_interpolation, (bool, np.bool_)):  # multiband.py:175-180
            raise TypeError("linear_interpolation must be a bool")
        self.linear_interpolation = bool(linear_interpolation)
        self.interferometers = list(interferometers)
        self.duration = float(duration)
        self.reference_chirp_mass = float(reference_chirp_mass)
        self.highest_mode = highest_mode
        self.accuracy_factor = accuracy_factor
        self.time_offset = float(time_offset)
        self.delta_f_end = float(delta_f_end)
        self.minimum_banding_duration = minimum_banding_duration
        # multiband.py:264-287: where f - 1/sqrt(-dtau/df) stops increasing
        fmax = ((15 / 968) ** (3 / 5) * (self.highest_mode / (2 * np.pi)) ** (8 / 5)
                / self.reference_chirp_mass_in_second)
        if maximum_banding_frequency is not None and maximum_banding_frequency < fmax:
            fmax = maximum_banding_frequency
        self.maximum_banding_frequency = fmax
        self._setup_frequency_bands()
        self._setup_integers()
        self._setup_waveform_frequency_points()
        self._setup_linear_coefficients()
        if self.linear_interpolation:
            self._setup_quadratic_coefficients()
        else:
            self._setup_quadratic_coefficients_ifft_fft()

    @property
    def reference_chirp_mass_in_second(self):
        return gravitational_constant * self.reference_chirp_mass * solar_mass / speed_of_light ** 3

    @property
    def minimum_frequency(self):
        return min(i.minimum_frequency for i in self.interferometers)

    @property
    def maximum_frequency(self):
        return max(i.maximum_frequency for i in self.interferometers)

    @property
    def number_of_bands(self):
        return len(self.durations)

    def _tau(self, f):
        """0PN time to merger of the reference chirp mass at frequency f of
        the highest mode (multiband.py:322-340)."""
        mc = self.reference_chirp_mass_in_second
        return 5 / 256 * mc * (np.pi * mc * (2 * f / self.highest_mode)) ** (-8 / 3)

    def _dtaudf(self, f):
        mc = self.reference_chirp_mass_in_second
        return -5 / 96 * mc * (np.pi * mc * (2 * f / self.highest_mode)) ** (-8 / 3) / f

    def _find_starting_frequency(self, duration, fnow):
        """Bisection for the lowest frequency at which a band of ``duration``
        holds the signal (Eqs. 10 and 51 of arXiv:2104.07813;
        multiband.py:362-400)."""
        def above(f):
            width = np.sqrt(-self._dtaudf(f))
            return ((duration - self.time_offset - self._tau(f) - self.accuracy_factor * width > 0)
                    and (f - 1.0 / width - fnow > 0))
        lo, hi = fnow, self.maximum_banding_frequency
        if not above(hi):
            return None, None
        f = hi
        while hi - lo > 1e-2 / duration:
            f = (lo + hi) / 2.0
            if above(f):
                hi = f
            else:
                lo = f
        return f, 1.0 / np.sqrt(-self._dtaudf(f))

    def _setup_frequency_bands(self):
        """multiband.py:402-426."""
        durations = [self.duration]
        fb_dfb = [[self.minimum_frequency, 0.0]]
        dnext = self.duration / 2
        while dnext > max(self.time_offset, self.minimum_banding_duration):
            fnext, dfnext = self._find_starting_frequency(dnext, fb_dfb[-1][0])
            if fnext is not None and fnext < min(self.maximum_frequency,
                                                 self.maximum_banding_frequency):
                durations.append(dnext)
                fb_dfb.append([fnext, dfnext])
                dnext /= 2
            else:
                break
        fb_dfb.append([self.maximum_frequency + self.delta_f_end, self.delta_f_end])
        self.durations = np.array(durations)
        self.fb_dfb = np.array(fb_dfb)

    def _setup_integers(self):
        """multiband.py:428-447: N_b (down-sampled length), M_b = N_b / 2^b
        (shortened length), [K_s, K_e] frequency index range of each band."""
        nbs, mbs, ks_ke = [], [], []
        for b in range(self.number_of_bands):
            dnow = self.durations[b]
            fnow, dfnow = self.fb_dfb[b]
            fnext = self.fb_dfb[b + 1][0]
            nb = max(round_up_to_power_of_two(2.0 * (fnext * self.duration + 1.0)), 2 ** b)
            nbs.append(nb)
            mbs.append(nb // 2 ** b)
            ks_ke.append([math.ceil((fnow - dfnow) * dnow), math.floor(fnext * dnow)])
        self.Nbs = np.array(nbs, dtype=int)
        self.Mbs = np.array(mbs, dtype=int)
        self.Ks_Ke = np.array(ks_ke, dtype=int)

    def _setup_waveform_frequency_points(self):
        """multiband.py:449-478."""
        pts, se = [], []
        start = 0
        for b in range(self.number_of_bands):
            ks, ke = self.Ks_Ke[b]
            pts.append(np.arange(ks, ke + 1) / self.durations[b])
            se.append([start, start + ke - ks])
            start += ke - ks + 1
        self.banded_frequency_points = np.concatenate(pts)
        self.start_end_idxs = np.array(se, dtype=int)
        self.waveform_frequencies, self.unique_to_original_frequencies = np.unique(
            self.banded_frequency_points, return_inverse=True)

    def _window(self, delta_f, start_idx, length, b):
        """multiband.py:480-527: 1 inside band b, raised-cosine tapers of
        widths dfnow / dfnext at its ends, on frequencies (start_idx + i)
        delta_f."""
        fnow, dfnow = self.fb_dfb[b]
        fnext, dfnext = self.fb_dfb[b + 1]
        w = np.zeros(length)
        inc0 = int(np.clip(math.floor((fnow - dfnow) / delta_f) - start_idx + 1, 0, length))
        one0 = int(np.clip(math.ceil(fnow / delta_f) - start_idx, 0, length))
        dec0 = int(np.clip(math.floor((fnext - dfnext) / delta_f) - start_idx + 1, 0, length))
        dec1 = int(np.clip(math.ceil(fnext / delta_f) - start_idx, 0, length))
        w[one0:dec0] = 1.0
        if inc0 < one0:
            f = (np.arange(inc0, one0) + start_idx) * delta_f
            w[inc0:one0] = (1.0 + np.cos(np.pi * (f - fnow) / dfnow)) / 2.0
        if dec0 < dec1:
            f = (np.arange(dec0, dec1) + start_idx) * delta_f
            w[dec0:dec1] = (1.0 - np.cos(np.pi * (f - fnext) / dfnext)) / 2.0
        return w

    def _setup_linear_coefficients(self):
        """multiband.py:529-549: per band, the whitened data d/S truncated at
        the band's Nyquist frequency, transformed to the time domain, cut to
        its last M_b samples and back, times 4/T_b and the window."""
        self.linear_coeffs = {}
        n_full = self.Nbs[-1]
        for ifo in self.interferometers:
            fd = np.zeros(n_full // 2 + 1, dtype=complex)
            mask = np.asarray(ifo.frequency_mask)
            m = mask[:len(fd)]
            fd[:len(ifo.frequency_domain_strain)][m] += np.asarray(
                ifo.frequency_domain_strain[mask] / ifo.power_spectral_density_array[mask])
            parts = []
            for b in range(self.number_of_bands):
                ks, ke = self.Ks_Ke[b]
                window = self._window(1.0 / self.durations[b], ks, ke - ks + 1, b)
                band = np.copy(fd[:int(self.Nbs[b] / 2 + 1)])
                band[-1] = 0.0
                td = np.fft.irfft(band)[-self.Mbs[b]:]
                shortened = np.fft.rfft(td)[ks:ke + 1]
                parts.append((4.0 / self.durations[b]) * window * np.conj(shortened))
            self.linear_coeffs[ifo.name] = np.concatenate(parts)

    def _setup_quadratic_coefficients(self):
        """multiband.py:551-611: |h|^2 linearly interpolated between banded
        points; each original-grid term (window / S) is split between the two
        banded points around it in proportion to its distance from each.
        The grid terms between consecutive banded points are summed per
        interval (np.bincount over the interval index)."""
        T = self.duration
        self.quadratic_coeffs = {ifo.name: [] for ifo in self.interferometers}
        for b in range(self.number_of_bands):
            s, e = self.start_end_idxs[b]
            fb = self.banded_frequency_points[s:e + 1]
            prefactor = 4 * self.durations[b] / T
            fnow, dfnow = self.fb_dfb[b]
            fnext = self.fb_dfb[b + 1][0]
            lo = math.ceil((fnow - dfnow) * T)
            window = self._window(1 / T, lo, math.floor(fnext * T) - lo + 1, b)
            n = len(fb)
            for ifo in self.interferometers:
                psd = ifo.power_spectral_density_array
                hi = min(lo + len(window) - 1, len(psd) - 1)
                mask = np.asarray(ifo.frequency_mask[lo:hi + 1])
                wps = np.zeros(hi + 1 - lo)
                wps[mask] = 1.0 / psd[lo:hi + 1][mask]
                wps *= window[:len(wps)]
                f = np.arange(lo, hi + 1) / T
                start = np.maximum(lo, np.ceil(T * fb[:-1]).astype(int))
                start[0] = lo
                stop = np.minimum(hi, np.ceil(T * fb[1:]).astype(int) - 1)
                stop[-1] = hi
                a = np.clip(start - lo, 0, len(wps))
                z = np.maximum(np.clip(stop - lo + 1, 0, len(wps)), a)
                # the intervals tile the grid terms: each term belongs to one k
                owner = np.repeat(np.arange(n - 1), z - a)
                if len(owner) != len(wps) or not np.array_equal(a[1:], z[:-1]):
                    raise RuntimeError("multiband: interpolation intervals do not tile band %d" % b)
                coeffs = np.zeros(n)
                coeffs[:-1] += prefactor * np.bincount(owner, weights=(fb[owner + 1] - f) * wps,
                                                       minlength=n - 1)
                coeffs[1:] += prefactor * np.bincount(owner, weights=(f - fb[owner]) * wps,
                                                      minlength=n - 1)
                self.quadratic_coeffs[ifo.name].append(coeffs)
        for name in self.quadratic_coeffs:
            self.quadratic_coeffs[name] = np.concatenate(self.quadratic_coeffs[name])

    def _setup_quadratic_coefficients_ifft_fft(self):
        """multiband.py:613-646: for each band the inverse PSD (zero off the
        mask) truncated at the band's Nyquist bin and taken to the time
        domain (N_b samples), kept at lags |n| <= N^_b / 2 with
        N^_b = min(2 M_b, N_b), and transformed back (Ibc); T^_b = T N^_b / N_b;
        the windows and their square roots on the banded points; zeroed work
        arrays wths / hbcs as the reference keeps them."""
        n_full = int(self.Nbs[-1])
        nhat = [int(min(2 * m, n)) for m, n in zip(self.Mbs, self.Nbs)]
        self.Tbhats = np.array([self.duration * nh / n for nh, n in zip(nhat, self.Nbs)])
        self.Ibcs, self.hbcs, self.wths = {}, {}, {}
        for ifo in self.interferometers:
            inv = np.zeros(n_full // 2 + 1)
            psd = np.asarray(ifo.power_spectral_density_array)
            mask = np.asarray(ifo.frequency_mask, dtype=bool)
            inv[:len(psd)][mask[:len(inv)]] = 1.0 / psd[mask]
            ibc, hbc, wth = [], [], []
            for b in range(self.number_of_bands):
                lag = np.fft.irfft(inv[:int(self.Nbs[b]) // 2 + 1])
                half = nhat[b] // 2
                tail = nhat[b] - half - 1
                kept = np.concatenate([lag[:half + 1], lag[len(lag) - tail:] if tail else []])
                ibc.append(np.fft.rfft(kept))
                hbc.append(np.zeros(nhat[b]))
                wth.append(np.zeros(int(self.Mbs[b]) // 2 + 1, dtype=complex))
            self.Ibcs[ifo.name], self.hbcs[ifo.name], self.wths[ifo.name] = ibc, hbc, wth
        win = [self._window(1.0 / self.durations[b], int(self.Ks_Ke[b][0]),
                            int(self.Ks_Ke[b][1] - self.Ks_Ke[b][0]) + 1, b)
               for b in range(self.number_of_bands)]
        self.windows = np.concatenate(win)
        self.square_root_windows = np.sqrt(self.windows)

    def ifft_fft_tables(self):
        """The device form of the IFFT-FFT <h|h> (multiband.py:766-783):
        per detector the diagonal coefficients on the banded points (the
        first band's 4/T mask w / S, and (4 / T^_b) Re Ibc_b[2 j] w_j at bin j
        of band b >= 1: the even bins of that band's 2 M_b-point transform
        are its half spectrum sqrt(w) h), and per band b >= 1 the odd-bin
        weights (4 / T^_b) Re Ibc_b[2 j + 1], j < M_b / 2."""
        T = self.duration
        diag = {}
        for ifo in self.interferometers:
            q = np.zeros(len(self.banded_frequency_points))
            psd = np.asarray(ifo.power_spectral_density_array)
            mask = np.asarray(ifo.frequency_mask, dtype=bool)
            for b in range(self.number_of_bands):
                s, e = (int(v) for v in self.start_end_idxs[b])
                ks, ke = (int(v) for v in self.Ks_Ke[b])
                w = self.windows[s:e + 1]
                if b == 0:
                    m = mask[ks:ke + 1]
                    q[s:e + 1] = np.where(m, 4.0 / T * w / np.where(m, psd[ks:ke + 1], 1.0), 0.0)
                else:
                    q[s:e + 1] = 4.0 / self.Tbhats[b] * self.Ibcs[ifo.name][b][2 * ks:2 * ke + 1:2].real * w
            diag[ifo.name] = q
        bands = []
        for b in range(1, self.number_of_bands):
            s, e = (int(v) for v in self.start_end_idxs[b])
            ks, ke = (int(v) for v in self.Ks_Ke[b])
            mh = int(self.Mbs[b]) // 2
            if ks < 1 or ke >= mh or mh < 16:
                raise NotImplementedError("IFFT-FFT band %d reaches the zero or Nyquist bin of its "
                                          "transform" % b)
            odd = np.stack([4.0 / self.Tbhats[b] * self.Ibcs[ifo.name][b][1:2 * mh:2].real
                            for ifo in self.interferometers])
            bands.append(dict(half_length=mh, first_point=s, n_points=e - s + 1, first_bin=ks,
                              sqrt_window=np.ascontiguousarray(self.square_root_windows[s:e + 1]),
                              odd_weights=np.ascontiguousarray(odd)))
        return diag, bands

    @property
    def weights(self):
        """multiband.py:648-671."""
        keys = ["reference_chirp_mass", "highest_mode", "accuracy_factor", "time_offset",
                "delta_f_end", "maximum_banding_frequency", "minimum_banding_duration",
                "durations", "fb_dfb", "Nbs", "Mbs", "Ks_Ke", "banded_frequency_points",
                "start_end_idxs", "unique_to_original_frequencies"]
        out = {k: getattr(self, k) for k in keys}
        out["linear_interpolation"] = self.linear_interpolation
        out["waveform_frequencies"] = self.waveform_frequencies
        out["linear_coeffs"] = dict(self.linear_coeffs)
        if self.linear_interpolation:
            out["quadratic_coeffs"] = dict(self.quadratic_coeffs)
        else:
            for k in ("Tbhats", "windows", "square_root_windows"):
                out[k] = getattr(self, k)
            for k in ("wths", "hbcs", "Ibcs"):
                out[k] = {name: {str(b): v for b, v in enumerate(arrs)}
                          for name, arrs in getattr(self, k).items()}
        return out

    def save_weights(self, filename):
        """multiband.py:673-687, as an .npz archive (bilby writes HDF5);
        nested dictionaries become keys joined with '/'."""
        flat = {}

        def put(prefix, v):
            if isinstance(v, dict):
                for name, sub in v.items():
                    put("%s/%s" % (prefix, name), sub)
            else:
                flat[prefix] = np.asarray(v)

        for k, v in self.weights.items():
            put(k, v)
        np.savez(filename, **flat)

    @classmethod
    def from_weights(cls, weights, interferometers):
        """multiband.py:689-712; ``weights`` a dict or an .npz file name."""
        if isinstance(weights, str):
            with np.load(weights, allow_pickle=False) as f:
                flat = {k: f[k] for k in f.files}
            weights = {}
            for k, v in flat.items():
                *path, leaf = k.split("/")
                node = weights
                for part in path:
                    node = node.setdefault(part, {})
                node[leaf] = v[()] if v.ndim == 0 else v
        self = cls.__new__(cls)
        self.interferometers = list(interferometers)
        self.duration = float(weights["durations"][0])
        self.linear_interpolation = bool(weights.get("linear_interpolation", True))
        for k, v in weights.items():
            if k == "linear_interpolation":
                continue
            if k in ("wths", "hbcs", "Ibcs"):  # multiband.py:704-708: per band lists
                v = {name: [d[str(b)] for b in range(len(d))] for name, d in v.items()}
            setattr(self, k, dict(v) if isinstance(v, dict) else v)
        return self


class MBGravitationalWaveTransient(GravitationalWaveTransient):
    """multiband.py:19-127 constructor semantics (reference chirp mass from
    the chirp-mass prior minimum when not given; time offset and high-end
    taper from the geocent_time prior, multiband.py:194-259), plus
    ``precision`` and ``device``."""

    def __init__(self, interferometers, waveform_generator, reference_chirp_mass=None,
                 highest_mode=2, linear_interpolation=True, accuracy_factor=5, time_offset=None,
                 delta_f_end=None, maximum_banding_frequency=None, minimum_banding_duration=0.0,
                 weights=None, distance_marginalization=False, phase_marginalization=False,
                 priors=None, time_marginalization=False, jitter_time=True,
                 distance_marginalization_lookup_table=None, reference_frame="sky",
                 time_reference="geocenter", precision="fp64", device=0):
        self._mb = None
        super().__init__(interferometers, waveform_generator,
                         time_marginalization=time_marginalization,
                         distance_marginalization=distance_marginalization,
                         phase_marginalization=phase_marginalization, priors=priors,
                         distance_marginalization_lookup_table=distance_marginalization_lookup_table,
                         jitter_time=jitter_time, reference_frame=reference_frame,
                         time_reference=time_reference, precision=precision, device=device)
        if self.waveform_generator.approximant != _lib.TAYLORF2:
            raise NotImplementedError("multibanding is provided for TaylorF2 only")
        ifos = self.interferometers
        if weights is not None:
            self.setup = MultibandSetup.from_weights(weights, ifos)
        else:
            safety = radius_of_earth / speed_of_light
            p = self.priors
            if time_offset is None:
                time_offset = (ifos.start_time + ifos.duration - p["geocent_time"].minimum + safety
                               if p is not None and "geocent_time" in p
                               and hasattr(p["geocent_time"], "minimum") else 2.12)
            if delta_f_end is None:
                delta_f_end = (100 / (ifos.start_time + ifos.duration - p["geocent_time"].maximum
                                      - safety)
                               if p is not None and "geocent_time" in p
                               and hasattr(p["geocent_time"], "maximum") else 53.0)
            if reference_chirp_mass is None:
                reference_chirp_mass = self._minimum_chirp_mass()
            self.setup = MultibandSetup(ifos, ifos.duration, reference_chirp_mass, highest_mode,
                                        accuracy_factor, time_offset, delta_f_end,
                                        maximum_banding_frequency, minimum_banding_duration,
                                        linear_interpolation)
        if self.time_marginalization:
            self._setup_time_marginalization_multiband()
        self._upload()

    def _setup_time_marginalization_multiband(self):
        """multiband.py:714-726: the time grid start + j T_0 / N, N = Nbs[-1] / 2,
        with prior weight spacing T_0 / N; the grid index int(f T_0) of each
        banded point; the antenna patterns at the middle of the geocent_time
        prior.  bilby sets that time as ``interferometers.reference_time``
        (which also changes later evaluations of other likelihoods sharing the
        interferometers); here it applies to this likelihood only.

        bilby fills the transformed array band by band with
        ``a[idx] += v`` (multiband.py:792-798): within one band a repeated
        index keeps only the last point's term, across bands the terms add.
        ``_time_grid_index`` is that map with -1 for a point whose term is
        dropped."""
        s = self.setup
        N = int(s.Nbs[-1]) // 2
        jitter_bound = self._delta_tc / 2
        self._delta_tc = float(s.durations[0]) / N
        self._times = self.interferometers.start_time + np.arange(N) * self._delta_tc
        prior = self.priors["geocent_time"]
        self.time_prior_array = prior.prob(self._times) * self._delta_tc
        self._full_to_multiband = [int(f * s.durations[0]) for f in s.banded_frequency_points]
        grid = np.asarray(self._full_to_multiband, dtype=np.int64)
        if len(grid) and grid.max() >= N:
            raise ValueError("a banded frequency point lies beyond the time-marginalisation grid")
        index = grid.copy()
        for b in range(s.number_of_bands):
            lo, hi = s.start_end_idxs[b]
            band = grid[lo:hi + 1]
            # the last occurrence of each index within the band survives
            _, last = np.unique(band[::-1], return_index=True)
            keep = np.zeros(len(band), dtype=bool)
            keep[len(band) - 1 - last] = True
            index[lo:hi + 1] = np.where(keep, band, -1)
        self._time_grid_index = np.ascontiguousarray(index)
        self._beam_pattern_reference_time = (prior.minimum + prior.maximum) / 2
        margin = max(jitter_bound, self._delta_tc)
        start = self.interferometers.start_time
        j_lo = int(np.floor((prior.minimum - start - margin) / self._delta_tc)) - 3
        j_hi = int(np.ceil((prior.maximum - start + margin) / self._delta_tc)) + 3
        j_lo, j_hi = max(0, j_lo), min(N - 1, j_hi)
        if j_hi < j_lo:
            raise ValueError("the geocent_time prior lies outside the segment")
        self._time_window = (N, j_lo, j_hi - j_lo + 1)

    def _minimum_chirp_mass(self):
        """gw/prior.py:721-737."""
        p = self.priors or {}
        if "chirp_mass" in p:
            return float(p["chirp_mass"].minimum)
        if "mass_1" in p and ("mass_2" in p or "mass_ratio" in p):
            m1 = p["mass_1"].minimum
            m2 = p["mass_2"].minimum if "mass_2" in p else m1 * p["mass_ratio"].minimum
            return float((m1 * m2) ** 0.6 / (m1 + m2) ** 0.2)
        raise ValueError("reference_chirp_mass is not given and the priors do not fix a minimum "
                         "chirp mass")

    # multiband.py:128-309: the banding settings, read from the set-up
    reference_chirp_mass = property(lambda self: self.setup.reference_chirp_mass)
    highest_mode = property(lambda self: self.setup.highest_mode)
    accuracy_factor = property(lambda self: self.setup.accuracy_factor)
    time_offset = property(lambda self: self.setup.time_offset)
    delta_f_end = property(lambda self: self.setup.delta_f_end)
    maximum_banding_frequency = property(lambda self: self.setup.maximum_banding_frequency)
    minimum_banding_duration = property(lambda self: self.setup.minimum_banding_duration)
    linear_interpolation = property(lambda self: self.setup.linear_interpolation)

    @property
    def weights(self):
        return self.setup.weights

    def save_weights(self, filename):
        self.setup.save_weights(filename)

    def _upload(self):
        s = self.setup
        nd, n = len(self.interferometers), len(s.banded_frequency_points)
        lin = np.zeros((nd, n, 2))
        quad = np.zeros((nd, n))
        fft_bands = []
        if s.linear_interpolation:
            diag = s.quadratic_coeffs
        else:
            diag, fft_bands = s.ifft_fft_tables()
        for k, ifo in enumerate(self.interferometers):
            lin[k, :, 0] = np.real(s.linear_coeffs[ifo.name])
            lin[k, :, 1] = np.imag(s.linear_coeffs[ifo.name])
            quad[k] = diag[ifo.name]
        freqs = np.ascontiguousarray(s.banded_frequency_points, dtype=np.float64)
        lin = np.ascontiguousarray(lin)
        quad = np.ascontiguousarray(quad)
        mb = ctypes.c_void_p()
        check(LIB.gwl_mb_create(self._event, n, dptr(freqs), dptr(lin), dptr(quad),
                                ctypes.byref(mb)), "gwl_mb_create")
        if fft_bands:
            descs = (_lib.MbFftBand * len(fft_bands))()
            for k, band in enumerate(fft_bands):
                for key in ("half_length", "first_point", "n_points", "first_bin"):
                    setattr(descs[k], key, band[key])
                descs[k].sqrt_window = dptr(band["sqrt_window"])
                descs[k].odd_weights = dptr(band["odd_weights"])
            rc = LIB.gwl_mb_set_fft(mb, len(fft_bands), descs)
            if rc != 0:
                LIB.gwl_mb_destroy(mb)
                check(rc, "gwl_mb_set_fft")
        if self.time_marginalization:
            N, j_first, n_window = self._time_window
            td = _lib.MbTimeDesc()
            td.grid_index = self._time_grid_index.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
            td.n_fft = N
            td.j_first = j_first
            td.n_window = n_window
            td.t_step = self._delta_tc
            td.antenna_time = self._beam_pattern_reference_time
            rc = LIB.gwl_mb_set_time(mb, ctypes.byref(td))
            if rc != 0:
                LIB.gwl_mb_destroy(mb)
                check(rc, "gwl_mb_set_time")
        self._mb = mb

    def _plain_inner_products(self, parameters):
        """<d|h>, <h|h> of this likelihood's approximation at the parameters'
        own time (its calculate_snrs, as the reconstruction of distance and
        phase uses them, base.py:642-742)."""
        q = dict(parameters)
        q["time_jitter"] = 0.0
        _, aux = self.log_likelihood_ratio_batch(self.rows_from_parameters(q)[:1],
                                                 return_inner_products=True)
        return complex(aux[0, 0], aux[0, 1]), float(aux[0, 2])

    def _time_posterior(self, parameters):
        """multiband.py:813-848: d_inner_h_array of the waveform at the
        sample's geocent_time (the jitter is not added to it) over the time
        window, the time-conditioned lnL on _times plus the jitter, times
        the prior, normalised.  Outside the window the prior is zero."""
        q = dict(parameters)
        for key, value in self._defaults().items():
            q.setdefault(key, value)
        jitter = float(q.get("time_jitter", 0.0)) if self.jitter_time else 0.0
        q["time_jitter"] = 0.0
        rows = torch.from_numpy(self.rows_from_parameters(q)[:1]).to("cuda:%d" % self.device)
        N, j_first, L = self._time_window
        dh = torch.empty((1, L, 2), dtype=torch.float64, device=rows.device)
        hh = torch.empty(1, dtype=torch.float64, device=rows.device)
        st = torch.cuda.current_stream(rows.device).cuda_stream
        check(LIB.gwl_mb_time_series(self._mb, ctypes.c_void_p(rows.data_ptr()), 1, _lib.NPARAM,
                                     _lib.FP64, ctypes.c_void_p(dh.data_ptr()),
                                     ctypes.c_void_p(hh.data_ptr()), ctypes.c_void_p(st)),
              "gwl_mb_time_series")
        dh = dh.cpu().numpy()[0]
        times = self._times[j_first:j_first + L] + jitter
        return times, grid_time_posterior(self, dh[:, 0] + 1j * dh[:, 1],
                                          float(hh.cpu().numpy()[0]), self._frame(q), times)

    def generate_time_sample_from_marginalized_likelihood(self, parameters, rng=None):
        """multiband.py:813-848: one of the grid times, drawn with the
        posterior's probabilities."""
        times, post = self._time_posterior(parameters)
        return float((np.random.default_rng() if rng is None else rng).choice(times, p=post))

    def _device_eval(self, rows_t, precision, aux=False, stream=None):
        n = rows_t.shape[0]
        dev = rows_t.device
        out = torch.empty(n, dtype=torch.float64, device=dev)
        aux_t = torch.empty((n, 3), dtype=torch.float64, device=dev) if aux else None
        st = torch.cuda.current_stream(dev).cuda_stream if stream is None else stream
        check(LIB.gwl_mb_eval(self._mb, ctypes.c_void_p(rows_t.data_ptr()), n, _lib.NPARAM,
                              precision, ctypes.c_void_p(out.data_ptr()),
                              ctypes.c_void_p(aux_t.data_ptr()) if aux else None,
                              ctypes.c_void_p(st)), "gwl_mb_eval")
        return out, aux_t

    def __del__(self):
        try:
            if getattr(self, "_mb", None) is not None:
                LIB.gwl_mb_destroy(self._mb)
                self._mb = None
        except Exception:
            pass
        super().__del__()
