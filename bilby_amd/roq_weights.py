"""ROQ weights from reduced bases (bilby/gw/likelihood/roq.py:736-1005,
1165-1229) for ROQGravitationalWaveTransient.

For a linear basis B (L elements on the masked frequencies) and a quadratic
basis (Q elements), the weights dict holds the ROQ time samples, the
frequency nodes and per detector
  <ifo>_linear    [n_times, L]:  4/T sum_f (d / S) conj(B_l) e^{2 pi i f t}
  <ifo>_quadratic [Q]:           4/T sum_f Re(B_q) / S
The linear weights, bilby's dominant set-up cost (one inverse FFT of the
full segment per basis element and detector), are computed on the GPU
(gwl_roq_weights_linear: a float64 GEMM over frequency chunks against the
exponentials at the window of time samples).  The bases are given on each
detector's masked frequencies, or on the ROQ frequency grid of roq_params
(flow, fhigh, seglen, scaled by roq_scale_factor), whose intersection with
the masked frequencies the sums run over (roq.py:800-838); multibanded
bases are not supported.
"""
import ctypes

import numpy as np
from scipy.integrate import simpson

from .constants import radius_of_earth, speed_of_light
from ._lib import LIB, check, dptr


def time_resolution(interferometers, optimal_snrs=None):
    """roq.py:1165-1229: 1 / (5 f_high) with f_high = (pi^2 rho^2 / 6)^(1/3)
    times the PSD-weighted f^(2/3) moment (rho the network SNR, at least 10
    and by default 30 per detector), then rounded so that T / dt is a power of
    two no smaller than f_max T + 1."""
    snrs = optimal_snrs or {}
    rho2 = sum(max(10.0, snrs.get(ifo.name, 30.0)) ** 2 for ifo in interferometers)
    ifo = list(interferometers)[-1]
    m = ifo.frequency_mask
    freq = ifo.frequency_array[m]
    psd = ifo.power_spectral_density_array[m]
    i1 = simpson(y=freq ** (-7.0 / 3.0) / psd, x=freq)
    f3 = simpson(y=freq ** (2.0 / 3.0) / (psd * i1), x=freq)
    fhigh = (np.pi ** 2 * rho2 / 6) ** (1 / 3) * f3 ** (1 / 3)
    duration = interferometers.duration
    n = max(duration / (1.0 / fhigh / 5), interferometers.frequency_array[-1] * duration + 1)
    n = int(2 ** np.ceil(np.log2(n)))
    return duration / n


def time_samples(interferometers, time_prior_minimum, time_prior_maximum, delta_t):
    """roq.py:746-766: ROQ times (seconds after the segment start) covering
    the time prior padded by the Earth's light-crossing time and 5 samples.
    Returns (times, first sample index, number of samples per segment)."""
    n = int(interferometers.duration / delta_t)
    pad = 2 * radius_of_earth / speed_of_light + 5 * delta_t
    start = interferometers.start_time
    lo = max(0, int(np.floor((time_prior_minimum - pad - start) / delta_t)))
    hi = min(n - 1, int(np.ceil((time_prior_maximum + pad - start) / delta_t)))
    return np.arange(lo, hi + 1) * float(delta_t), lo, n


def roq_frequency_indices(interferometers, roq_params, roq_scale_factor=1.0):
    """roq.py:800-838: per detector, the indices into the ROQ frequency grid
    (create_frequency_series(2 fhigh s, seglen / s) from flow s, s the scale
    factor) and into the detector's masked frequencies of the frequencies
    both hold.  Returns ({name: roq_idxs}, {name: ifo_idxs})."""
    from .detector import create_frequency_series
    s = float(roq_scale_factor)
    fmin = float(roq_params["flow"]) * s
    fmax = float(roq_params["fhigh"]) * s
    seglen = float(roq_params["seglen"]) / s
    f_roq = create_frequency_series(sampling_frequency=fmax * 2, duration=seglen)
    f_roq = f_roq[f_roq >= fmin]
    roq_idxs, ifo_idxs = {}, {}
    for ifo in interferometers:
        _, i_ifo, i_roq = np.intersect1d(np.asarray(ifo.frequency_array[ifo.frequency_mask]),
                                         f_roq, return_indices=True)
        roq_idxs[ifo.name] = i_roq
        ifo_idxs[ifo.name] = i_ifo
    return roq_idxs, ifo_idxs


def _selection(interferometer, basis, roq_idxs, ifo_idxs):
    """The basis columns and masked-frequency indices the sums run over:
    every masked frequency when no ROQ grid is given (the basis must then
    cover exactly those), else the grid intersection."""
    n_m = int(np.sum(interferometer.frequency_mask))
    if roq_idxs is None:
        if basis.shape[1] != n_m:
            raise ValueError("Mismatch between ROQ basis and frequency array for %s"
                             % interferometer.name)
        return np.arange(n_m), np.arange(n_m)
    return np.asarray(roq_idxs[interferometer.name]), np.asarray(ifo_idxs[interferometer.name])


def linear_weights(interferometer, basis, duration, k0, n_times, n_samples, device=0,
                   roq_idxs=None, ifo_idxs=None):
    """roq.py:862-918 for one detector: [n_times, L] complex on the GPU."""
    from .likelihood import get_context
    ifo = interferometer
    m = ifo.frequency_mask
    basis = np.ascontiguousarray(basis, dtype=complex)
    cols, sel = _selection(ifo, basis, roq_idxs, ifo_idxs)
    d_over_s = (ifo.frequency_domain_strain[m] / ifo.power_spectral_density_array[m])[sel]
    basis = np.ascontiguousarray(basis[:, cols])
    x = np.ascontiguousarray((d_over_s[None, :] * np.conj(basis)).view(np.float64))
    # roq.py:885-886: the masked bins start at int(minimum_frequency * T)
    bins = np.ascontiguousarray(sel.astype(np.int64) + int(ifo.minimum_frequency * duration))
    out = np.zeros((n_times, basis.shape[0], 2))
    check(LIB.gwl_roq_weights_linear(get_context(device), basis.shape[0], len(d_over_s), dptr(x),
                                     bins.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                     int(n_samples), int(k0), int(n_times), 4.0 / duration,
                                     dptr(out)), "gwl_roq_weights_linear")
    return out[..., 0] + 1j * out[..., 1]


def quadratic_weights(interferometer, basis, duration, roq_idxs=None, ifo_idxs=None):
    """roq.py:976-1005: 4/T Re(B_q) . (1 / S) over the masked frequencies."""
    ifo = interferometer
    m = ifo.frequency_mask
    basis = np.asarray(basis)
    cols, sel = _selection(ifo, basis, roq_idxs, ifo_idxs)
    return 4.0 / duration * basis[:, cols].real @ (1.0 / ifo.power_spectral_density_array[m][sel])


def build_roq_weights(interferometers, linear_basis, quadratic_basis, frequency_nodes_linear,
                      frequency_nodes_quadratic, time_prior_minimum, time_prior_maximum,
                      delta_t=None, optimal_snrs=None, device=0, roq_params=None,
                      roq_scale_factor=1.0):
    """The weights dict of roq.py:736-847 for one linear and one quadratic
    basis given on each detector's masked frequencies, or with roq_params on
    the ROQ frequency grid (see roq_frequency_indices)."""
    if delta_t is None:
        delta_t = time_resolution(interferometers, optimal_snrs)
    times, k0, n = time_samples(interferometers, time_prior_minimum, time_prior_maximum, delta_t)
    weights = {"time_samples": times,
               "frequency_nodes_linear": np.asarray(frequency_nodes_linear, float),
               "frequency_nodes_quadratic": np.asarray(frequency_nodes_quadratic, float)}
    roq_idxs = ifo_idxs = None
    if roq_params is not None:
        roq_idxs, ifo_idxs = roq_frequency_indices(interferometers, roq_params, roq_scale_factor)
    T = interferometers.duration
    for ifo in interferometers:
        weights[ifo.name + "_linear"] = linear_weights(ifo, linear_basis, T, k0, len(times), n,
                                                       device, roq_idxs, ifo_idxs)
        weights[ifo.name + "_quadratic"] = quadratic_weights(ifo, quadratic_basis, T, roq_idxs,
                                                             ifo_idxs)
    return weights
