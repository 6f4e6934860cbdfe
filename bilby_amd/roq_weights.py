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
exponentials at the window of time samples).  One basis (no prior-range
selection among several) and bases on the full masked grid (not
multibanded bases) are supported.
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


def linear_weights(interferometer, basis, duration, k0, n_times, n_samples, device=0):
    """roq.py:862-918 for one detector: [n_times, L] complex on the GPU."""
    from .likelihood import get_context
    ifo = interferometer
    m = ifo.frequency_mask
    basis = np.ascontiguousarray(basis, dtype=complex)
    d_over_s = ifo.frequency_domain_strain[m] / ifo.power_spectral_density_array[m]
    if basis.shape[1] != len(d_over_s):
        raise ValueError("Mismatch between ROQ basis and frequency array for %s" % ifo.name)
    x = np.ascontiguousarray((d_over_s[None, :] * np.conj(basis)).view(np.float64))
    # roq.py:885-886: the masked bins start at int(minimum_frequency * T)
    bins = np.ascontiguousarray(np.arange(len(d_over_s), dtype=np.int64)
                                + int(ifo.minimum_frequency * duration))
    out = np.zeros((n_times, basis.shape[0], 2))
    check(LIB.gwl_roq_weights_linear(get_context(device), basis.shape[0], len(d_over_s), dptr(x),
                                     bins.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                     int(n_samples), int(k0), int(n_times), 4.0 / duration,
                                     dptr(out)), "gwl_roq_weights_linear")
    return out[..., 0] + 1j * out[..., 1]


def quadratic_weights(interferometer, basis, duration):
    """roq.py:976-1005: 4/T Re(B_q) . (1 / S) over the masked frequencies."""
    ifo = interferometer
    m = ifo.frequency_mask
    return 4.0 / duration * np.asarray(basis).real @ (1.0 / ifo.power_spectral_density_array[m])


def build_roq_weights(interferometers, linear_basis, quadratic_basis, frequency_nodes_linear,
                      frequency_nodes_quadratic, time_prior_minimum, time_prior_maximum,
                      delta_t=None, optimal_snrs=None, device=0):
    """The weights dict of roq.py:736-847 for one linear and one quadratic
    basis given on each detector's masked frequencies."""
    if delta_t is None:
        delta_t = time_resolution(interferometers, optimal_snrs)
    times, k0, n = time_samples(interferometers, time_prior_minimum, time_prior_maximum, delta_t)
    weights = {"time_samples": times,
               "frequency_nodes_linear": np.asarray(frequency_nodes_linear, float),
               "frequency_nodes_quadratic": np.asarray(frequency_nodes_quadratic, float)}
    T = interferometers.duration
    for ifo in interferometers:
        weights[ifo.name + "_linear"] = linear_weights(ifo, linear_basis, T, k0, len(times), n,
                                                       device)
        weights[ifo.name + "_quadratic"] = quadratic_weights(ifo, quadratic_basis, T)
    return weights
