"""FP64 restatement of bilby's ROQ weight construction
(bilby/gw/likelihood/roq.py:849-918 _set_weights_linear, :976-1005
_set_weights_quadratic) for one basis: one inverse FFT of the full segment
per basis element, as the reference does with numpy.fft.

Test infrastructure only (see oracle/__init__.py).
"""
import numpy as np


def roq_indices(frequency_array, mask, flow, fhigh, seglen, scale=1.0):
    """roq.py:800-826: indices into the ROQ frequency grid (the frequencies
    of a segment of seglen / s at sampling rate 2 fhigh s, series.py:115-134,
    from flow s) and into the masked frequencies of the frequencies both hold
    (np.intersect1d, as the reference)."""
    fs, duration = 2 * fhigh * scale, seglen / scale
    n = int(np.round(duration * fs))
    f_roq = np.linspace(0, fs / 2, int(np.round(n / 2) + 1))
    f_roq = f_roq[f_roq >= flow * scale]
    _, ifo_idxs, roq_idxs = np.intersect1d(frequency_array[mask], f_roq, return_indices=True)
    return roq_idxs, ifo_idxs


def linear_weights(frequency_array, mask, strain, psd, minimum_frequency, duration, basis,
                   time_samples, roq_idxs=None, ifo_idxs=None):
    """roq.py:862-918: weights [n_times, L] (complex); roq_idxs / ifo_idxs
    select the basis columns and masked frequencies (all of them when
    omitted)."""
    time_space = time_samples[1] - time_samples[0]
    n_samples = int(duration / time_space)
    start_idx = int(time_samples[0] / time_space)
    end_idx = int(time_samples[-1] / time_space)
    if ifo_idxs is None:
        ifo_idxs = np.arange(int(np.sum(mask)))
        roq_idxs = ifo_idxs
    nonzero = ifo_idxs + int(minimum_frequency * duration)
    d_over_s = strain[mask][ifo_idxs] / psd[mask][ifo_idxs]
    weights = np.zeros((len(basis), len(time_samples)), dtype=complex)
    buf = np.zeros(n_samples, dtype=complex)
    for i in range(len(basis)):
        buf[:] = 0
        buf[nonzero] = d_over_s * np.conj(basis[i][roq_idxs])
        weights[i] = np.fft.ifft(buf)[start_idx:end_idx + 1]
    return weights.T * 4.0 * n_samples / float(duration)


def quadratic_weights(mask, psd, duration, basis, roq_idxs=None, ifo_idxs=None):
    """roq.py:997-1004."""
    basis = np.asarray(basis)
    if ifo_idxs is None:
        return 4.0 / duration * basis.real @ (1 / psd[mask])
    return 4.0 / duration * basis[:, roq_idxs].real @ (1 / psd[mask][ifo_idxs])
