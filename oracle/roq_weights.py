"""FP64 restatement of bilby's ROQ weight construction
(bilby/gw/likelihood/roq.py:849-918 _set_weights_linear, :976-1005
_set_weights_quadratic) for one basis: one inverse FFT of the full segment
per basis element, as the reference does with numpy.fft.

Test infrastructure only (see oracle/__init__.py).
"""
import numpy as np


def linear_weights(frequency_array, mask, strain, psd, minimum_frequency, duration, basis,
                   time_samples):
    """roq.py:862-918: weights [n_times, L] (complex)."""
    time_space = time_samples[1] - time_samples[0]
    n_samples = int(duration / time_space)
    start_idx = int(time_samples[0] / time_space)
    end_idx = int(time_samples[-1] / time_space)
    ifo_idxs = np.arange(int(np.sum(mask)))
    nonzero = ifo_idxs + int(minimum_frequency * duration)
    d_over_s = strain[mask][ifo_idxs] / psd[mask][ifo_idxs]
    weights = np.zeros((len(basis), len(time_samples)), dtype=complex)
    buf = np.zeros(n_samples, dtype=complex)
    for i in range(len(basis)):
        buf[:] = 0
        buf[nonzero] = d_over_s * np.conj(basis[i])
        weights[i] = np.fft.ifft(buf)[start_idx:end_idx + 1]
    return weights.T * 4.0 * n_samples / float(duration)


def quadratic_weights(mask, psd, duration, basis):
    """roq.py:997-1004."""
    return 4.0 / duration * np.asarray(basis).real @ (1 / psd[mask])
