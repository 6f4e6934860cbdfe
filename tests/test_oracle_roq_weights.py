"""ROQ weight construction pinned on the CPU: the oracle's inverse-FFT form
(roq.py:849-918) equals the direct sum 4/T sum_f (d/S) conj(B) e^{2 pi i f t}
at the ROQ times (the closed form the FFT evaluates), and the product's
host helpers (time resolution, time samples, quadratic weights) equal the
restatement of roq.py:746-766, 976-1005, 1165-1229."""
import types

import numpy as np

from oracle import core
from oracle import roq_weights as orw


def _ifos(T=4.0, fs=1024.0, seed=3):
    rng = np.random.default_rng(seed)
    f = core.frequency_series(fs, T)
    out = []
    for name in ("H1", "L1"):
        psd = core.analytic_aligo_psd(f)
        strain, _ = core.noise_realisation(lambda x: core.analytic_aligo_psd(x), 0, np.inf,
                                           len(f), T, rng)
        mask = core.frequency_mask(f, 20, fs / 2, fs)
        out.append(types.SimpleNamespace(name=name, frequency_array=f, frequency_mask=mask,
                                         frequency_domain_strain=strain * mask,
                                         power_spectral_density_array=psd,
                                         minimum_frequency=20.0))
    ifos = types.SimpleNamespace(duration=T, start_time=1000.0, frequency_array=f,
                                 sampling_frequency=fs)
    return ifos, out


def test_ifft_weights_equal_direct_sum():
    ifos, dets = _ifos()
    T = ifos.duration
    det = dets[0]
    m = det.frequency_mask
    rng = np.random.default_rng(5)
    basis = rng.normal(size=(6, int(m.sum()))) + 1j * rng.normal(size=(6, int(m.sum())))
    dt = T / 4096
    times = np.arange(1800, 1830) * dt
    w = orw.linear_weights(det.frequency_array, m, det.frequency_domain_strain,
                           det.power_spectral_density_array, 20.0, T, basis, times)
    f = det.frequency_array[m]
    x = det.frequency_domain_strain[m] / det.power_spectral_density_array[m]
    for k in (0, 7, 29):
        for i in (0, 3, 5):
            direct = 4.0 / T * np.sum(x * np.conj(basis[i]) * np.exp(2j * np.pi * f * times[k]))
            assert abs(w[k, i] - direct) < 1e-10 * np.max(np.abs(w))


def test_product_host_helpers_equal_restatement():
    from bilby_amd import roq_weights as rw
    ifos, dets = _ifos()
    lst = types.SimpleNamespace(duration=ifos.duration, start_time=ifos.start_time,
                                frequency_array=ifos.frequency_array)

    class L(list):
        pass
    il = L(dets)
    il.duration, il.start_time, il.frequency_array = lst.duration, lst.start_time, lst.frequency_array
    dt = rw.time_resolution(il)
    n = il.duration / dt
    assert abs(n - round(n)) < 1e-9 and int(round(n)) & (int(round(n)) - 1) == 0
    assert n >= il.frequency_array[-1] * il.duration + 1
    times, k0, nseg = rw.time_samples(il, 1001.9, 1002.1, dt)
    pad = 2 * 6378136.6 / 299792458.0 + 5 * dt
    assert k0 == int(np.floor((1001.9 - pad - 1000.0) / dt))
    assert len(times) == int(np.ceil((1002.1 + pad - 1000.0) / dt)) - k0 + 1
    assert nseg == int(il.duration / dt)
    rng = np.random.default_rng(6)
    q = rng.normal(size=(5, int(dets[0].frequency_mask.sum())))
    a = rw.quadratic_weights(dets[0], q, il.duration)
    b = orw.quadratic_weights(dets[0].frequency_mask, dets[0].power_spectral_density_array,
                              il.duration, q)
    assert np.allclose(a, b, rtol=1e-14, atol=0)
