"""GPU ROQ linear weights (gwl_roq_weights_linear: float64 GEMM against the
exponentials at the ROQ times) against the oracle's inverse FFT per basis
element (roq.py:849-918), relative 1e-10, on every ROQ time sample of a
4 s, 1024 Hz segment with a 24-element random basis."""
import sys
import types

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def test_linear_weights_match_oracle_ifft():
    from bilby_amd import roq_weights as rw
    from oracle import core
    from oracle import roq_weights as orw
    say("roq weights")
    T, fs = 4.0, 1024.0
    rng = np.random.default_rng(21)
    f = core.frequency_series(fs, T)
    dets = []
    for name in ("H1", "L1"):
        psd = core.analytic_aligo_psd(f)
        strain, _ = core.noise_realisation(lambda x: core.analytic_aligo_psd(x), 0, np.inf,
                                           len(f), T, rng)
        mask = core.frequency_mask(f, 20, fs / 2, fs)
        dets.append(types.SimpleNamespace(name=name, frequency_array=f, frequency_mask=mask,
                                          frequency_domain_strain=strain * mask,
                                          power_spectral_density_array=psd,
                                          minimum_frequency=20.0))

    class Ifos(list):
        pass
    ifos = Ifos(dets)
    ifos.duration, ifos.start_time, ifos.frequency_array = T, 1000.0, f
    n_m = int(dets[0].frequency_mask.sum())
    lin = rng.normal(size=(24, n_m)) + 1j * rng.normal(size=(24, n_m))
    quad = rng.normal(size=(8, n_m))
    w = rw.build_roq_weights(ifos, lin, quad, np.arange(24.0), np.arange(8.0), 1001.9, 1002.1)
    times = w["time_samples"]
    say("time step", times[1] - times[0], "samples", len(times))
    for d in dets:
        ref = orw.linear_weights(f, d.frequency_mask, d.frequency_domain_strain,
                                 d.power_spectral_density_array, 20.0, T, lin, times)
        err = np.max(np.abs(w[d.name + "_linear"] - ref)) / np.max(np.abs(ref))
        say(d.name, "relative error", err)
        assert w[d.name + "_linear"].shape == (len(times), 24)
        assert err < 1e-10
        assert np.allclose(w[d.name + "_quadratic"],
                           orw.quadratic_weights(d.frequency_mask, d.power_spectral_density_array,
                                                 T, quad), rtol=1e-13, atol=0)
