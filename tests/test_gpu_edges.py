"""GPU edge cases: empty batches on every evaluation path (Python API, the
C ABI with NULL buffers, the server); a NaN in a row's sky or time
parameters stays in that row; the largest segment of the uneven workload
(128 s @ 4096 Hz, three detectors, time + distance + phase marginalised,
262,145 bins per detector) against the FP64 oracle."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def small():
    from bilby_amd.configs import make_taylorf2_marginalised
    from bilby_amd.likelihood import GravitationalWaveTransient
    from tests.helpers import fixed_marginalised
    s = make_taylorf2_marginalised(n_draws=16, seed=31, duration=4.0, sampling_frequency=2048.0)
    priors = s["priors"].copy()
    lk = GravitationalWaveTransient(s["interferometers"], s["waveform_generator"],
                                    time_marginalization=True, distance_marginalization=True,
                                    phase_marginalization=True, priors=priors, precision="fp32")
    return s, lk, fixed_marginalised(s["draws"], priors)


def test_empty_batches(small):
    import torch
    from bilby_amd._lib import LIB
    from bilby_amd.relative import RelativeBinningGravitationalWaveTransient
    from bilby_amd.serving import LikelihoodServer
    s, lk, draws = small
    say("empty batches")
    empty = {k: v[:0] for k, v in draws.items()}
    assert lk.log_likelihood_ratio_batch(empty).shape == (0,)
    assert lk.log_likelihood_ratio_batch(np.zeros((0, 16))).shape == (0,)
    dev = torch.zeros((0, 16), dtype=torch.float64, device="cuda")
    assert tuple(lk.log_likelihood_ratio_batch(dev).shape) == (0,)
    out, aux = lk.log_likelihood_ratio_batch(empty, return_inner_products=True)
    assert out.shape == (0,) and aux.shape == (0, 3)
    # the C ABI: an empty batch needs no buffers
    assert LIB.gwl_eval(lk._event, None, 0, 16, 1, None, None, None) == 0
    rb = RelativeBinningGravitationalWaveTransient(
        s["interferometers"], s["waveform_generator"], fiducial_parameters=dict(s["injection"]),
        epsilon=0.5)
    assert rb.log_likelihood_ratio_batch({k: v[:0] for k, v in draws.items()}).shape == (0,)
    assert LIB.gwl_relbin_eval(rb._relbin, None, 0, 16, 0, None, None, None) == 0
    server = LikelihoodServer(hbm_budget=20e9, max_batch_rows=64)
    try:
        server.register("c2", lk)
        assert server.submit("c2", np.zeros((0, 16))).result().shape == (0,)
    finally:
        server.close()


def test_nan_parameters_stay_in_their_row(small):
    s, lk, draws = small
    say("nan rows")
    good = lk.log_likelihood_ratio_batch(draws)
    for key in ("ra", "geocent_time", "psi"):
        d = {k: v.copy() for k, v in draws.items()}
        d[key][3] = np.nan
        out = lk.log_likelihood_ratio_batch(d)
        keep = [i for i in range(len(good)) if i != 3]
        assert np.array_equal(out[keep], good[keep]), key
        assert not (np.isfinite(out[3]) and out[3] > np.nan_to_num(-np.inf)), (key, out[3])


def test_largest_segment_matches_oracle():
    from bilby_amd.configs import make_taylorf2_marginalised
    from bilby_amd.likelihood import GravitationalWaveTransient
    from tests.helpers import draws_as_dicts, fixed_marginalised, oracle_from_setup
    say("128 s @ 4096 Hz, H1+L1+V1, T+D+P")
    s = make_taylorf2_marginalised(n_draws=6, seed=41, duration=128.0, sampling_frequency=4096.0,
                                   chi_range=0.5)
    priors = s["priors"].copy()
    lk = GravitationalWaveTransient(s["interferometers"], s["waveform_generator"],
                                    time_marginalization=True, distance_marginalization=True,
                                    phase_marginalization=True, priors=priors, precision="fp32")
    draws = fixed_marginalised(s["draws"], priors)
    f32 = lk.log_likelihood_ratio_batch(draws)
    f64 = lk.log_likelihood_ratio_batch(draws, precision="fp64")
    olk, _ = oracle_from_setup(s)
    ref = np.array([olk.log_likelihood_ratio(p) for p in draws_as_dicts(draws, range(6))])
    say("bins", len(s["interferometers"].frequency_array), "fp32", np.max(np.abs(f32 - ref)),
        "fp64", np.max(np.abs(f64 - ref)), "lnL range", ref.min(), ref.max())
    assert np.all(np.isfinite(ref))
    assert np.max(np.abs(f64 - ref)) < 1e-6
    assert np.max(np.abs(f32 - ref)) < 1e-3
