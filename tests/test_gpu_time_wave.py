"""GPU checks of the wave-per-row TaylorF2 time transform (gwl_time_wave.hip):
the fused epilogue (one group of sub-sequences per row), the split-group
path and the tiled kernel agree with each other and with the FP64 oracle,
for a heavy-binary support (a few slots of the transform) and a light one
whose support spans every slot."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def _setup(masses, flags, n_draws=160, seed=11):
    from bilby_amd.configs import make_taylorf2_marginalised
    s = make_taylorf2_marginalised(n_draws=n_draws, seed=seed, duration=4.0,
                                   sampling_frequency=2048.0, detectors=("H1", "L1", "V1"),
                                   chi_range=0.5)
    rng = np.random.default_rng(seed + 1)
    s["draws"]["mass_1"] = rng.uniform(*masses, n_draws)
    s["draws"]["mass_2"] = rng.uniform(*masses, n_draws)
    s.update(time_marginalization=True, distance_marginalization=flags["distance"],
             phase_marginalization=flags["phase"])
    return s


@pytest.mark.parametrize("masses", [(5.0, 50.0), (1.0, 2.5)])
@pytest.mark.parametrize("flags", [dict(distance=True, phase=True),
                                   dict(distance=False, phase=True),
                                   dict(distance=False, phase=False)])
def test_wave_paths_agree_with_tiled_and_oracle(masses, flags):
    """fused (G = 1) and split (G > 1) wave paths vs the tiled kernel:
    |dlnL| < 2e-4 (float32 transforms summed in different orders); each vs
    the FP64 oracle on 24 rows: < 1e-3."""
    from bilby_amd.autotune import set_tile, set_transform
    from bilby_amd.likelihood import GravitationalWaveTransient
    from tests.helpers import draws_as_dicts, fixed_marginalised, oracle_from_setup
    say("time wave", masses, flags)
    s = _setup(masses, flags)
    priors = s["priors"].copy()
    lk = GravitationalWaveTransient(s["interferometers"], s["waveform_generator"],
                                    time_marginalization=True,
                                    distance_marginalization=flags["distance"],
                                    phase_marginalization=flags["phase"], priors=priors,
                                    precision="fp32")
    draws = fixed_marginalised(s["draws"], priors)
    set_transform(lk, 2)
    set_tile(lk, 0, 1)      # wave path, one group per row for 160 rows: fused
    fused = lk.log_likelihood_ratio_batch(draws)
    set_tile(lk, 0, 64)     # wave path, several groups per row: split
    split = lk.log_likelihood_ratio_batch(draws)
    set_transform(lk, 1)
    set_tile(lk, 8, 2)      # tiled kernel
    tiled = lk.log_likelihood_ratio_batch(draws)
    set_transform(lk, 0)
    set_tile(lk, 0, 0)
    n = 24
    olk, _ = oracle_from_setup(s)
    ref = np.array([olk.log_likelihood_ratio(p) for p in draws_as_dicts(draws, range(n))])
    say("fused-tiled", np.max(np.abs(fused - tiled)), "split-tiled", np.max(np.abs(split - tiled)),
        "fused-oracle", np.max(np.abs(fused[:n] - ref)), "lnL range", ref.min(), ref.max())
    assert np.all(np.isfinite(ref))
    assert np.max(np.abs(fused - tiled)) < 2e-4
    assert np.max(np.abs(split - tiled)) < 2e-4
    assert np.max(np.abs(fused[:n] - ref)) < 1e-3
    assert np.max(np.abs(split[:n] - ref)) < 1e-3


def test_wave_invalid_rows_and_repeat():
    """Invalid rows get nan_to_num(-inf) on the fused path without touching
    their neighbours, and a repeated call gives the same bits."""
    from bilby_amd.autotune import set_tile, set_transform
    from bilby_amd.likelihood import GravitationalWaveTransient
    from tests.helpers import fixed_marginalised
    s = _setup((5.0, 50.0), dict(distance=True, phase=True), n_draws=64)
    priors = s["priors"].copy()
    lk = GravitationalWaveTransient(s["interferometers"], s["waveform_generator"],
                                    time_marginalization=True, distance_marginalization=True,
                                    phase_marginalization=True, priors=priors)
    d0 = {k: v.copy() for k, v in fixed_marginalised(s["draws"], priors).items()}
    for transform in (2, 1):
        set_transform(lk, transform)
        set_tile(lk, 0, 1)
        d = {k: v.copy() for k, v in d0.items()}
        good = lk.log_likelihood_ratio_batch(d)
        again = lk.log_likelihood_ratio_batch(d)
        assert np.array_equal(good, again)
        d["mass_1"][7] = -3.0
        bad = lk.log_likelihood_ratio_batch(d)
        assert bad[7] == np.nan_to_num(-np.inf)
        keep = [i for i in range(64) if i != 7]
        assert np.array_equal(bad[keep], good[keep])
    set_transform(lk, 0)
