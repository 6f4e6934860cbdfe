"""Time marginalisation with geocent_time priors other than Uniform
(base.py:794-820 weighs each time by prior.prob(t) delta_tc): a Gaussian
prior (unbounded; its tails trimmed where the density vanishes) and a
truncated Gaussian with hard edges, TaylorF2 T+D+P and T alone on a 4 s /
2048 Hz H1+L1 event, against the FP64 oracle that evaluates prob at the
jittered times directly: float64 within 1e-6, float32 within 1e-3."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("kind", ["gaussian", "truncated"])
@pytest.mark.parametrize("marg", [dict(distance=True, phase=True), dict(distance=False, phase=False)])
def test_non_uniform_time_prior_matches_oracle(kind, marg):
    from bilby_amd.configs import make_taylorf2_marginalised
    from bilby_amd.likelihood import GravitationalWaveTransient
    from bilby_amd.prior import Gaussian, TruncatedGaussian
    from tests.helpers import draws_as_dicts, fixed_marginalised, oracle_from_setup
    say("time prior", kind, marg)
    s = make_taylorf2_marginalised(n_draws=48, seed=31, duration=4.0, sampling_frequency=2048.0,
                                   detectors=("H1", "L1"), distance=marg["distance"],
                                   phase=marg["phase"])
    tc = s["injection"]["geocent_time"]
    if kind == "gaussian":
        tp = Gaussian(tc + 0.004, 0.015, name="geocent_time")
    else:
        tp = TruncatedGaussian(tc - 0.01, 0.03, tc - 0.06, tc + 0.05, name="geocent_time")
    s["priors"]["geocent_time"] = tp
    priors = s["priors"].copy()
    lk = GravitationalWaveTransient(s["interferometers"], s["waveform_generator"],
                                    time_marginalization=True,
                                    distance_marginalization=marg["distance"],
                                    phase_marginalization=marg["phase"], priors=priors)
    draws = fixed_marginalised(s["draws"], priors)
    n = 24
    olk, _ = oracle_from_setup(s)
    ref = np.array([olk.log_likelihood_ratio(p) for p in draws_as_dicts(draws, range(n))])
    d = {k: v[:n] for k, v in draws.items()}
    out64 = lk.log_likelihood_ratio_batch(d, precision="fp64")
    out32 = lk.log_likelihood_ratio_batch(d, precision="fp32")
    say("fp64", np.max(np.abs(out64 - ref)), "fp32", np.max(np.abs(out32 - ref)),
        "lnL range", ref.min(), ref.max())
    assert np.all(np.isfinite(ref))
    assert np.max(np.abs(out64 - ref)) < 1e-6
    assert np.max(np.abs(out32 - ref)) < 1e-3


@pytest.mark.parametrize("marg", [dict(distance=False, phase=True), dict(distance=True, phase=True)])
def test_full_segment_default_time_prior_on_a_16s_4096Hz_event(marg):
    """bilby's default when no geocent_time prior is given: Uniform over the
    whole segment (base.py:362-369).  On a 16 s, 4096 Hz event the window is
    32768 samples, evaluated in four slices of one 8192-point transform each
    (float32 and float64), against the FP64 oracle with the same prior:
    float64 within 1e-6, float32 within 1e-3."""
    from bilby_amd.configs import make_taylorf2_marginalised
    from bilby_amd.likelihood import GravitationalWaveTransient
    from bilby_amd.prior import Uniform
    from tests.helpers import draws_as_dicts, fixed_marginalised, oracle_from_setup
    say("full-segment time prior, 16 s @ 4096 Hz", marg)
    s = make_taylorf2_marginalised(n_draws=6, seed=33, duration=16.0, sampling_frequency=4096.0,
                                   detectors=("H1", "L1"), distance=marg["distance"],
                                   phase=marg["phase"])
    ifos = s["interferometers"]
    del s["priors"]["geocent_time"]
    priors = s["priors"].copy()
    lk = GravitationalWaveTransient(ifos, s["waveform_generator"], time_marginalization=True,
                                    distance_marginalization=marg["distance"],
                                    phase_marginalization=marg["phase"], priors=priors)
    tp = lk.priors["geocent_time"]
    assert (tp.minimum, tp.maximum) == (ifos.start_time, ifos.start_time + ifos.duration)
    s["priors"]["geocent_time"] = Uniform(ifos.start_time, ifos.start_time + ifos.duration)
    olk, _ = oracle_from_setup(s)
    draws = fixed_marginalised(s["draws"], priors)
    ref = np.array([olk.log_likelihood_ratio(p) for p in draws_as_dicts(draws, range(6))])
    out64 = lk.log_likelihood_ratio_batch(draws, precision="fp64")
    out32 = lk.log_likelihood_ratio_batch(draws, precision="fp32")
    say("fp64", np.max(np.abs(out64 - ref)), "fp32", np.max(np.abs(out32 - ref)),
        "lnL range", ref.min(), ref.max())
    assert np.all(np.isfinite(ref))
    assert np.max(np.abs(out64 - ref)) < 1e-6
    assert np.max(np.abs(out32 - ref)) < 1e-3
