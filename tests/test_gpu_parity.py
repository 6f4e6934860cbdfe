"""GPU parity: libgwl (HIP, gfx950) vs the FP64 oracle on the same seeded
inputs.  Tolerances are stated per test."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    return torch.device("cuda:0")


@pytest.fixture(scope="module")
def ctx():
    from bilby_amd.likelihood import get_context
    return get_context(0)


# ------------------------------------------------------------------ operators
@pytest.mark.parametrize("P", [2, 8, 64, 512, 1024, 4096, 8192, 16384, 65536])
@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_fft_batch_matches_numpy(ctx, torch_dev, P, prec):
    """The LDS radix FFT vs numpy.fft.fft (the call at base.py:326); float64
    beyond 8192 points and float32 beyond 16384 through the four-step split.
    Tolerance: max |X - X_np| / max |X_np| < 1e-13 (fp64), 2e-6 (fp32)."""
    import torch
    from bilby_amd._lib import LIB, check
    say("fft", P, prec)
    rng = np.random.default_rng(P)
    batch = 7
    x = rng.normal(size=(batch, P)) + 1j * rng.normal(size=(batch, P))
    cdt = torch.complex64 if prec == "fp32" else torch.complex128
    xt = torch.from_numpy(x).to(cdt).to(torch_dev)
    yt = torch.empty_like(xt)
    check(LIB.gwl_fft_batch(ctx, ctypes.c_void_p(xt.data_ptr()), ctypes.c_void_p(yt.data_ptr()),
                            batch, P, 1 if prec == "fp32" else 0, None))
    torch.cuda.synchronize()
    y = yt.cpu().numpy()
    ref = np.fft.fft(x.astype(np.complex64 if prec == "fp32" else np.complex128).astype(complex),
                     axis=1)
    err = np.max(np.abs(y - ref)) / np.max(np.abs(ref))
    assert err < (2e-6 if prec == "fp32" else 1e-13), err


def test_ln_i0_device_matches_scipy(ctx, torch_dev):
    """Device ln I0 vs log(scipy i0e(x)) + |x| (bilby/gw/utils.py:1006-1022);
    the reference's own bar is 1e-10 on linspace(-10, 10, 101)
    (test/gw/utils_test.py:349-352); here 1e-10 absolute or 1e-13 relative
    over [-10, 10] and logspace(-6, 11), and on both sides of the series
    lengths' switches (8, 16) and of the large-argument form (30)."""
    import torch
    from scipy.special import i0e
    from bilby_amd._lib import LIB, check
    x = np.concatenate([np.linspace(-10, 10, 101), np.logspace(-6, 11, 2000),
                        [8.0, 8.000001, 16.0, 16.000001, 29.999999, 30.0, 30.000001]])
    xt = torch.from_numpy(x).to(torch_dev)
    yt = torch.empty_like(xt)
    check(LIB.gwl_ln_i0(ctx, ctypes.c_void_p(xt.data_ptr()), ctypes.c_void_p(yt.data_ptr()),
                        len(x), None))
    torch.cuda.synchronize()
    y = yt.cpu().numpy()
    ref = np.log(i0e(x)) + np.abs(x)
    assert np.all(np.abs(y - ref) <= np.maximum(1e-10, 1e-13 * np.abs(ref)))


def test_distance_table_matches_oracle(ctx):
    """GPU lookup table (base.py:994-1018) vs the oracle's C restatement:
    relative 1e-11 (values up to 1.5e11)."""
    from bilby_amd.likelihood import d_inner_h_ref_axis, optimal_snr_squared_ref_axis
    from bilby_amd._lib import LIB, check, dptr
    from bilby_amd.prior import UniformComovingVolume
    from oracle.likelihood import OracleGWTransient
    prior = UniformComovingVolume(100.0, 2000.0)
    dist = np.linspace(prior.minimum, prior.maximum, int(1e4))
    pa = np.array([prior.prob(d) for d in dist])
    ref_d = float(prior.rescale(0.5))
    for phase in (True, False):
        say("distance table phase", phase)
        x = d_inner_h_ref_axis(phase)
        y = optimal_snr_squared_ref_axis()
        out = np.zeros((400, 800))
        check(LIB.gwl_distance_table(ctx, dptr(dist), dptr(pa), len(dist), ref_d, int(phase),
                                     dptr(x), 800, dptr(y), 400, dptr(out)))
        o = OracleGWTransient.__new__(OracleGWTransient)
        o.phase_marginalization = phase
        o._distance_array = dist
        o.distance_prior_array = pa
        o._ref_dist = ref_d
        ref = o.create_lookup_table()
        err = np.max(np.abs(out - ref) / np.maximum(1.0, np.abs(ref)))
        assert err < 1e-11, (phase, err)


# ------------------------------------------------------------------ responses
def _sg_setup():
    from bilby_amd.configs import make_sine_gaussian_plain
    return make_sine_gaussian_plain(n_draws=256, seed=11)


def _tf2_setup(**kw):
    from bilby_amd.configs import make_taylorf2_marginalised
    return make_taylorf2_marginalised(n_draws=96, seed=12, **kw)


def test_sine_gaussian_response_matches_oracle():
    """get_detector_response (interferometer.py:303-368) of the sine-Gaussian,
    bin by bin: |h_gpu - h_oracle| <= 1e-9 max|h| (fp64), 1e-5 max|h| (fp32)."""
    from bilby_amd.likelihood import GravitationalWaveTransient
    from oracle.likelihood import detector_response, source_polarizations
    from tests.helpers import oracle_from_setup
    say("sine-Gaussian response")
    setup = _sg_setup()
    lk = GravitationalWaveTransient(setup["interferometers"], setup["waveform_generator"],
                                    precision="fp64")
    olk, _ = oracle_from_setup(setup)
    for i in range(4):
        p = {k: float(v[i]) for k, v in setup["draws"].items()}
        pols = source_polarizations("sinegaussian", olk.event.frequency_array, p, {})
        for k, det in enumerate(olk.event.detectors):
            ref = detector_response(det, pols, p, olk.event.start_time)
            for prec, tol in (("fp64", 1e-9), ("fp32", 1e-5)):
                h = lk.get_detector_response(p, k, precision=prec)
                err = np.max(np.abs(h - ref)) / np.max(np.abs(ref))
                assert err < tol, (i, k, prec, err)


@pytest.mark.parametrize("chi", [0.0, 0.8])
def test_taylorf2_response_matches_oracle(chi):
    """TaylorF2 detector response bin by bin (fp64 1e-9, fp32 1e-5 of max|h|),
    non-spinning and aligned-spin."""
    from bilby_amd.likelihood import GravitationalWaveTransient
    from oracle.likelihood import detector_response, source_polarizations
    from tests.helpers import oracle_from_setup
    say("taylorf2 response chi", chi)
    setup = _tf2_setup(chi_range=chi, time=False, distance=False, phase=False)
    lk = GravitationalWaveTransient(setup["interferometers"], setup["waveform_generator"],
                                    precision="fp64")
    olk, _ = oracle_from_setup(setup)
    for i in range(4):
        p = {k: float(v[i]) for k, v in setup["draws"].items()}
        pols = source_polarizations("taylorf2", olk.event.frequency_array, p,
                                    olk.waveform_arguments)
        for k, det in enumerate(olk.event.detectors):
            ref = detector_response(det, pols, p, olk.event.start_time)
            if np.max(np.abs(ref)) == 0:
                continue
            for prec, tol in (("fp64", 1e-9), ("fp32", 1e-5)):
                h = lk.get_detector_response(p, k, precision=prec)
                err = np.max(np.abs(h - ref)) / np.max(np.abs(ref))
                assert err < tol, (i, k, prec, err)


def test_taylorf2_isco_termination_opt_in():
    """waveform_arguments taylorf2_termination='isco' (an opt-in with no
    counterpart in the reference): the response is zero above the ISCO and
    equals the oracle's bin by bin, and the T+D+P likelihood agrees with the
    oracle's (fp64 1e-6, fp32 1e-3); without it the response of a heavy
    binary runs to the Nyquist bin, as lalsimulation's TaylorF2 does for
    bilby's fEnd = maximum_frequency."""
    from bilby_amd.likelihood import GravitationalWaveTransient
    from oracle.likelihood import detector_response, source_polarizations
    from tests.helpers import draws_as_dicts, fixed_marginalised, oracle_from_setup
    say("taylorf2 ISCO termination opt-in")
    p = None
    for end in ("isco", None):
        setup = _tf2_setup(duration=4.0, sampling_frequency=2048.0, detectors=("H1", "L1"),
                           taylorf2_termination=end)
        lk = GravitationalWaveTransient(setup["interferometers"], setup["waveform_generator"],
                                        precision="fp64")
        olk, _ = oracle_from_setup(setup)
        f = olk.event.frequency_array
        p = {k: float(v[0]) for k, v in setup["draws"].items()}
        p.update(mass_1=40.0, mass_2=35.0)
        fisco = 1.0 / (6.0 ** 1.5 * np.pi * 75.0 * 4.925490947641267e-06)
        pols = source_polarizations("taylorf2", f, p, olk.waveform_arguments)
        for k, det in enumerate(olk.event.detectors):
            ref = detector_response(det, pols, p, olk.event.start_time)
            h = lk.get_detector_response(p, k, precision="fp64")
            assert np.max(np.abs(h - ref)) < 1e-9 * np.max(np.abs(ref))
            if end == "isco":
                assert np.all(h[f > fisco] == 0) and np.any(h[f <= fisco] != 0)
            else:
                assert h[-1] != 0
        if end == "isco":
            priors = setup["priors"].copy()
            lkm = GravitationalWaveTransient(setup["interferometers"], setup["waveform_generator"],
                                             time_marginalization=True,
                                             distance_marginalization=True,
                                             phase_marginalization=True, priors=priors)
            draws = fixed_marginalised(setup["draws"], priors)
            olkm, _ = oracle_from_setup(setup)
            ref = np.array([olkm.log_likelihood_ratio(q) for q in draws_as_dicts(draws, range(16))])
            d16 = {k: v[:16] for k, v in draws.items()}
            out64 = lkm.log_likelihood_ratio_batch(d16, precision="fp64")
            out32 = lkm.log_likelihood_ratio_batch(d16, precision="fp32")
            say("isco T+D+P fp64", np.max(np.abs(out64 - ref)), "fp32", np.max(np.abs(out32 - ref)))
            assert np.max(np.abs(out64 - ref)) < 1e-6
            assert np.max(np.abs(out32 - ref)) < 1e-3


# ------------------------------------------------------------------ likelihoods
def test_fp32_escalation_meets_1e3_on_every_config1_row():
    """Config 1 (sine-Gaussian, H1+L1, 4 s at 2048 Hz, 10^4 prior draws with
    log-uniform hrss): without escalation float32 misses |dlnL| < 1e-3 on
    the loudest rows; its error stays below the bound coeff * Q used to
    select rows (Q = |<d|h>| + <h|h> from the float64 inner products); with
    escalation every row is within 1e-3 of the float64 GPU path, and the
    escalated rows are returned bit for bit as float64."""
    from bilby_amd.configs import make_sine_gaussian_plain
    from bilby_amd.likelihood import GravitationalWaveTransient
    say("fp32 escalation, config 1, 10^4 rows")
    setup = make_sine_gaussian_plain(n_draws=10000, seed=1)
    lk = GravitationalWaveTransient(setup["interferometers"], setup["waveform_generator"],
                                    precision="fp32")
    d = setup["draws"]
    out64, aux64 = lk.log_likelihood_ratio_batch(d, precision="fp64", return_inner_products=True)
    lk.set_precision_escalation(0.0)
    raw32 = lk.log_likelihood_ratio_batch(d, precision="fp32")
    q = np.hypot(aux64[:, 0], aux64[:, 1]) + aux64[:, 2]
    ratio = np.abs(raw32 - out64) / np.maximum(q, 1e-300)
    lk.set_precision_escalation()
    out32 = lk.log_likelihood_ratio_batch(d, precision="fp32")
    n_esc = lk.escalated_rows()
    err = np.abs(out32 - out64)
    say("raw fp32 max", np.max(np.abs(raw32 - out64)), "rows > 1e-3",
        int(np.sum(np.abs(raw32 - out64) > 1e-3)), "max err/Q", ratio.max(),
        "escalated", n_esc, "max after", err.max(), "max |lnL|", np.max(np.abs(out64)))
    assert ratio.max() < 2e-7        # the default coefficient bounds the float32 error
    assert err.max() < 1e-3
    # selection uses the float32 inner products: allow a 1 % band at the threshold
    sure = 2e-7 * q > 8e-4 * 1.01
    maybe = 2e-7 * q > 8e-4 * 0.99
    assert int(np.sum(sure)) <= n_esc <= int(np.sum(maybe))
    assert np.array_equal(out32[sure], out64[sure])


def test_sine_gaussian_plain_lnl():
    """Plain likelihood (no marginalisation), sine-Gaussian, H1+L1, 4 s at
    2048 Hz, 256 prior draws.  fp64: |dlnL| < 1e-6 + 1e-10 |lnL| (|lnL| up
    to ~1.4e5, where lnL is the difference of inner products of that size
    summed in different orders by the two codes: float64 agreement measured
    at 1e-11 relative); fp32 (rows beyond the float32 error bound escalated
    to float64 inside gwl_eval): |dlnL| < 1e-3.  Inner products also
    checked."""
    from bilby_amd.likelihood import GravitationalWaveTransient
    from tests.helpers import oracle_from_setup, draws_as_dicts
    say("sine-Gaussian plain lnL")
    setup = _sg_setup()
    olk, _ = oracle_from_setup(setup)
    n = 256
    ref = np.array([olk.log_likelihood_ratio(p) for p in draws_as_dicts(setup["draws"], range(n))])
    lk = GravitationalWaveTransient(setup["interferometers"], setup["waveform_generator"],
                                    precision="fp64")
    d = {k: v[:n] for k, v in setup["draws"].items()}
    out64, aux = lk.log_likelihood_ratio_batch(d, precision="fp64", return_inner_products=True)
    out32 = lk.log_likelihood_ratio_batch(d, precision="fp32")
    say("fp64 max", np.max(np.abs(out64 - ref)), "fp32 max", np.max(np.abs(out32 - ref)),
        "max|lnL|", np.max(np.abs(ref)))
    assert np.all(np.abs(out64 - ref) < 1e-6 + 1e-10 * np.abs(ref))
    assert np.all(np.abs(out32 - ref) < 1e-3)
    # noise log-likelihood, base.py:402-413
    assert abs(lk.noise_log_likelihood() - olk.noise_log_likelihood()) < 1e-6
    # single-dict API returns the same float as the batch
    p0 = draws_as_dicts(setup["draws"], [3])[0]
    assert abs(lk.log_likelihood_ratio(p0) - out64[3]) < 1e-9


@pytest.mark.parametrize("flags", [
    dict(time=True, distance=True, phase=True),
    dict(time=True, distance=False, phase=True),
    dict(time=True, distance=False, phase=False),
    dict(time=False, distance=True, phase=True),
    dict(time=False, distance=True, phase=False),
    dict(time=False, distance=False, phase=True),
    dict(time=True, distance=True, phase=False),
])
def test_taylorf2_marginalised_lnl(flags):
    """TaylorF2, H1+L1+V1, 8 s at 4096 Hz, each marginalisation combination:
    |dlnL| < 1e-3 for fp32 and < 1e-6 for fp64 against the FP64 oracle."""
    from bilby_amd.likelihood import GravitationalWaveTransient
    from tests.helpers import oracle_from_setup, draws_as_dicts, fixed_marginalised
    say("taylorf2 marginalised", flags)
    setup = _tf2_setup(**flags)
    olk, opri = oracle_from_setup(setup)
    priors = setup["priors"].copy()
    lk = GravitationalWaveTransient(setup["interferometers"], setup["waveform_generator"],
                                    time_marginalization=flags["time"],
                                    distance_marginalization=flags["distance"],
                                    phase_marginalization=flags["phase"], priors=priors,
                                    precision="fp32")
    draws = fixed_marginalised(setup["draws"], priors)
    n = 48
    ref = np.array([olk.log_likelihood_ratio(p) for p in draws_as_dicts(draws, range(n))])
    d = {k: v[:n] for k, v in draws.items()}
    out32 = lk.log_likelihood_ratio_batch(d, precision="fp32")
    out64 = lk.log_likelihood_ratio_batch(d, precision="fp64")
    say("fp64 max", np.max(np.abs(out64 - ref)), "fp32 max", np.max(np.abs(out32 - ref)),
        "lnL range", ref.min(), ref.max())
    assert np.all(np.isfinite(ref))
    assert np.all(np.abs(out64 - ref) < 1e-6)
    assert np.all(np.abs(out32 - ref) < 1e-3)


def test_invalid_rows_get_sentinel_and_do_not_leak():
    """A row whose waveform is undefined gets nan_to_num(-inf) like
    base.py:424-425; neighbours in the same tile are unchanged."""
    from bilby_amd.likelihood import GravitationalWaveTransient
    say("invalid rows")
    setup = _tf2_setup(time=False, distance=False, phase=False)
    lk = GravitationalWaveTransient(setup["interferometers"], setup["waveform_generator"])
    d = {k: v[:16].copy() for k, v in setup["draws"].items()}
    good = lk.log_likelihood_ratio_batch(d)
    d["mass_1"][5] = -1.0
    d["mass_2"][9] = np.nan
    bad = lk.log_likelihood_ratio_batch(d)
    assert bad[5] == np.nan_to_num(-np.inf) and bad[9] == np.nan_to_num(-np.inf)
    keep = [i for i in range(16) if i not in (5, 9)]
    assert np.array_equal(bad[keep], good[keep])


def test_alone_equals_in_batch():
    """Determinism: a row evaluated alone and inside a larger batch gives the
    same bits without time marginalisation (the reduction order does not
    depend on the co-batched rows), and agrees to 1e-9 with it (the
    transform is split into a batch-dependent number of groups)."""
    from bilby_amd.likelihood import GravitationalWaveTransient
    from tests.helpers import fixed_marginalised
    say("alone vs batch")
    for flags, tol in ((dict(time=False, distance=False, phase=True), 0.0),
                       (dict(time=True, distance=True, phase=True), 1e-9)):
        setup = _tf2_setup(**flags)
        priors = setup["priors"].copy()
        lk = GravitationalWaveTransient(setup["interferometers"], setup["waveform_generator"],
                                        time_marginalization=flags["time"],
                                        distance_marginalization=flags["distance"],
                                        phase_marginalization=True, priors=priors)
        draws = fixed_marginalised(setup["draws"], priors)
        d = {k: v[:90] for k, v in draws.items()}
        batch = lk.log_likelihood_ratio_batch(d)
        for i in (0, 17, 89):
            one = lk.log_likelihood_ratio_batch({k: v[i:i + 1] for k, v in d.items()})
            assert abs(one[0] - batch[i]) <= tol, (flags, i, one[0], batch[i])


@pytest.mark.parametrize("time", [True, False])
def test_four_detector_network_matches_oracle(time):
    """H1+L1+V1+K1 (the kernels' largest detector count), TaylorF2 4 s at
    2048 Hz, distance + phase with and without time: |dlnL| < 1e-6 (fp64)
    and < 1e-3 (fp32) against the FP64 oracle."""
    from bilby_amd.likelihood import GravitationalWaveTransient
    from tests.helpers import oracle_from_setup, draws_as_dicts, fixed_marginalised
    say("four-detector network, time", time)
    setup = _tf2_setup(duration=4.0, sampling_frequency=2048.0,
                       detectors=("H1", "L1", "V1", "K1"), time=time)
    olk, _ = oracle_from_setup(setup)
    priors = setup["priors"].copy()
    lk = GravitationalWaveTransient(setup["interferometers"], setup["waveform_generator"],
                                    time_marginalization=time, distance_marginalization=True,
                                    phase_marginalization=True, priors=priors, precision="fp32")
    draws = fixed_marginalised(setup["draws"], priors)
    n = 24
    ref = np.array([olk.log_likelihood_ratio(p) for p in draws_as_dicts(draws, range(n))])
    d = {k: v[:n] for k, v in draws.items()}
    out32 = lk.log_likelihood_ratio_batch(d, precision="fp32")
    out64 = lk.log_likelihood_ratio_batch(d, precision="fp64")
    say("fp64 max", np.max(np.abs(out64 - ref)), "fp32 max", np.max(np.abs(out32 - ref)))
    assert np.all(np.isfinite(ref))
    assert np.all(np.abs(out64 - ref) < 1e-6)
    assert np.all(np.abs(out32 - ref) < 1e-3)
