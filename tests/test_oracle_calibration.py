"""The calibration-marginalisation oracle (oracle/calibration.py) and the
product's host-side calibration models pinned on the CPU:
* marginalised ln L equals the logsumexp over curves (weights 1/C) of the
  un-marginalised ln L with each curve applied to the signal, the
  reference's own check (marginalization_test.py:525-555, 7 places);
* all-identity curves give the uncalibrated likelihood exactly;
* the cubic-spline response equals scipy's not-a-knot cubic spline in
  log10 f (the end rows M0 - 2 M1 + M2 = 0 are the not-a-knot condition for
  equally spaced nodes) composed as (1 + dA)(2 + i dphi)/(2 - i dphi);
* bilby_amd.calibration's spline and curve lookup equal the oracle's."""
import numpy as np
import pytest
from scipy.interpolate import CubicSpline as ScipyCubicSpline
from scipy.special import logsumexp

from oracle import core
from oracle.calibration import OracleCalibrated, cubic_spline_factor
from oracle.likelihood import OracleDetector, OracleEvent, OracleGWTransient, detector_response

WF = dict(reference_frequency=50.0, minimum_frequency=20.0)


def _event(seed=2, T=4.0, fs=1024.0, dets=("H1", "L1")):
    rng = np.random.default_rng(seed)
    f = core.frequency_series(fs, T)
    start = 1126259642.413 - T / 2
    inj = dict(mass_1=30.0, mass_2=25.0, chi_1=0.0, chi_2=0.0, luminosity_distance=500.0,
               theta_jn=0.5, phase=1.1, ra=1.2, dec=-0.4, psi=0.8, geocent_time=start + T / 2)
    pols = core.taylorf2(f, inj["mass_1"], inj["mass_2"], inj["luminosity_distance"],
                         inj["theta_jn"], inj["phase"], **WF)
    out = []
    for name in dets:
        psd = core.analytic_aligo_psd(f)
        noise, _ = core.noise_realisation(lambda x: core.analytic_aligo_psd(x), 0, np.inf,
                                          len(f), T, rng)
        mask = core.frequency_mask(f, 20, fs / 2, fs)
        det = OracleDetector(name, f, noise, psd, mask)
        det.strain = det.strain + detector_response(det, pols, inj, start) * mask
        out.append(det)
    return OracleEvent(out, T, fs, start), inj


def _curves(ev, n, seed, sigma=(0.1, 0.1), nodes=5):
    rng = np.random.default_rng(seed)
    out = {}
    for det in ev.detectors:
        mf = det.frequency_array[det.mask]
        cs = [cubic_spline_factor(mf, rng.normal(0, sigma[0], nodes), rng.normal(0, sigma[1], nodes),
                                  mf[0], mf[-1]) for _ in range(n)]
        out[det.name] = np.array(cs)
    return out


@pytest.mark.parametrize("phase", [False, True])
def test_marginalised_equals_logsumexp_of_per_curve_likelihoods(phase):
    ev, inj = _event()
    curves = _curves(ev, 100, seed=3)
    kw = dict(waveform_arguments=WF, phase_marginalization=phase,
              priors={"phase": 0.0} if phase else None)
    lk = OracleCalibrated(ev, "taylorf2", curves, **kw)
    p = dict(inj)
    p["mass_1"] *= 1.003
    if phase:
        p["phase"] = 0.0
    marg = lk.log_likelihood_ratio(p)
    per = [lk.log_likelihood_ratio_with_curve(p, i) for i in range(100)]
    assert marg == pytest.approx(logsumexp(per, b=1 / 100), abs=1e-7)
    assert np.allclose(lk.per_curve_log_likelihoods(p), per, rtol=0, atol=1e-8)


def test_identity_curves_give_the_uncalibrated_likelihood():
    ev, inj = _event(seed=4)
    ones = {det.name: np.ones((7, int(det.mask.sum())), complex) for det in ev.detectors}
    lk = OracleCalibrated(ev, "taylorf2", ones, waveform_arguments=WF)
    full = OracleGWTransient(ev, "taylorf2", waveform_arguments=WF)
    assert lk.log_likelihood_ratio(inj) == pytest.approx(full.log_likelihood_ratio(inj),
                                                         rel=1e-12, abs=1e-9)


class _U:
    def __init__(self, a, b):
        self.minimum, self.maximum = a, b

    def prob(self, x):
        x = np.asarray(x)
        return ((x >= self.minimum) & (x <= self.maximum)) / (self.maximum - self.minimum)

    def rescale(self, u):
        return self.minimum + u * (self.maximum - self.minimum)


@pytest.mark.parametrize("phase", [False, True])
def test_time_marginalised_equals_logsumexp_of_per_curve_likelihoods(phase):
    """The reference's check with time marginalisation
    (marginalization_test.py:524-554, "time", jitter off): the calibration-
    and time-marginalised ln L (one FFT per curve of the tiled, curve-
    multiplied integrand, base.py:305-324) equals the logsumexp over curves
    (weights 1/C) of the time-marginalised ln L with each curve applied to
    the signal (base.py:286-290, 325-330), to 7 places."""
    ev, inj = _event()
    curves = _curves(ev, 100, seed=5)
    tc = inj["geocent_time"]
    pri = {"geocent_time": _U(tc - 0.1, tc + 0.1)}
    if phase:
        pri["phase"] = 0.0
    lk = OracleCalibrated(ev, "taylorf2", curves, waveform_arguments=WF,
                          time_marginalization=True, jitter_time=False,
                          phase_marginalization=phase, priors=pri)
    p = dict(inj, geocent_time=ev.start_time)
    p["mass_1"] *= 1.003
    if phase:
        p["phase"] = 0.0
    marg = lk.log_likelihood_ratio(p)
    per = [lk.log_likelihood_ratio_with_curve(p, i) for i in range(100)]
    assert max(per) > 20
    assert marg == pytest.approx(logsumexp(per, b=1 / 100), abs=1e-7)
    assert np.allclose(lk.per_curve_log_likelihoods(p), per, rtol=0, atol=1e-8)


def test_time_marginalised_identity_curves_give_the_time_marginalised_likelihood():
    ev, inj = _event(seed=4)
    ones = {det.name: np.ones((5, int(det.mask.sum())), complex) for det in ev.detectors}
    tc = inj["geocent_time"]
    lk = OracleCalibrated(ev, "taylorf2", ones, waveform_arguments=WF, time_marginalization=True,
                          priors={"geocent_time": _U(tc - 0.1, tc + 0.1)})
    full = OracleGWTransient(ev, "taylorf2", waveform_arguments=WF, time_marginalization=True,
                             priors={"geocent_time": _U(tc - 0.1, tc + 0.1)})
    p = dict(inj, geocent_time=ev.start_time, time_jitter=1e-4)
    assert lk.log_likelihood_ratio(p) == pytest.approx(full.log_likelihood_ratio(p),
                                                       rel=1e-12, abs=1e-9)


def test_cubic_spline_is_not_a_knot_in_log_frequency():
    rng = np.random.default_rng(5)
    f = np.linspace(20.0, 512.0, 2000)
    for n in (4, 5, 10):
        amp, ph = rng.normal(0, 0.1, n), rng.normal(0, 0.1, n)
        got = cubic_spline_factor(f, amp, ph, 20.0, 512.0)
        x = np.linspace(np.log10(20.0), np.log10(512.0), n)
        sa = ScipyCubicSpline(x, amp, bc_type="not-a-knot")(np.log10(f))
        sp = ScipyCubicSpline(x, ph, bc_type="not-a-knot")(np.log10(f))
        want = (1 + sa) * (2 + 1j * sp) / (2 - 1j * sp)
        assert np.max(np.abs(got - want)) < 1e-12
        assert np.allclose(cubic_spline_factor(f, np.zeros(n), np.zeros(n), 20.0, 512.0), 1.0)


def test_product_calibration_models_equal_the_oracle(tmp_path):
    from bilby_amd import calibration as cal
    from bilby_amd.prior import PriorDict
    f = np.linspace(20.0, 512.0, 977)
    spline = cal.CubicSpline("recalib_H1_", 20.0, 512.0, 6)
    rng = np.random.default_rng(6)
    amp, ph = rng.normal(0, 0.1, 6), rng.normal(0, 0.1, 6)
    params = {"recalib_H1_amplitude_%d" % i: amp[i] for i in range(6)}
    params.update({"recalib_H1_phase_%d" % i: ph[i] for i in range(6)})
    got = spline.get_calibration_factor(f, **params)
    assert np.max(np.abs(got - cubic_spline_factor(f, amp, ph, 20.0, 512.0))) < 1e-14
    # curves from priors: each curve is the spline of its own draw
    priors = cal.constant_uncertainty_spline_priors(0.1, 0.1, 20.0, 512.0, 5, "H1")
    draws = PriorDict(priors).sample(12, rng=np.random.default_rng(7))
    curves = cal.curves_from_spline_and_prior(draws, "H1", 5, f, 12)
    for i in range(12):
        a = [draws["recalib_H1_amplitude_%d" % k][i] for k in range(5)]
        p = [draws["recalib_H1_phase_%d" % k][i] for k in range(5)]
        assert np.max(np.abs(curves[i] - cubic_spline_factor(f, a, p, f[0], f[-1]))) < 1e-14
    # lookup files round trip (start index respected)
    fn = str(tmp_path / "H1.npz")
    cal.write_calibration_file(fn, f, curves, draws)
    back, bp = cal._read_lookup(fn, f, 5, 3)
    assert np.array_equal(back, curves[3:8])
    assert np.array_equal(bp["recalib_H1_amplitude_0"], draws["recalib_H1_amplitude_0"][3:8])


def test_inner_products_with_a_drawn_curve():
    """base.py:286-294 (signal[mask] *= curve) as the distance / phase
    reconstruction uses it after a calibration draw: a constant curve c
    scales <d|h> by conj(c) and <h|h> by |c|^2; the identity curve and no
    recalib_index give the plain inner products; a random curve's values
    equal the curve-applied signal's inner products computed from the
    per-curve arrays of calibration_snrs (conjugated: those hold
    conj(d) h cal / S)."""
    ev, inj = _event(seed=5)
    c = 0.9 * np.exp(0.3j)
    n = {det.name: int(det.mask.sum()) for det in ev.detectors}
    rand = _curves(ev, 3, 11)
    curves = {k: np.stack([np.full(n[k], c), np.ones(n[k]), rand[k][2]]) for k in n}
    o = OracleCalibrated(ev, "taylorf2", curves, waveform_arguments=WF)
    dh, hh = o.plain_inner_products(dict(inj))
    a, b = o.plain_inner_products(dict(inj, recalib_index=0))
    assert abs(a - np.conj(c) * dh) < 1e-12 * abs(dh) and abs(b - abs(c) ** 2 * hh) < 1e-12 * hh
    a, b = o.plain_inner_products(dict(inj, recalib_index=1))
    assert abs(a - dh) < 1e-12 * abs(dh) and abs(b - hh) < 1e-12 * hh
    a, b = o.plain_inner_products(dict(inj, recalib_index=2))
    arr, hh_arr = o.calibration_snrs(dict(inj))
    assert abs(a - np.conj(arr[2])) < 1e-10 * abs(a) and abs(b - hh_arr[2]) < 1e-10 * b
