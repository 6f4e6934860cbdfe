"""The relative-binning oracle (oracle/relative.py) pinned on the CPU:
* at the fiducial point the ratio is 1 at every edge and the binned
  likelihood equals the full one (relative_binning_test.py:118-126 asks
  1e-3; the summary sums make it exact up to round-off);
* the summary contraction equals the inner products of the piecewise-linear
  reconstruction r0 + r1 (f - f_c) summed on the grid, minus the dropped
  second-order |r1|^2 term for <h|h>;
* draws around the fiducial point: |binned - full| < 0.1 |full(ref) - full|
  (relative_binning_test.py:108-116);
* a very small epsilon still gives a finite value (:151-162);
* the product's vectorised host setup (bilby_amd.relative: bins and summary
  data) equals the oracle's bin-by-bin loops."""
import types

import numpy as np
import pytest

from oracle import core
from oracle.likelihood import OracleDetector, OracleEvent, OracleGWTransient, detector_response
from oracle.relative import OracleRelativeBinning

WF = dict(reference_frequency=50.0, minimum_frequency=20.0)


def _event(seed=3, T=4.0, fs=2048.0, dets=("H1", "L1", "V1"), masses=(30.0, 25.0),
           distance=600.0):
    rng = np.random.default_rng(seed)
    f = core.frequency_series(fs, T)
    start = 1126259642.413 - T / 2
    inj = dict(mass_1=masses[0], mass_2=masses[1], chi_1=0.1, chi_2=-0.2,
               luminosity_distance=distance, theta_jn=0.5, phase=1.1, ra=1.2, dec=-0.4, psi=0.8,
               geocent_time=start + T / 2)
    pols = core.taylorf2(f, inj["mass_1"], inj["mass_2"], inj["luminosity_distance"],
                         inj["theta_jn"], inj["phase"], chi_1=inj["chi_1"], chi_2=inj["chi_2"],
                         **WF)
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


def _inner(a, b, psd, T):
    return 4 / T * np.sum(np.conj(a) * b / psd)


def _without_last_masked_bin(ev):
    """The event with each detector's last masked bin removed from its mask.
    TaylorF2 runs to the end of the band (lalsimulation's rule for bilby's
    fEnd = maximum_frequency), so the last bin edge is the last masked
    point, which the summary data leave out (relative.py:333-340); the full
    likelihood over this mask is what the binned one equals at the fiducial
    point."""
    dets = []
    for det in ev.detectors:
        mask = det.mask.copy()
        mask[np.nonzero(mask)[0][-1]] = False
        dets.append(OracleDetector(det.name, det.frequency_array, det.strain, det.psd, mask))
    return OracleEvent(dets, ev.duration, ev.sampling_frequency, ev.start_time)


@pytest.mark.parametrize("phase", [False, True])
def test_fiducial_point_equals_full_likelihood(phase):
    ev, inj = _event()
    priors = {"phase": 0.0} if phase else None
    rb = OracleRelativeBinning(ev, "taylorf2", inj, waveform_arguments=WF,
                               phase_marginalization=phase, priors=priors, epsilon=0.05)
    assert rb.maximum_frequency == ev.frequency_array[-1]
    trimmed = _without_last_masked_bin(ev)
    full = OracleGWTransient(trimmed, "taylorf2", waveform_arguments=WF,
                             phase_marginalization=phase,
                             priors={"phase": 0.0} if phase else None)
    p = dict(inj)
    if phase:
        p["phase"] = 0.0
    a, b = rb.log_likelihood_ratio(p), full.log_likelihood_ratio(p)
    assert rb.number_of_bins > 50
    assert abs(a - b) < 1e-8 * max(1.0, abs(b)), (a, b)
    # per detector, the ratio-1 contraction is the full inner products
    for det, tdet in zip(ev.detectors, trimmed.detectors):
        dh, hh = rb.snrs_from_ratio(det, np.ones(len(rb.bin_freqs), complex))
        h0, m = rb.fiducial[det.name], tdet.mask
        assert abs(dh - _inner(h0[m], det.strain[m], det.psd[m], ev.duration)) < 1e-10 * abs(dh)
        assert abs(hh - _inner(h0[m], h0[m], det.psd[m], ev.duration).real) < 1e-10 * hh


def test_summary_contraction_equals_grid_reconstruction():
    ev, inj = _event(seed=4)
    rb = OracleRelativeBinning(ev, "taylorf2", inj, waveform_arguments=WF, epsilon=0.3)
    rng = np.random.default_rng(8)
    x = (rb.bin_freqs - rb.bin_freqs[0]) / (rb.bin_freqs[-1] - rb.bin_freqs[0])
    for det in ev.detectors:
        # a smooth, non-linear complex ratio at the edges
        c = rng.normal(size=4) + 1j * rng.normal(size=4)
        ratio = c[0] + c[1] * x + c[2] * x ** 2 + c[3] * np.sin(3 * x)
        dh, hh = rb.snrs_from_ratio(det, ratio)
        r0 = (ratio[1:] + ratio[:-1]) / 2
        r1 = (ratio[1:] - ratio[:-1]) / np.diff(rb.bin_freqs)
        m = det.mask
        mf = det.frequency_array[m]
        inds = [int(np.searchsorted(mf, e)) for e in rb.bin_freqs]
        if inds[-1] < len(mf) - 1:
            inds[-1] += 1
        h0, d, S = rb.fiducial[det.name][m], det.strain[m], det.psd[m]
        dh_grid, hh_grid, dropped = 0j, 0.0, 0.0
        for i in range(rb.number_of_bins):
            s, e = inds[i], inds[i + 1]
            fc = (mf[s] + mf[e]) / 2
            delta = mf[s:e] - fc
            h = h0[s:e] * (r0[i] + r1[i] * delta)
            dh_grid += _inner(h, d[s:e], S[s:e], ev.duration)
            hh_grid += _inner(h, h, S[s:e], ev.duration).real
            dropped += _inner(h0[s:e] * r1[i] * delta, h0[s:e] * r1[i] * delta, S[s:e],
                              ev.duration).real
        assert abs(dh - dh_grid) < 1e-10 * abs(dh_grid)
        assert abs(hh - (hh_grid - dropped)) < 1e-10 * hh_grid


def test_many_draws_match_full_likelihood():
    """Draws around the fiducial point of a light binary: |binned - full| <
    0.1 |full(ref) - full| (relative_binning_test.py:108-116).  TaylorF2 runs
    smoothly to the end of the band for the fiducial point and every draw,
    so the relative-binning ansatz (a ratio linear within a bin) holds."""
    ev, inj = _event(seed=6, masses=(1.5, 1.3), distance=60.0)
    rb = OracleRelativeBinning(ev, "taylorf2", inj, waveform_arguments=WF, epsilon=0.05)
    full = OracleGWTransient(ev, "taylorf2", waveform_arguments=WF)
    ref = full.log_likelihood_ratio(inj)
    rng = np.random.default_rng(10)
    worst = 0.0
    for _ in range(40):
        p = dict(inj)
        scale = 1 + rng.normal(0, 2e-5)
        p["mass_1"] = inj["mass_1"] * scale
        p["mass_2"] = inj["mass_2"] * scale
        p["ra"] = inj["ra"] + rng.normal(0, 0.05)
        p["dec"] = inj["dec"] + rng.normal(0, 0.05)
        p["geocent_time"] = inj["geocent_time"] + rng.normal(0, 2e-4)
        p["luminosity_distance"] = inj["luminosity_distance"] * rng.uniform(0.8, 1.2)
        regular = full.log_likelihood_ratio(p)
        binned = rb.log_likelihood_ratio(p)
        worst = max(worst, abs(regular - binned) / abs(ref - regular))
    assert worst < 0.1, worst


def test_very_small_epsilon_is_finite():
    ev, inj = _event(seed=7, T=4.0, fs=1024.0, dets=("H1", "L1"))
    rb = OracleRelativeBinning(ev, "taylorf2", inj, waveform_arguments=WF, epsilon=0.001)
    assert np.isfinite(rb.log_likelihood_ratio(inj))


def test_product_host_setup_equals_oracle_loops():
    """bilby_amd.relative's setup_bins / compute_summary_data (vectorised)
    against the oracle's loops, on a stand-in object (no GPU needed)."""
    from bilby_amd.relative import RelativeBinningGravitationalWaveTransient as RB
    cases = [(_event(seed=9), eps) for eps in (0.5, 0.05, 0.003)]
    # a band-limited fiducial whose last edge is the last masked point
    cases.append((_event(seed=9, masses=(1.5, 1.3), distance=60.0), 0.05))
    for (ev, inj), eps in cases:
        orb = OracleRelativeBinning(ev, "taylorf2", inj, waveform_arguments=WF, epsilon=eps)
        ifos = []
        for det in ev.detectors:
            mf = det.frequency_array[det.mask]
            ifos.append(types.SimpleNamespace(
                name=det.name, frequency_array=det.frequency_array, frequency_mask=det.mask,
                minimum_frequency=mf[0], maximum_frequency=mf[-1], duration=ev.duration,
                frequency_domain_strain=det.strain, power_spectral_density_array=det.psd))
        stub = types.SimpleNamespace(
            waveform_generator=types.SimpleNamespace(frequency_array=ev.frequency_array,
                                                     waveform_arguments={}),
            interferometers=ifos, maximum_frequency=orb.maximum_frequency, chi=1, epsilon=eps,
            gamma=np.array([-5 / 3, -2 / 3, 1, 5 / 3, 7 / 3]),
            per_detector_fiducial_waveforms=dict(orb.fiducial),
            per_detector_fiducial_waveform_points={})
        RB.setup_bins(stub)
        assert np.array_equal(stub.bin_freqs, orb.bin_freqs)
        assert np.array_equal(stub.bin_inds, orb.bin_inds)
        RB.compute_summary_data(stub)
        for det in ev.detectors:
            for mine, theirs in zip(stub.summary_data[det.name], orb.summary_data[det.name]):
                scale = np.max(np.abs(theirs))
                assert np.max(np.abs(mine - theirs)) < 1e-12 * scale


class _U:
    def __init__(self, a, b):
        self.minimum, self.maximum = a, b

    def prob(self, x):
        x = np.asarray(x)
        return ((x >= self.minimum) & (x <= self.maximum)) / (self.maximum - self.minimum)

    def rescale(self, u):
        return self.minimum + u * (self.maximum - self.minimum)


def _time_pair(ev, inj, epsilon, jitter_time=True, flags=None, full_event=None, wf=None):
    flags = flags or {}
    wf = WF if wf is None else wf
    tc = inj["geocent_time"]
    pri = {"geocent_time": _U(tc - 0.1, tc + 0.1)}
    if flags.get("phase"):
        pri["phase"] = _U(0, 2 * np.pi)
    rb = OracleRelativeBinning(ev, "taylorf2", inj, waveform_arguments=wf, epsilon=epsilon,
                               time_marginalization=True, jitter_time=jitter_time,
                               phase_marginalization=flags.get("phase", False),
                               priors=dict(pri))
    full = OracleGWTransient(ev if full_event is None else full_event, "taylorf2",
                             waveform_arguments=wf, time_marginalization=True,
                             jitter_time=jitter_time,
                             phase_marginalization=flags.get("phase", False), priors=dict(pri))
    return rb, full, pri


@pytest.mark.parametrize("phase", [False, True])
def test_time_marginalised_at_fiducial_equals_full(phase):
    """At the fiducial point (geocent_time = start, no jitter) every edge
    ratio is 1, the rebuilt waveform (relative.py:376-391) is the fiducial
    response itself on the binned range, and the time-marginalised binned
    lnL equals the full likelihood's (base.py:325-330, 794-820) up to
    round-off.  The last bin edge is the last masked point, which the summary
    data leave out (relative.py:336-340), so the full likelihood is taken
    over the masks without it."""
    ev, inj = _event(seed=6, masses=(2.6, 2.2), distance=100.0)
    rb, full, _ = _time_pair(ev, inj, 0.1, flags={"phase": phase},
                             full_event=_without_last_masked_bin(ev))
    p = dict(inj, geocent_time=ev.start_time, time_jitter=0.0)
    if phase:
        p["phase"] = 0.0
    a, b = rb.log_likelihood_ratio(p), full.log_likelihood_ratio(p)
    assert b > 20
    assert abs(a - b) < 1e-8 * abs(b), (a, b)


def test_time_marginalised_matches_brute_force_integral():
    """The reference's own check (marginalization_test.py:392-413, 416-433
    for relbin): the time-marginalised binned lnL against
    log trapz(exp(lnL(t)) prior(t)) of the un-marginalised binned lnL over
    1000 times of the prior, within 0.5.  As there, a 31 + 29 Msun binary at
    4000 Mpc.  The check holds only where the transform's grid (2 / f_s)
    resolves the likelihood's peak in time; TaylorF2 as lalsimulation
    computes it keeps f^-7/6 power up to maximum_frequency (the reference's
    IMRPhenomPv2 signal fades after merger), so the waveform band ends at
    128 Hz here."""
    wf = dict(WF, maximum_frequency=128.0)
    ev, inj = _event(seed=6, masses=(31.0, 29.0), distance=4000.0)
    rb_t, _, pri = _time_pair(ev, inj, 0.1, jitter_time=False, wf=wf)
    rb = OracleRelativeBinning(ev, "taylorf2", inj, waveform_arguments=wf, epsilon=0.1)
    tp = pri["geocent_time"]
    ts = np.linspace(tp.minimum, tp.maximum, 1000)
    lls = np.array([rb.log_likelihood_ratio(dict(inj, geocent_time=t)) for t in ts])
    brute = np.log(np.trapezoid(np.exp(lls - lls.max()) * tp.prob(ts), ts)) + lls.max()
    marg = rb_t.log_likelihood_ratio(dict(inj, geocent_time=ev.start_time))
    assert lls.max() > 15
    assert abs(marg - brute) < 0.5, (marg, brute)


def test_time_marginalised_draws_near_fiducial():
    """Draws around the fiducial point with time jitter: the time-
    marginalised binned lnL follows the full one (the bound of
    relative_binning_test.py:108-116, 0.1 |full(ref) - full|, plus 0.05 for
    the jitter's linear interpolation across bins)."""
    ev, inj = _event(seed=6, masses=(1.6, 1.3), distance=60.0)
    rb, full, _ = _time_pair(ev, inj, 0.05)
    rng = np.random.default_rng(9)
    ref = full.log_likelihood_ratio(dict(inj, geocent_time=ev.start_time, time_jitter=0.0))
    for _ in range(6):
        p = dict(inj, geocent_time=ev.start_time,
                 time_jitter=rng.uniform(-1, 1) / ev.sampling_frequency)
        p["mass_1"] *= 1 + rng.normal(0, 1e-4)
        p["ra"] += rng.normal(0, 0.02)
        a, b = rb.log_likelihood_ratio(p), full.log_likelihood_ratio(p)
        assert abs(a - b) < 0.1 * abs(ref - b) + 0.05, (a, b, ref)


@pytest.mark.parametrize("phase", [False, True])
def test_time_reconstruction_at_fiducial_equals_full(phase):
    """The time posterior of the reconstruction (base.py:580-640, inherited
    by relative binning with relative.py:380-393 as the waveform) at the
    fiducial point: every edge ratio is 1, the rebuilt waveform is the
    fiducial response on the binned range, and the posterior on the
    1/16384 s grid equals the full likelihood's."""
    ev, inj = _event(seed=6, masses=(2.6, 2.2), distance=100.0)
    rb, full, _ = _time_pair(ev, inj, 0.1, flags={"phase": phase})
    p = dict(inj, geocent_time=ev.start_time, time_jitter=0.0)
    if phase:
        p["phase"] = 0.0
    t_a, post_a = rb.time_posterior(p)
    t_b, post_b = full.time_posterior(p)
    assert np.array_equal(t_a, t_b) and len(t_a) > 1000
    assert np.max(np.abs(post_a - post_b)) < 1e-8 * np.max(post_b)
    # away from the fiducial point the rebuilt waveform is an approximation
    q = dict(p, geocent_time=ev.start_time + 2e-4, mass_1=inj["mass_1"] * (1 + 1e-4))
    t_a, post_a = rb.time_posterior(q)
    t_b, post_b = full.time_posterior(q)
    assert abs(t_a[np.argmax(post_a)] - t_b[np.argmax(post_b)]) < 2 / 16384
