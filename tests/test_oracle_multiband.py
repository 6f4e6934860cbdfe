"""The multibanded likelihood's host setup (bilby_amd.multiband.MultibandSetup)
and its oracle evaluation (oracle/multiband.py) pinned on the CPU:
* multibanded ln L vs the full oracle ln L on the reference test's
  configuration (16 s, 2048 Hz, H1+L1+V1, geocent prior +-0.1 s, chirp mass
  6, mass ratio 0.5, network SNR ~50, noise seed 70817;
  likelihood_test.py:1313-1428): |dlnL| < 1.5e-2 (its tolerance for the
  linear-interpolation algorithm); TaylorF2 runs on to the band's end as
  lalsimulation's does for bilby's fEnd (the reference uses IMRPhenomD);
  with the analytic PSD here, 280 Mpc gives the reference's SNR ~50;
* accuracy_factor 50 at least halves the error of accuracy_factor 5
  (likelihood_test.py:1430-1475), same configuration;
* the prefix-sum quadratic coefficients equal the reference's loop form;
* weights saved and loaded give the same setup."""
import types

import numpy as np
import pytest

from oracle import core
from oracle.likelihood import OracleDetector, OracleEvent, OracleGWTransient, detector_response
from oracle.multiband import OracleMultiband, quadratic_coefficients_loop

WF = dict(reference_frequency=20.0, minimum_frequency=20.0)
T, FS = 16.0, 2048.0


# likelihood_test.py:1317-1333: chirp_mass 6, mass_ratio 0.5
_MC, _Q = 6.0, 0.5
REF_MASSES = (_MC * (1 + _Q) ** 0.2 / _Q ** 0.6, _Q * _MC * (1 + _Q) ** 0.2 / _Q ** 0.6)


def _event(seed=70817, masses=REF_MASSES, distance=280.0):
    rng = np.random.default_rng(seed)
    f = core.frequency_series(FS, T)
    tc = 1187008882.0
    start = tc - T + 2.0
    inj = dict(mass_1=masses[0], mass_2=masses[1], chi_1=0.0, chi_2=0.0,
               luminosity_distance=distance, theta_jn=0.4, phase=1.3, ra=1.3, dec=-1.2, psi=0.659,
               geocent_time=tc)
    pols = core.taylorf2(f, inj["mass_1"], inj["mass_2"], inj["luminosity_distance"],
                         inj["theta_jn"], inj["phase"], **WF)
    dets, ifos = [], []
    for name in ("H1", "L1", "V1"):
        psd = core.analytic_aligo_psd(f)
        noise, _ = core.noise_realisation(lambda x: core.analytic_aligo_psd(x), 0, np.inf,
                                          len(f), T, rng)
        mask = core.frequency_mask(f, 20, FS / 2, FS)
        det = OracleDetector(name, f, noise, psd, mask)
        det.strain = det.strain + detector_response(det, pols, inj, start) * mask
        dets.append(det)
        ifos.append(types.SimpleNamespace(
            name=name, frequency_array=f, frequency_mask=mask, frequency_domain_strain=det.strain,
            power_spectral_density_array=psd, minimum_frequency=20.0, maximum_frequency=FS / 2))
    return OracleEvent(dets, T, FS, start), ifos, inj


def _setup(ifos, inj, **kw):
    from bilby_amd.constants import radius_of_earth, speed_of_light
    from bilby_amd.multiband import MultibandSetup
    start = inj["geocent_time"] - T + 2.0
    safety = radius_of_earth / speed_of_light
    time_offset = start + T - (inj["geocent_time"] - 0.1) + safety     # multiband.py:194-225
    delta_f_end = 100 / (start + T - (inj["geocent_time"] + 0.1) - safety)
    mc = (inj["mass_1"] * inj["mass_2"]) ** 0.6 / (inj["mass_1"] + inj["mass_2"]) ** 0.2
    return MultibandSetup(ifos, T, mc, time_offset=time_offset, delta_f_end=delta_f_end, **kw)


def test_multiband_matches_full_likelihood():
    ev, ifos, inj = _event()
    setup = _setup(ifos, inj)
    assert setup.number_of_bands >= 3
    assert len(setup.waveform_frequencies) < 0.5 * len(ev.frequency_array)
    mb = OracleMultiband(ev, "taylorf2", setup, waveform_arguments=WF)
    full = OracleGWTransient(ev, "taylorf2", waveform_arguments=WF)
    ll, llmb = full.log_likelihood_ratio(inj), mb.log_likelihood_ratio(inj)
    assert ll > 100
    assert abs(ll - llmb) < 1.5e-2, (ll, llmb)
    rng = np.random.default_rng(3)
    for _ in range(5):
        p = dict(inj)
        p["ra"] += rng.normal(0, 0.05)
        p["luminosity_distance"] *= rng.uniform(0.8, 1.2)
        p["geocent_time"] += rng.normal(0, 1e-3)
        assert abs(full.log_likelihood_ratio(p) - mb.log_likelihood_ratio(p)) < 1.5e-2


def test_larger_accuracy_factor_is_more_accurate():
    ev, ifos, inj = _event()
    full = OracleGWTransient(ev, "taylorf2", waveform_arguments=WF).log_likelihood_ratio(inj)
    errs = []
    for acc in (5, 50):
        setup = _setup(ifos, inj, accuracy_factor=acc)
        mb = OracleMultiband(ev, "taylorf2", setup, waveform_arguments=WF)
        errs.append(abs(full - mb.log_likelihood_ratio(inj)))
    assert errs[1] < errs[0] / 2, errs


def test_quadratic_coefficients_equal_loop_form_and_weights_round_trip(tmp_path):
    from bilby_amd.multiband import MultibandSetup
    ev, ifos, inj = _event(seed=6)
    setup = _setup(ifos, inj)
    loop = quadratic_coefficients_loop(setup, ifos)
    for name, q in loop.items():
        assert np.max(np.abs(setup.quadratic_coeffs[name] - q)) < 1e-12 * np.max(np.abs(q))
    fn = str(tmp_path / "mb.npz")
    setup.save_weights(fn)
    back = MultibandSetup.from_weights(fn, ifos)
    for key in ("banded_frequency_points", "unique_to_original_frequencies", "durations", "Ks_Ke"):
        assert np.array_equal(getattr(back, key), getattr(setup, key))
    for name in setup.linear_coeffs:
        assert np.array_equal(back.linear_coeffs[name], setup.linear_coeffs[name])
        assert np.array_equal(back.quadratic_coeffs[name], setup.quadratic_coeffs[name])


class _U:
    def __init__(self, a, b):
        self.minimum, self.maximum = a, b

    def prob(self, x):
        x = np.asarray(x)
        return ((x >= self.minimum) & (x <= self.maximum)) / (self.maximum - self.minimum)

    def rescale(self, u):
        return self.minimum + u * (self.maximum - self.minimum)


def test_time_marginalised_array_matches_direct_transform():
    """The multibanded d_inner_h_array (multiband.py:792-800: the h L terms
    scattered to int(f T_0) and transformed with N = Nbs[-1] / 2) against a
    direct DFT of the full-grid integrand 4/T h conj(d) / S at the same times
    t_j = j T_0 / N (antenna patterns at the prior's middle, as
    multiband.py:724-726 sets): within 1e-3 of the peak |X| at the window
    bins around the peak, and the peak at the same bin."""
    from oracle.likelihood import source_polarizations
    ev, ifos, inj = _event()
    setup = _setup(ifos, inj)
    tc = inj["geocent_time"]
    pri = {"geocent_time": _U(tc - 0.1, tc + 0.1)}
    mb = OracleMultiband(ev, "taylorf2", setup, waveform_arguments=WF, time_marginalization=True,
                         priors=dict(pri))
    p = dict(inj, geocent_time=ev.start_time, time_jitter=0.0)
    wa = dict(WF, maximum_frequency=ev.frequency_array[-1])
    pols = source_polarizations("taylorf2", mb.w["waveform_frequencies"], p, wa)
    pols_full = source_polarizations("taylorf2", ev.frequency_array, p, WF)
    N = len(mb._times)
    assert N == setup.Nbs[-1] // 2
    j_peak = int(round((tc - ev.start_time) / mb._delta_tc))
    js = np.arange(j_peak - 40, j_peak + 41)
    x_mb = 0
    x_direct = np.zeros(len(js), dtype=complex)
    for det in ev.detectors:
        _, _, arr = mb.calculate_snrs_mb(pols, det, p)
        x_mb = x_mb + arr[js]
        sig = detector_response(det, pols_full, p, ev.start_time, antenna_time=mb.antenna_time)
        m = det.mask
        integrand = 4 / T * sig[m] * np.conj(det.strain[m]) / det.psd[m]
        f = det.frequency_array[m]
        x_direct += np.exp(-2j * np.pi * np.outer(js * mb._delta_tc, f)) @ integrand
    scale = np.max(np.abs(x_direct))
    assert scale > 100
    assert np.argmax(np.abs(x_mb)) == np.argmax(np.abs(x_direct))
    assert np.max(np.abs(x_mb - x_direct)) < 1e-3 * scale, np.max(np.abs(x_mb - x_direct)) / scale


def test_time_grid_index_follows_fancy_index_accumulation():
    """The product's map from banded points to transform bins
    (MBGravitationalWaveTransient._setup_time_marginalization_multiband,
    used by the GEMM) reproduces the oracle's literal band-by-band
    ``a[idx] += v`` (multiband.py:792-798) -- on the real setup and on one
    with indices repeated inside a band (numpy keeps the last point's term
    there) and across bands (terms add)."""
    from bilby_amd.multiband import MBGravitationalWaveTransient
    ev, ifos, inj = _event()
    real = _setup(ifos, inj)
    tc = inj["geocent_time"]
    rng = np.random.default_rng(2)
    fake_points = np.array([20.0, 20.0, 20.5, 21.0, 21.0, 21.0, 22.0, 20.5, 22.0, 23.0]) / 1.0
    fake = types.SimpleNamespace(Nbs=np.array([64, 128]), durations=np.array([2.0, 1.0]),
                                 banded_frequency_points=fake_points,
                                 start_end_idxs=np.array([[0, 6], [7, 9]]), number_of_bands=2)
    for setup in (real, fake):
        obj = types.SimpleNamespace(setup=setup, priors={"geocent_time": _U(tc - 0.1, tc + 0.1)},
                                    interferometers=types.SimpleNamespace(start_time=tc - 1.0),
                                    _delta_tc=2 / FS)
        if setup is fake:
            obj.priors = {"geocent_time": _U(tc - 0.9, tc - 0.5)}
        MBGravitationalWaveTransient._setup_time_marginalization_multiband(obj)
        n = len(setup.banded_frequency_points)
        y = rng.normal(size=n) + 1j * rng.normal(size=n)
        N = int(setup.Nbs[-1]) // 2
        literal = np.zeros(N, dtype=complex)
        fm = [int(f * setup.durations[0]) for f in setup.banded_frequency_points]
        for b in range(len(setup.durations)):
            s, e = setup.start_end_idxs[b]
            literal[fm[s:e + 1]] += y[s:e + 1]
        idx = obj._time_grid_index
        gemm_form = np.zeros(N, dtype=complex)
        np.add.at(gemm_form, idx[idx >= 0], y[idx >= 0])
        assert np.allclose(gemm_form, literal, rtol=0, atol=1e-14)
        if setup is fake:
            assert (idx < 0).sum() == 3   # 20.0, 21.0, 21.0 dropped in band 0
        N_, j_first, n_window = obj._time_window
        assert N_ == N and 0 <= j_first and j_first + n_window <= N


# ------------------------------------------------------------------ IFFT-FFT <h|h>

def _band_strain(mb, det, params):
    """The detector response at the banded points, as the oracle's
    calculate_snrs_mb forms it."""
    from oracle.likelihood import source_polarizations
    w = mb.w
    wa = dict(WF, maximum_frequency=mb.event.frequency_array[-1])
    pols = source_polarizations("taylorf2", w["waveform_frequencies"], params, wa)
    idx = w["unique_to_original_frequencies"]
    pts = types.SimpleNamespace(tensor=det.tensor, vertex=det.vertex,
                                frequency_array=w["banded_frequency_points"],
                                mask=np.ones(len(idx), dtype=bool))
    return detector_response(pts, {k: v[idx] for k, v in pols.items()}, params,
                             mb.event.start_time)


def test_ifft_fft_matches_full_likelihood():
    """linear_interpolation=False on the reference test's configuration:
    |dlnL| against the full likelihood below 5e-3, the reference's tolerance
    for the IFFT-FFT algorithm (likelihood_test.py:1380-1386), and the
    setup flags and arrays as multiband.py:613-646 leaves them."""
    ev, ifos, inj = _event()
    setup = _setup(ifos, inj, linear_interpolation=False)
    assert setup.linear_interpolation is False and not hasattr(setup, "quadratic_coeffs")
    nb = setup.number_of_bands
    assert nb >= 3 and len(setup.Tbhats) == nb and len(setup.windows) == len(
        setup.banded_frequency_points)
    for b in range(nb):
        nhat = min(2 * setup.Mbs[b], setup.Nbs[b])
        assert setup.Tbhats[b] == T * nhat / setup.Nbs[b]
        assert len(setup.Ibcs["H1"][b]) == nhat // 2 + 1
    mb = OracleMultiband(ev, "taylorf2", setup, waveform_arguments=WF)
    full = OracleGWTransient(ev, "taylorf2", waveform_arguments=WF)
    rng = np.random.default_rng(3)
    for k in range(6):
        p = dict(inj)
        if k:
            p["ra"] += rng.normal(0, 0.05)
            p["luminosity_distance"] *= rng.uniform(0.8, 1.2)
            p["geocent_time"] += rng.normal(0, 1e-3)
        ll, llmb = full.log_likelihood_ratio(p), mb.log_likelihood_ratio(p)
        assert ll > 100
        assert abs(ll - llmb) < 5e-3, (k, ll, llmb)


def test_ifft_fft_band_sums_equal_the_lag_form():
    """Each band's IFFT-FFT sum against the quadratic form it equals by
    Parseval, written without transforms of the waveform:
      sum_{k <= M_b} Re Ibc_k |thbc_k|^2
        = (N^ sum_{n,m} x_n x_m I[|n - m|] + Ibc_0 t_0^2 + Ibc_M t_M^2) / 2
    with x the band's time samples, I the inverse PSD's lags (the inverse
    real transform of 1/S truncated at the band's Nyquist bin), t_0 = sum x,
    t_M = sum (-1)^(n + M) x_n and Ibc_0, Ibc_M the sums of the kept lags
    with signs 1 and (-1)^n.  The double sum runs over the autocorrelation of
    x (np.correlate), so a wrong lag window, band placement or half-spectrum
    weighting in the restatement shows as a mismatch."""
    ev, ifos, inj = _event()
    setup = _setup(ifos, inj, linear_interpolation=False)
    mb = OracleMultiband(ev, "taylorf2", setup, waveform_arguments=WF)
    for det, ifo in zip(ev.detectors, ifos):
        strain = _band_strain(mb, det, inj)
        inv = np.zeros(int(setup.Nbs[-1]) // 2 + 1)
        inv[:len(det.psd)][det.mask] = 1 / det.psd[det.mask]
        s0, e0 = setup.start_end_idxs[0]
        k0, k1 = setup.Ks_Ke[0]
        lag_form = 4 / T * np.sum(np.abs(strain[s0:e0 + 1]) ** 2 * det.mask[k0:k1 + 1]
                                  * setup.windows[s0:e0 + 1] / det.psd[k0:k1 + 1])
        for b in range(1, setup.number_of_bands):
            s, e = setup.start_end_idxs[b]
            ks, ke = setup.Ks_Ke[b]
            m = int(setup.Mbs[b])
            nhat = min(2 * m, int(setup.Nbs[b]))
            wth = np.zeros(m // 2 + 1, dtype=complex)
            wth[ks:ke + 1] = setup.square_root_windows[s:e + 1] * strain[s:e + 1]
            x = np.fft.irfft(wth)
            lag = np.fft.irfft(inv[:int(setup.Nbs[b]) // 2 + 1])
            ac = np.correlate(x, x, mode="full")[m - 1:]          # sum_n x_n x_{n+d}, d >= 0
            quad = lag[0] * ac[0] + 2 * np.dot(lag[1:m], ac[1:m])
            kept = lag[np.minimum(np.arange(nhat), nhat - np.arange(nhat))]
            sign = (-1.0) ** np.arange(nhat)
            t0, tm = np.sum(x), (-1.0) ** m * np.dot((-1.0) ** np.arange(m), x)
            S = (nhat * quad + np.sum(kept) * t0 ** 2 + np.dot(sign, kept) * tm ** 2) / 2
            lag_form += 4 / setup.Tbhats[b] * S
        oracle = mb.ifft_fft_optimal_snr_squared(strain, det)
        assert oracle > 10
        assert abs(oracle - lag_form) < 1e-10 * oracle, (det.name, oracle, lag_form)


def test_ifft_fft_device_tables_reproduce_the_oracle():
    """The device split of the IFFT-FFT <h|h> (MultibandSetup.ifft_fft_tables:
    diagonal coefficients on the banded points, odd-bin weights per band)
    evaluated with the GPU kernels' algebra in numpy -- packed half-length
    inverse transform, half-bin modulation, forward transform, even / odd
    separation -- equals the oracle's <h|h> per detector."""
    ev, ifos, inj = _event()
    setup = _setup(ifos, inj, linear_interpolation=False)
    mb = OracleMultiband(ev, "taylorf2", setup, waveform_arguments=WF)
    diag, bands = setup.ifft_fft_tables()
    assert len(bands) == setup.number_of_bands - 1
    for di, det in enumerate(ev.detectors):
        strain = _band_strain(mb, det, inj)
        hh = np.sum(diag[det.name] * np.abs(strain) ** 2)
        for band in bands:
            mh, p0, L, ks = (band[k] for k in ("half_length", "first_point", "n_points",
                                              "first_bin"))
            X = np.zeros(mh + 1, dtype=complex)
            X[ks:ks + L] = band["sqrt_window"] * strain[p0:p0 + L]
            k = np.arange(mh)
            xm = np.conj(X[mh - k])
            zc = (X[k] + xm) / 2 + 1j * np.exp(1j * np.pi * k / mh) * (X[k] - xm) / 2
            u = np.fft.ifft(zc) * mh * np.exp(-1j * np.pi * k / mh)
            U = np.fft.fft(u)
            ur = np.conj(U[mh - 1 - k])
            Y = ((U + ur) / 2 + np.exp(-1j * np.pi * (2 * k + 1) / (2 * mh)) * (U - ur) / 2j) / mh
            hh += np.sum(band["odd_weights"][di] * np.abs(Y) ** 2)
        oracle = mb.ifft_fft_optimal_snr_squared(strain, det)
        assert abs(hh - oracle) < 1e-11 * oracle, (det.name, hh, oracle)


def test_ifft_fft_weights_round_trip(tmp_path):
    """IFFT-FFT weights saved and read back (multiband.py:648-712): the
    algorithm flag, T^_b, the windows and each detector's Ibc per band."""
    from bilby_amd.multiband import MultibandSetup
    ev, ifos, inj = _event(seed=6)
    setup = _setup(ifos, inj, linear_interpolation=False)
    w = setup.weights
    assert w["linear_interpolation"] is False and "quadratic_coeffs" not in w
    assert set(w["Ibcs"]["L1"]) == {str(b) for b in range(setup.number_of_bands)}
    fn = str(tmp_path / "mb_ifft.npz")
    setup.save_weights(fn)
    back = MultibandSetup.from_weights(fn, ifos)
    assert back.linear_interpolation is False
    assert np.array_equal(back.Tbhats, setup.Tbhats)
    assert np.array_equal(back.square_root_windows, setup.square_root_windows)
    for name in setup.Ibcs:
        for b in range(setup.number_of_bands):
            assert np.array_equal(back.Ibcs[name][b], setup.Ibcs[name][b])
    d0, b0 = setup.ifft_fft_tables()
    d1, b1 = back.ifft_fft_tables()
    for name in d0:
        assert np.array_equal(d0[name], d1[name])
    for x, y in zip(b0, b1):
        assert np.array_equal(x["odd_weights"], y["odd_weights"])
    with pytest.raises(TypeError):
        _setup(ifos, inj, linear_interpolation=1)
