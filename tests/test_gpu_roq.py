"""GPU parity of the ROQ path: the MFMA GEMM kernels against numpy, and the
ROQ log-likelihood (roq.py) against the FP64 oracle."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def _split_bf16(x):
    import torch
    t = torch.from_numpy(x.astype(np.float32))
    hi = t.to(torch.bfloat16)
    lo = (t - hi.to(torch.float32)).to(torch.bfloat16)
    return hi, lo


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_gemm_modes_match_numpy(mode):
    """C = A Bt^T on MFMA (fp32 32x32x2, split-bf16 x3 32x32x16) and the
    float64 reference GEMM vs numpy: relative Frobenius error < 1e-14 (fp64),
    2e-6 (fp32), 3e-5 (bf16x3)."""
    import torch
    from bilby_amd._lib import LIB, check
    from bilby_amd.likelihood import get_context
    ctx = get_context(0)
    say("gemm mode", mode)
    rng = np.random.default_rng(mode)
    M, N, K = 256, 384, 96
    A = rng.normal(size=(M, K))
    Bt = rng.normal(size=(N, K))
    ref = A @ Bt.T
    dev = torch.device("cuda:0")
    if mode == 0:
        a = torch.from_numpy(A).to(dev)
        b = torch.from_numpy(Bt).to(dev)
        c = torch.empty((M, N), dtype=torch.float64, device=dev)
        check(LIB.gwl_gemm(ctx, 0, ctypes.c_void_p(a.data_ptr()), None, ctypes.c_void_p(b.data_ptr()),
                           None, ctypes.c_void_p(c.data_ptr()), M, N, K, None))
    elif mode == 1:
        a = torch.from_numpy(A.astype(np.float32)).to(dev)
        b = torch.from_numpy(Bt.astype(np.float32)).to(dev)
        c = torch.empty((M, N), dtype=torch.float32, device=dev)
        check(LIB.gwl_gemm(ctx, 1, ctypes.c_void_p(a.data_ptr()), None, ctypes.c_void_p(b.data_ptr()),
                           None, ctypes.c_void_p(c.data_ptr()), M, N, K, None))
    else:
        ah, al = (t.to(dev) for t in _split_bf16(A))
        bh, bl = (t.to(dev) for t in _split_bf16(Bt))
        c = torch.empty((M, N), dtype=torch.float32, device=dev)
        check(LIB.gwl_gemm(ctx, 2, ctypes.c_void_p(ah.data_ptr()), ctypes.c_void_p(al.data_ptr()),
                           ctypes.c_void_p(bh.data_ptr()), ctypes.c_void_p(bl.data_ptr()),
                           ctypes.c_void_p(c.data_ptr()), M, N, K, None))
    torch.cuda.synchronize()
    out = c.cpu().numpy().astype(np.float64)
    err = np.linalg.norm(out - ref) / np.linalg.norm(ref)
    tol = {0: 1e-14, 1: 2e-6, 2: 3e-5}[mode]
    say("gemm err", err)
    assert err < tol, err


def test_gemm_bf16x3_wide_tiles_match_narrow_bit_for_bit():
    """N a multiple of 256 runs the 128 x 256-tile split-bfloat16 kernel, other
    N the 128 x 128 one; an output element's products and their order are the
    same in both, so the first 384 columns of an N = 512 product equal the
    N = 384 product bit for bit (and numpy within 3e-5)."""
    import torch
    from bilby_amd._lib import LIB, check
    from bilby_amd.likelihood import get_context
    ctx = get_context(0)
    say("gemm bf16x3 wide vs narrow tiles")
    rng = np.random.default_rng(11)
    M, K = 384, 160
    A = rng.normal(size=(M, K))
    Bt = rng.normal(size=(512, K))
    dev = torch.device("cuda:0")
    ah, al = (t.to(dev) for t in _split_bf16(A))
    outs = {}
    for N in (512, 384):
        bh, bl = (t.to(dev).contiguous() for t in _split_bf16(Bt[:N]))
        c = torch.empty((M, N), dtype=torch.float32, device=dev)
        check(LIB.gwl_gemm(ctx, 2, ctypes.c_void_p(ah.data_ptr()), ctypes.c_void_p(al.data_ptr()),
                           ctypes.c_void_p(bh.data_ptr()), ctypes.c_void_p(bl.data_ptr()),
                           ctypes.c_void_p(c.data_ptr()), M, N, K, None))
        torch.cuda.synchronize()
        outs[N] = c.cpu().numpy()
    assert np.array_equal(outs[512][:, :384], outs[384])
    ref = A @ Bt.T
    err = np.linalg.norm(outs[512].astype(np.float64) - ref) / np.linalg.norm(ref)
    say("gemm wide err", err)
    assert err < 3e-5, err


def _roq_setup():
    from bilby_amd.configs import make_roq
    return make_roq(n_draws=128, seed=77)


@pytest.mark.parametrize("time,phase", [(True, True), (True, False), (False, True),
                                        (False, False)])
def test_roq_lnl_matches_oracle(time, phase):
    """ROQ, H1+L1, 400/120 nodes, 1000 ROQ time samples, time and/or phase
    marginalised or neither (geocent_time drawn within the prior):
    fp64 |dlnL| < 1e-8 against the FP64 oracle, fp32 MFMA < 1e-3; split-bf16
    x3 is reported and bounded by 0.05."""
    from bilby_amd.roq import ROQGravitationalWaveTransient
    from oracle.roq import OracleROQ
    from tests.helpers import fixed_marginalised, draws_as_dicts
    say("roq time", time, "phase", phase)
    s = _roq_setup()
    w = s["weights"]
    opri = dict(s["priors"])
    o = OracleROQ(["H1", "L1"], s["interferometers"].start_time, w, w["frequency_nodes_linear"],
                  w["frequency_nodes_quadratic"], opri, time_marginalization=time,
                  phase_marginalization=phase, delta_tc=s["delta_tc"],
                  waveform_arguments=s["waveform_generator"].waveform_arguments)
    priors = s["priors"].copy()
    lk = ROQGravitationalWaveTransient(s["interferometers"], s["waveform_generator"], priors,
                                       weights=w, time_marginalization=time,
                                       phase_marginalization=phase, delta_tc=s["delta_tc"])
    draws = fixed_marginalised(s["draws"], priors)
    n = 64
    ref = np.array([o.log_likelihood_ratio(p) for p in draws_as_dicts(draws, range(n))])
    d = {k: v[:n] for k, v in draws.items()}
    out = {m: lk.log_likelihood_ratio_batch(d, precision=m) for m in ("fp64", "fp32", "bf16x3")}
    errs = {m: float(np.max(np.abs(v - ref))) for m, v in out.items()}
    say("roq errors", errs, "lnL range", ref.min(), ref.max())
    assert np.all(np.isfinite(ref))
    assert errs["fp64"] < 1e-8
    assert errs["fp32"] < 1e-3
    assert errs["bf16x3"] < 1e-3


def test_roq_from_bases_matches_oracle(tmp_path):
    """ROQGravitationalWaveTransient from linear_matrix / quadratic_matrix
    (roq.py:146-190, 334-366, 736-1005): the bases as arrays [n_freq, n_basis]
    and as a .npy file [n_basis, n_freq]; the weights it builds (time
    resolution, time samples, linear weights on the GPU) against the oracle's
    restatement of the weight sums, and its lnL (time + phase marginalised)
    against the ROQ oracle on those weights."""
    from bilby_amd.configs import make_taylorf2_marginalised
    from bilby_amd.roq import ROQGravitationalWaveTransient
    from oracle import roq_weights as orw
    from oracle.roq import OracleROQ
    from tests.helpers import draws_as_dicts, fixed_marginalised
    say("roq from bases")
    s = make_taylorf2_marginalised(n_draws=16, seed=31, duration=4.0, sampling_frequency=1024.0,
                                   detectors=("H1", "L1"), time=True, distance=False, phase=True)
    ifos, wfg = s["interferometers"], s["waveform_generator"]
    rng = np.random.default_rng(5)
    n_m = int(ifos[0].frequency_mask.sum())
    L, Q = 40, 12
    lin = np.linalg.qr(rng.normal(size=(n_m, L)) + 1j * rng.normal(size=(n_m, L)))[0]
    quad = np.abs(rng.normal(size=(n_m, Q)))
    wfg.waveform_arguments.update(frequency_nodes_linear=np.sort(rng.uniform(20, 480, L)),
                                  frequency_nodes_quadratic=np.sort(rng.uniform(20, 480, Q)))
    opri = dict(s["priors"])
    priors = s["priors"].copy()
    np.save(tmp_path / "quad.npy", quad.T)
    lk = ROQGravitationalWaveTransient(ifos, wfg, priors, linear_matrix=lin,
                                       quadratic_matrix=str(tmp_path / "quad.npy"),
                                       time_marginalization=True, phase_marginalization=True,
                                       precision="fp64")
    w = lk.weights
    ts = w["time_samples"]
    for ifo in ifos:
        ref = orw.linear_weights(ifo.frequency_array, ifo.frequency_mask,
                                 ifo.frequency_domain_strain, ifo.power_spectral_density_array,
                                 ifo.minimum_frequency, ifos.duration, lin.T, ts)
        err = np.max(np.abs(w[ifo.name + "_linear"] - ref)) / np.max(np.abs(ref))
        q_ref = orw.quadratic_weights(ifo.frequency_mask, ifo.power_spectral_density_array,
                                      ifos.duration, quad.T)
        say(ifo.name, "linear weights relative error", err, "samples", len(ts))
        assert err < 1e-10
        assert np.allclose(w[ifo.name + "_quadratic"], q_ref, rtol=1e-13, atol=0)
    o = OracleROQ([i.name for i in ifos], ifos.start_time, w, w["frequency_nodes_linear"],
                  w["frequency_nodes_quadratic"], opri, time_marginalization=True,
                  phase_marginalization=True, delta_tc=ts[1] - ts[0],
                  waveform_arguments=wfg.waveform_arguments)
    draws = fixed_marginalised(s["draws"], priors)
    n = 16
    ref = np.array([o.log_likelihood_ratio(p) for p in draws_as_dicts(draws, range(n))])
    got = lk.log_likelihood_ratio_batch({k: v[:n] for k, v in draws.items()})
    say("lnL max |diff|", float(np.max(np.abs(got - ref))), "range", ref.min(), ref.max())
    assert np.all(np.isfinite(ref)) and np.max(np.abs(got - ref)) < 1e-8 * max(1.0, np.max(np.abs(ref)))
    with pytest.raises(ValueError):
        ROQGravitationalWaveTransient(ifos, wfg, s["priors"].copy(), linear_matrix=lin)
    # roq.py:1055-1160: the weights saved as npz and read back give the same likelihood
    lk.save_weights(str(tmp_path / "w"), format="npz")
    lk2 = ROQGravitationalWaveTransient(ifos, wfg, s["priors"].copy(),
                                        weights=str(tmp_path / "w.npz"),
                                        time_marginalization=True, phase_marginalization=True,
                                        precision="fp64")
    assert np.array_equal(lk2.log_likelihood_ratio_batch({k: v[:n] for k, v in draws.items()}),
                          got)
    again = lk.load_weights(str(tmp_path / "w.npz"))
    assert np.array_equal(again["H1_linear"], w["H1_linear"])
    with pytest.raises(NotImplementedError):
        lk.save_weights(str(tmp_path / "w"))


def _two_basis_weights():
    """Weights with two linear and two quadratic bases (different node sets,
    seeded synthetic weights on the same time samples) and prior ranges on
    the chirp mass that overlap, as bilby's multi-basis files hold them."""
    from bilby_amd.configs import make_roq
    a = make_roq(n_draws=256, seed=77, n_linear=120, n_quadratic=40, n_time_samples=400)
    b = make_roq(n_draws=8, seed=78, n_linear=90, n_quadratic=56, n_time_samples=400)
    wa, wb = a["weights"], b["weights"]
    assert np.array_equal(wa["time_samples"], wb["time_samples"])
    w = {"time_samples": wa["time_samples"],
         "frequency_nodes_linear": [wa["frequency_nodes_linear"], wb["frequency_nodes_linear"]],
         "frequency_nodes_quadratic": [wa["frequency_nodes_quadratic"],
                                       wb["frequency_nodes_quadratic"]],
         "prior_range_linear": {"chirp_mass": np.array([[4.0, 20.0], [18.0, 60.0]])},
         "prior_range_quadratic": {"chirp_mass": np.array([[4.0, 30.0], [25.0, 60.0]])}}
    for name in ("H1", "L1"):
        for kind in ("linear", "quadratic"):
            w["%s_%s" % (name, kind)] = [wa["%s_%s" % (name, kind)], wb["%s_%s" % (name, kind)]]
    return a, w


def _pair_weights(w, bl, bq):
    out = {"time_samples": w["time_samples"],
           "frequency_nodes_linear": w["frequency_nodes_linear"][bl],
           "frequency_nodes_quadratic": w["frequency_nodes_quadratic"][bq]}
    for name in ("H1", "L1"):
        out[name + "_linear"] = w[name + "_linear"][bl]
        out[name + "_quadratic"] = w[name + "_quadratic"][bq]
    return out


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_roq_several_bases_by_prior_range(precision):
    """roq.py:192-217, 402-439: each row evaluated with the first linear and
    quadratic bases whose chirp-mass ranges hold it.  The batch equals, row by
    row, the one-basis likelihood of the pair it selects (fp64 1e-10
    relative, fp32 1e-6), the three pairs the ranges allow occur, two rows per pair agree
    with the FP64 ROQ oracle on that pair's weights (fp64 1e-8, fp32 1e-3);
    log_likelihood_ratio records the pair (basis_number_*).  The draws carry
    chirp_mass and mass_ratio, the sampled keys the ranges name (bilby reads
    the ranges' parameters from the sample as given); parameter rows give the
    same result through their component masses."""
    from bilby_amd.roq import ROQGravitationalWaveTransient
    from oracle.roq import OracleROQ
    from tests.helpers import draws_as_dicts, fixed_marginalised
    say("roq two bases", precision)
    s, w = _two_basis_weights()
    kw = dict(time_marginalization=True, phase_marginalization=True, delta_tc=s["delta_tc"],
              precision=precision)
    opri = dict(s["priors"])
    priors = s["priors"].copy()
    lk = ROQGravitationalWaveTransient(s["interferometers"], s["waveform_generator"], priors,
                                       weights=w, **kw)
    assert lk.number_of_bases_linear == 2 and lk.number_of_bases_quadratic == 2
    draws = fixed_marginalised(s["draws"], priors)
    m1, m2 = draws.pop("mass_1"), draws.pop("mass_2")
    mc = (m1 * m2) ** 0.6 / (m1 + m2) ** 0.2
    draws.update(chirp_mass=mc, mass_ratio=np.minimum(m1, m2) / np.maximum(m1, m2))
    odraws = dict(draws, mass_1=m1, mass_2=m2)   # the oracle reads component masses
    got = lk.log_likelihood_ratio_batch(draws)
    from_rows = lk.log_likelihood_ratio_batch(lk.rows_from_parameters(draws))
    assert np.array_equal(from_rows, got)
    bl = np.where(mc <= 20.0, 0, 1)
    bq = np.where(mc <= 30.0, 0, 1)
    pairs = set(zip(bl.tolist(), bq.tolist()))
    say("pairs", sorted(pairs), "chirp mass range", mc.min(), mc.max())
    assert pairs == {(0, 0), (1, 0), (1, 1)}
    rtol = 1e-10 if precision == "fp64" else 1e-6
    worst = 0.0
    for pair in sorted(pairs):
        idx = np.nonzero((bl == pair[0]) & (bq == pair[1]))[0]
        single = ROQGravitationalWaveTransient(s["interferometers"], s["waveform_generator"],
                                               s["priors"].copy(), weights=_pair_weights(w, *pair),
                                               **kw)
        ref = single.log_likelihood_ratio_batch({k: v[idx] for k, v in draws.items()})
        worst = max(worst, float(np.max(np.abs(got[idx] - ref) / np.maximum(1.0, np.abs(ref)))))
        o = OracleROQ(["H1", "L1"], s["interferometers"].start_time, _pair_weights(w, *pair),
                      w["frequency_nodes_linear"][pair[0]], w["frequency_nodes_quadratic"][pair[1]],
                      dict(opri), time_marginalization=True, phase_marginalization=True,
                      delta_tc=s["delta_tc"],
                      waveform_arguments=s["waveform_generator"].waveform_arguments)
        sel = idx[:2]
        oref = np.array([o.log_likelihood_ratio(p) for p in draws_as_dicts(odraws, sel)])
        oerr = float(np.max(np.abs(got[sel] - oref)))
        say("pair", pair, "rows", len(idx), "vs oracle", oerr)
        assert oerr < (1e-8 if precision == "fp64" else 1e-3)
        # the single-dict call records the pair, as _update_basis does
        one = draws_as_dicts(draws, [int(idx[0])])[0]
        assert abs(lk.log_likelihood_ratio(one) - got[idx[0]]) < 1e-9 * max(1, abs(got[idx[0]]))
        assert (lk.basis_number_linear, lk.basis_number_quadratic) == pair
    say("batch vs per-pair worst relative", worst)
    assert worst < rtol


def test_roq_params_grid_and_checks(tmp_path):
    """roq.py:86-145, 653-734, 800-838: bases on the ROQ frequency grid of
    roq_params (flow, fhigh, seglen with roq_scale_factor 2) wider than the
    detectors' bands; the weights sum over the intersection (against the
    oracle's restatement, 1e-10) and the lnL matches the ROQ oracle on them
    (1e-8); the band, duration and CBC chirp-mass checks raise
    BilbyROQParamsRangeError."""
    from bilby_amd.configs import make_taylorf2_marginalised
    from bilby_amd.prior import CBCPriorDict, Uniform
    from bilby_amd.roq import BilbyROQParamsRangeError, ROQGravitationalWaveTransient
    from oracle import roq_weights as orw
    from oracle.roq import OracleROQ
    from tests.helpers import draws_as_dicts, fixed_marginalised
    say("roq params grid")
    s = make_taylorf2_marginalised(n_draws=8, seed=41, duration=4.0, sampling_frequency=1024.0,
                                   detectors=("H1", "L1"), time=True, distance=False, phase=True)
    ifos, wfg = s["interferometers"], s["waveform_generator"]
    ifos[1].maximum_frequency = 400.0
    dt = [("flow", float), ("fhigh", float), ("seglen", float), ("chirpmassmin", float),
          ("chirpmassmax", float)]
    params = np.array((8.0, 256.0, 8.0, 6.0, 200.0), dtype=dt)   # x2: 16-512 Hz, 4 s, Mc >= 3
    n_grid = int(np.sum(np.linspace(0, 512, 2049) >= 16.0))
    rng = np.random.default_rng(6)
    L, Q = 36, 10
    lin = np.linalg.qr(rng.normal(size=(n_grid, L)) + 1j * rng.normal(size=(n_grid, L)))[0]
    quad = np.abs(rng.normal(size=(n_grid, Q)))
    wfg.waveform_arguments.update(frequency_nodes_linear=np.sort(rng.uniform(20, 390, L)),
                                  frequency_nodes_quadratic=np.sort(rng.uniform(20, 390, Q)))
    opri = dict(s["priors"])
    kw = dict(time_marginalization=True, phase_marginalization=True, precision="fp64")
    lk = ROQGravitationalWaveTransient(ifos, wfg, s["priors"].copy(), linear_matrix=lin,
                                       quadratic_matrix=quad, roq_params=params,
                                       roq_scale_factor=2.0, **kw)
    w = lk.weights
    ts = w["time_samples"]
    for ifo in ifos:
        r, i = orw.roq_indices(ifo.frequency_array, ifo.frequency_mask, 8.0, 256.0, 8.0, 2.0)
        assert len(i) == int(ifo.frequency_mask.sum()) and len(r) < n_grid
        ref = orw.linear_weights(ifo.frequency_array, ifo.frequency_mask,
                                 ifo.frequency_domain_strain, ifo.power_spectral_density_array,
                                 ifo.minimum_frequency, ifos.duration, lin.T, ts, r, i)
        err = np.max(np.abs(w[ifo.name + "_linear"] - ref)) / np.max(np.abs(ref))
        q_ref = orw.quadratic_weights(ifo.frequency_mask, ifo.power_spectral_density_array,
                                      ifos.duration, quad.T, r, i)
        say(ifo.name, "grid columns", len(r), "of", n_grid, "linear weights error", err)
        assert err < 1e-10
        assert np.allclose(w[ifo.name + "_quadratic"], q_ref, rtol=1e-13, atol=0)
    o = OracleROQ([i.name for i in ifos], ifos.start_time, w, w["frequency_nodes_linear"],
                  w["frequency_nodes_quadratic"], opri, time_marginalization=True,
                  phase_marginalization=True, delta_tc=ts[1] - ts[0],
                  waveform_arguments=wfg.waveform_arguments)
    draws = fixed_marginalised(s["draws"], s["priors"].copy())
    ref = np.array([o.log_likelihood_ratio(p) for p in draws_as_dicts(draws, range(8))])
    got = lk.log_likelihood_ratio_batch(draws)
    say("lnL max |diff|", float(np.max(np.abs(got - ref))))
    assert np.max(np.abs(got - ref)) < 1e-8 * max(1.0, np.max(np.abs(ref)))
    for bad in (np.array((8.0, 128.0, 8.0, 6.0, 200.0), dtype=dt),   # fhigh below the band
                np.array((12.0, 256.0, 8.0, 6.0, 200.0), dtype=dt),  # flow above it
                np.array((8.0, 256.0, 16.0, 6.0, 200.0), dtype=dt)):  # another duration
        with pytest.raises(BilbyROQParamsRangeError):
            ROQGravitationalWaveTransient(ifos, wfg, s["priors"].copy(), linear_matrix=lin,
                                          quadratic_matrix=quad, roq_params=bad,
                                          roq_scale_factor=2.0, **kw)
    cbc = CBCPriorDict(s["priors"].copy(), chirp_mass=Uniform(2.0, 40.0))
    with pytest.raises(BilbyROQParamsRangeError, match="chirp mass"):
        ROQGravitationalWaveTransient(ifos, wfg, cbc, linear_matrix=lin, quadratic_matrix=quad,
                                      roq_params=params, roq_scale_factor=2.0, **kw)
    # the check can be switched off
    ROQGravitationalWaveTransient(ifos, wfg, cbc.copy(), linear_matrix=lin, quadratic_matrix=quad,
                                  roq_params=params, roq_scale_factor=2.0, roq_params_check=False,
                                  **kw)
