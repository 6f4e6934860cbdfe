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
