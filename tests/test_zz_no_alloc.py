"""gwl_eval_noalloc (include/gwl.h): the allocation-free evaluation through
the C ABI gives the same results as gwl_eval, on two calls in a row (the
second with fewer rows, so the shared scratch holds the first call's
contents), for the inner-product path with float32 escalation and the
time-marginalised path; a batch whose workspace exceeds the scratch is
rejected with an error and no evaluation."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def _eval(fn, lk, rows_t, precision, aux=False):
    import torch
    from bilby_amd import _lib
    n = rows_t.shape[0]
    out = torch.full((n,), np.nan, dtype=torch.float64, device=rows_t.device)
    aux_t = torch.full((n, 3), np.nan, dtype=torch.float64, device=rows_t.device) if aux else None
    st = torch.cuda.current_stream(rows_t.device).cuda_stream
    rc = fn(lk._event, ctypes.c_void_p(rows_t.data_ptr()), n, _lib.NPARAM, precision,
            ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(aux_t.data_ptr()) if aux else None,
            ctypes.c_void_p(st))
    torch.cuda.synchronize()
    return rc, out.cpu().numpy(), (aux_t.cpu().numpy() if aux else None)


def _rows(lk, draws):
    import torch
    return torch.from_numpy(lk.rows_from_parameters(draws)).to("cuda:0")


def test_noalloc_matches_gwl_eval_plain_with_escalation():
    from bilby_amd._lib import LIB, FP32, FP64
    from bilby_amd.configs import make_sine_gaussian_plain
    from bilby_amd.likelihood import GravitationalWaveTransient
    say("noalloc: sine-Gaussian plain, 3000 then 700 rows")
    setup = make_sine_gaussian_plain(n_draws=3000, seed=4)
    lk = GravitationalWaveTransient(setup["interferometers"], setup["waveform_generator"],
                                    precision="fp32")
    rows = _rows(lk, setup["draws"])
    for n in (3000, 700):
        r = rows[:n].contiguous()
        for prec in (FP32, FP64):
            rc0, ref, aux0 = _eval(LIB.gwl_eval, lk, r, prec, aux=True)
            rc1, got, aux1 = _eval(LIB.gwl_eval_noalloc, lk, r, prec, aux=True)
            assert rc0 == 0 and rc1 == 0
            say(n, prec, "max diff", np.max(np.abs(got - ref)))
            assert np.array_equal(got, ref)
            assert np.array_equal(aux1, aux0)


def test_noalloc_matches_gwl_eval_time_marginalised():
    from bilby_amd._lib import LIB, FP32
    from bilby_amd.configs import make_taylorf2_marginalised
    from bilby_amd.likelihood import GravitationalWaveTransient
    from tests.helpers import fixed_marginalised
    say("noalloc: TaylorF2 T+D+P, 1024 then 300 rows")
    s = make_taylorf2_marginalised(n_draws=1024, seed=8, duration=4.0, sampling_frequency=2048.0,
                                   detectors=("H1", "L1"))
    priors = s["priors"].copy()
    lk = GravitationalWaveTransient(s["interferometers"], s["waveform_generator"],
                                    time_marginalization=True, distance_marginalization=True,
                                    phase_marginalization=True, priors=priors, precision="fp32")
    rows = _rows(lk, fixed_marginalised(s["draws"], priors))
    for n in (1024, 300):
        r = rows[:n].contiguous()
        rc0, ref, _ = _eval(LIB.gwl_eval, lk, r, FP32)
        rc1, got, _ = _eval(LIB.gwl_eval_noalloc, lk, r, FP32)
        assert rc0 == 0 and rc1 == 0
        say(n, "max diff", np.max(np.abs(got - ref)))
        assert np.array_equal(got, ref)


def test_noalloc_rejects_a_batch_beyond_the_scratch():
    import torch
    from bilby_amd._lib import LIB, FP32
    from bilby_amd.configs import make_sine_gaussian_plain
    from bilby_amd.likelihood import GravitationalWaveTransient
    setup = make_sine_gaussian_plain(n_draws=4, seed=4)
    lk = GravitationalWaveTransient(setup["interferometers"], setup["waveform_generator"],
                                    precision="fp32")
    rows = _rows(lk, setup["draws"])
    # 2^20 rows need more than the 64 MiB scratch; the rows buffer itself is
    # never read because the size check comes first
    n = 1 << 20
    big = torch.zeros((n, 16), dtype=torch.float64, device="cuda:0")
    big[:4] = rows
    out = torch.zeros(n, dtype=torch.float64, device="cuda:0")
    rc = LIB.gwl_eval_noalloc(lk._event, ctypes.c_void_p(big.data_ptr()), n, 16, FP32,
                              ctypes.c_void_p(out.data_ptr()), None, None)
    torch.cuda.synchronize()
    assert rc != 0
    assert b"scratch" in LIB.gwl_last_error()
    assert float(out.abs().max()) == 0.0
