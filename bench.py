#!/usr/bin/env python
"""Benchmark: lnL evaluations per second on one MI355X.

Workload (one "step" = one batched evaluation of the hot path): TaylorF2
non-spinning, H1+L1+V1, 8 s at 4096 Hz (16385 bins), time + distance + phase
marginalisation, a batch of 4096 prior draws, float32 bulk arithmetic with
float64 phase arguments and accumulation.  Accuracy is audited against the
float64 GPU path on every row and against the FP64 oracle on a sample.

Prints one JSON line (see the keys below).  --gpus N > 1 is refused: this
build measures one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "lnL evals/s at |ΔlnL|<1e-3 vs FP64 (1 GPU); p50/p99 latency under load"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def bytes_per_eval(lk):
    """Algorithmic HBM bytes of one evaluation: per detector and masked bin a
    complex64 d/S and a float32 1/S (12 B), plus the parameter row (16
    float64) and the output (float64)."""
    n_mask = sum(int(ifo.frequency_mask.sum()) for ifo in lk.interferometers)
    return n_mask * 12 + 16 * 8 + 8


def cpu_baseline(setup, priors_fixed, seconds):
    """The FP64 oracle driven like bilby drives a likelihood: one parameter
    dict per call, one core."""
    import numpy as np
    from tests.helpers import draws_as_dicts, oracle_from_setup
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    olk, _ = oracle_from_setup(setup)
    from tests.helpers import fixed_marginalised
    draws = fixed_marginalised(setup["draws"], priors_fixed)
    dicts = draws_as_dicts(draws, range(min(2048, len(next(iter(draws.values()))))))
    olk.log_likelihood_ratio(dicts[0])
    n = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        olk.log_likelihood_ratio(dicts[n % len(dicts)])
        n += 1
    dt = time.perf_counter() - t0
    return olk, dict(value=n / dt, unit="lnL evals/s", cores=1, kind="port",
                     sample="%d single-dict oracle calls (%.1f s) of the same 8 s / 4096 Hz "
                            "H1+L1+V1 TaylorF2 T+D+P-marginalised workload; TaylorF2 in "
                            "numpy, not lalsimulation" % (n, dt))


_POOL_LK = None


def _pool_init(blob):
    global _POOL_LK
    import pickle
    _POOL_LK = pickle.loads(blob)


def _pool_run(job):
    dicts, seconds = job
    _POOL_LK.log_likelihood_ratio(dicts[0])
    n = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        _POOL_LK.log_likelihood_ratio(dicts[n % len(dicts)])
        n += 1
    return n, time.perf_counter() - t0


def profile_figure(tag, key):
    """A figure recorded from a counter run in a summary under profiles/ (a
    line '<tag> <key>: <number>'); None if no summary records it."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        with open(path) as fh:
            for line in fh:
                line = line.strip().strip("`")
                if line.startswith("%s %s:" % (tag, key)):
                    return float(line.split(":", 1)[1])
    return None


# VALU issue peak of the MI355X: 256 CUs x 4 SIMDs, a wave64 VALU instruction
# every 4 cycles on a 16-lane SIMD, at the 2.4 GHz peak engine clock
VALU_PEAK_GIPS = 256 * 4 / 4 * 2.4


def roofline_block(kname, wave, k_ms, achieved_gbps, bpe, traffic, key):
    """The roofline of the headline kernel.  Its event data (0.6 MB) stays in
    L2/MALL and every row re-reads it, so the algorithmic bytes over the
    kernel time exceed what HBM alone delivers and HBM traffic is a small
    fraction of them: the bound is VALU issue.  achieved = the kernel's VALU
    wave-instructions per launch (rocprofv3 SQ_INSTS_VALU, a property of the
    kernel and the batch, in profiles/) over its launch time measured here
    with HIP events; the HBM figures are kept under "hbm"."""
    desc = (" (wave per row: waveform + projection + register radix-8 transform + window + "
            "fused epilogue)" if wave else " (tiled: waveform + projection + LDS FFT)")
    hbm = {"achieved": achieved_gbps, "peak": 8000.0, "unit": "GB/s",
           "frac": achieved_gbps / 8000.0, "bytes_per_eval": bpe,
           "note": "algorithmic bytes / kernel time; L2/MALL-resident data re-read per row"}
    insts = profile_figure("bench-valu-insts", key)
    out = {"kernel": kname + desc, "kernel_ms": k_ms, "traffic": traffic,
           "traffic_source": "rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE x1.0 per launch, "
                             "profiles/ (bench-traffic %s)" % key, "hbm": hbm}
    if insts is None:
        out.update(bound="hbm", achieved=achieved_gbps, peak=8000.0, unit="GB/s",
                   frac=achieved_gbps / 8000.0)
        return out
    gips = insts / (k_ms * 1e-3) / 1e9
    out.update(bound="valu", achieved=gips, peak=VALU_PEAK_GIPS,
               unit="G VALU wave-instructions/s", frac=gips / VALU_PEAK_GIPS,
               valu_insts_per_launch=insts,
               valu_source="rocprofv3 SQ_INSTS_VALU per dispatch, profiles/ "
                           "(bench-valu-insts %s); launch time from HIP events in this run" % key)
    return out


def measured_traffic(key):
    """HBM bytes per launch of the kernel ``key`` from the FETCH_SIZE /
    WRITE_SIZE counter runs ('bench-traffic' lines)."""
    v = profile_figure("bench-traffic", key)
    return None if v is None else int(v)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline_pool(olk, dicts, seconds, workers):
    """The same oracle over a process pool, as bilby's npool does
    (bilby/core/sampler/base_sampler.py:772-800): every worker evaluates one
    dict per call on one thread for ``seconds``; the rate is the sum."""
    import multiprocessing
    import pickle
    blob = pickle.dumps(olk)
    old = os.environ.get("OMP_NUM_THREADS")
    os.environ["OMP_NUM_THREADS"] = "1"
    try:
        ctx = multiprocessing.get_context("spawn")
        with ctx.Pool(workers, initializer=_pool_init, initargs=(blob,)) as pool:
            res = pool.map(_pool_run, [(dicts[k::workers], seconds) for k in range(workers)],
                           chunksize=1)
    finally:
        if old is None:
            os.environ.pop("OMP_NUM_THREADS", None)
        else:
            os.environ["OMP_NUM_THREADS"] = old
    n = sum(r[0] for r in res)
    rate = sum(r[0] / r[1] for r in res)
    return dict(value=rate, unit="lnL evals/s", cores=workers, kind="port",
                cpu=cpu_model(), host_cpus=os.cpu_count(),
                sample="%d single-dict oracle calls over %d worker processes, %.0f s each, same "
                       "workload as cpu_baseline" % (n, workers, seconds))


def timed(fn, steps, warmup=2):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def bench_sine_gaussian(steps, n_oracle=256):
    """Plain likelihood, sine-Gaussian, H1+L1, 4 s @ 2048 Hz, 10^4 draws:
    float32 with per-row float64 escalation (the default), float32 without
    it (not a metric figure: it misses 1e-3 on loud rows), float64."""
    import numpy as np
    import torch
    from bilby_amd.configs import make_sine_gaussian_plain
    from bilby_amd.likelihood import GravitationalWaveTransient
    from tests.helpers import draws_as_dicts, oracle_from_setup
    setup = make_sine_gaussian_plain(n_draws=10000, seed=1234)
    lk = GravitationalWaveTransient(setup["interferometers"], setup["waveform_generator"])
    rows = torch.from_numpy(lk.rows_from_parameters(setup["draws"])).cuda()
    out = {}
    res = {}
    for name, prec, esc in (("fp32_escalated", "fp32", True), ("fp32_raw", "fp32", False),
                            ("fp64", "fp64", True)):
        if esc:
            lk.set_precision_escalation()
        else:
            lk.set_precision_escalation(0.0)
        dt = timed(lambda: lk.log_likelihood_ratio_batch(rows, precision=prec), steps)
        out[name + "_evals_per_s"] = rows.shape[0] / dt
        res[name] = lk.log_likelihood_ratio_batch(rows, precision=prec).cpu().numpy()
        if name == "fp32_escalated":
            out["escalated_fraction"] = lk.escalated_rows() / rows.shape[0]
    lk.set_precision_escalation()
    olk, _ = oracle_from_setup(setup)
    ref = np.array([olk.log_likelihood_ratio(p)
                    for p in draws_as_dicts(setup["draws"], range(n_oracle))])
    for name in res:
        d = np.abs(res[name][:n_oracle] - ref)
        out[name + "_max_abs_dlnL_vs_oracle_%drows" % n_oracle] = float(d.max())
        out[name + "_rows_within_1e-3_vs_fp64_gpu"] = float(
            np.mean(np.abs(res[name] - res["fp64"]) < 1e-3))
    out["fp32_escalated_max_abs_dlnL_vs_fp64_gpu"] = float(
        np.max(np.abs(res["fp32_escalated"] - res["fp64"])))
    out["max_abs_lnL"] = float(np.max(np.abs(res["fp64"])))
    return out


def bench_uneven(steps):
    """64 events, durations log-uniform over {4..128} s, 2048/4096 Hz, 2-3
    detectors, aligned-spin TaylorF2, T+D+P marginalised, 256-4096 rows each,
    one persistent launch replayed from a hipGraph; tile per bin-count bucket
    chosen by bilby_amd.autotune.autotune_multi."""
    import numpy as np
    import torch
    from bilby_amd.configs import make_uneven_events
    from bilby_amd.likelihood import GravitationalWaveTransient
    from bilby_amd.multi import MultiEventLikelihood
    from tests.helpers import fixed_marginalised
    events = make_uneven_events(n_events=64, seed=97531)
    lks, draws = [], []
    for ev in events:
        pri = ev["priors"].copy()
        lks.append(GravitationalWaveTransient(ev["interferometers"], ev["waveform_generator"],
                                              time_marginalization=True,
                                              distance_marginalization=True,
                                              phase_marginalization=True, priors=pri))
        draws.append(fixed_marginalised(ev["draws"], pri))
    multi = MultiEventLikelihood(lks, [ev["n_rows"] for ev in events])
    for e, d in enumerate(draws):
        multi.set_parameters(e, d)
    total_rows = sum(ev["n_rows"] for ev in events)
    from bilby_amd.autotune import autotune_multi
    steps_of, sweep = autotune_multi(multi, use_cache=False)
    multi.capture()
    dt = timed(lambda: multi.run(), steps)
    outs = [o.cpu().numpy() for o in multi.run()]
    # agreement with the per-event float64 path, and with the FP64 oracle on
    # 4 rows of every event (one distance table: every event has the same
    # distance prior and phase flag)
    from tests.helpers import draws_as_dicts, oracle_from_setup
    dmax = 0.0
    for lk, d, o in zip(lks[:16], draws[:16], outs[:16]):
        ref = lk.log_likelihood_ratio_batch(d, precision="fp64")
        dmax = max(dmax, float(np.max(np.abs(o - ref))))
    omax, table = 0.0, None
    for ev, d, o in zip(events, draws, outs):
        olk, _ = oracle_from_setup(ev, distance_table=table)
        table = olk._dist_margd_loglikelihood_array
        idx = [0, 1, ev["n_rows"] // 2, ev["n_rows"] - 1]
        ref = np.array([olk.log_likelihood_ratio(p) for p in draws_as_dicts(d, idx)])
        omax = max(omax, float(np.max(np.abs(o[idx] - ref))))
    n_bins = [len(lk.interferometers.frequency_array) for lk in lks]
    # roofline from the seeded event list: the algorithmic bytes of every
    # row (float32 d/S and 1/S on each detector's masked bins, and the row),
    # and the VALU issue of the launch from its counted instructions
    abytes = sum(ev["n_rows"] * (sum(int(np.sum(np.asarray(i.frequency_mask)))
                                     for i in lk.interferometers) * 12 + 72)
                 for ev, lk in zip(events, lks))
    valu = profile_figure("bench-valu-insts", "k_time_w_multi_uneven_64_events")
    roof = {"kernel": "k_time_w_multi (persistent, all 64 events in one launch)",
            "algorithmic_bytes_per_launch": abytes,
            "hbm": {"achieved": abytes / dt / 1e9, "peak": 8000.0, "unit": "GB/s",
                    "frac": abytes / dt / 1e9 / 8000.0,
                    "note": "each event's data is L2/MALL-resident and re-read by its rows"}}
    if valu:
        roof.update(bound="valu", valu_insts_per_launch=valu,
                    achieved=valu / dt / 1e9, peak=VALU_PEAK_GIPS,
                    unit="G VALU wave-instructions/s", frac=valu / dt / 1e9 / VALU_PEAK_GIPS,
                    valu_source="rocprofv3 SQ_INSTS_VALU per dispatch, profiles/ "
                                "(bench-valu-insts k_time_w_multi_uneven_64_events); "
                                "launch time of the graph replay in this run")
    return dict(events=len(events), rows=int(total_rows), ms_per_launch=dt * 1e3,
                evals_per_s=total_rows / dt, bins_min=int(min(n_bins)), bins_max=int(max(n_bins)),
                tile_steps_by_bins={str(k): int(v) for k, v in steps_of.items()},
                tile_sweep_ms=sweep, fp32_vs_fp64_gpu_max_16_events=dmax,
                fp32_vs_fp64_oracle_max_4_rows_x_64_events=omax, graph=True, roofline=roof)


def bench_relbin(steps):
    """Relative binning on the headline event (TaylorF2 H1+L1+V1, 8 s @
    4096 Hz), distance + phase marginalised, fiducial at the injection,
    epsilon 0.5 (bilby's default) and 0.05, 65536 draws around the
    fiducial point; the full (unbinned) likelihood with the same
    marginalisations timed beside it."""
    import ctypes
    import numpy as np
    import torch
    from bilby_amd._lib import LIB, check
    from bilby_amd.configs import make_taylorf2_marginalised
    from bilby_amd.likelihood import GravitationalWaveTransient
    from bilby_amd.relative import RelativeBinningGravitationalWaveTransient
    from oracle.likelihood import OracleDetector, OracleEvent
    from oracle.relative import OracleRelativeBinning
    from tests.helpers import draws_as_dicts
    s = make_taylorf2_marginalised(n_draws=8, seed=4321, time=False)
    inj = dict(s["injection"])
    n = 65536
    rng = np.random.default_rng(77)
    draws = {k: np.full(n, float(v)) for k, v in inj.items()}
    for k, sd in (("mass_1", 0.003), ("mass_2", 0.003)):
        draws[k] = draws[k] * (1 + rng.normal(0, sd, n))
    for k, sd in (("ra", 0.1), ("dec", 0.1), ("psi", 0.3), ("theta_jn", 0.1),
                  ("geocent_time", 5e-4)):
        draws[k] = draws[k] + rng.normal(0, sd, n)
    out = {"rows": n}
    ifos = s["interferometers"]
    for eps in (0.5, 0.05):
        lk = RelativeBinningGravitationalWaveTransient(
            ifos, s["waveform_generator"], fiducial_parameters=inj, distance_marginalization=True,
            phase_marginalization=True, priors=s["priors"].copy(), epsilon=eps, precision="fp32")
        d = dict(draws)
        d["luminosity_distance"] = np.full(n, lk._ref_dist)
        d["phase"] = np.zeros(n)
        rows = torch.from_numpy(lk.rows_from_parameters(d)).cuda()
        key = "eps%g" % eps
        out[key + "_bins"] = lk.number_of_bins
        for prec in ("fp32", "fp64"):
            dt = timed(lambda: lk.log_likelihood_ratio_batch(rows, precision=prec), steps)
            out["%s_%s_evals_per_s" % (key, prec)] = n / dt
        check(LIB.gwl_event_timing(lk._event, 1))
        for _ in range(steps):
            lk.log_likelihood_ratio_batch(rows)
        torch.cuda.synchronize()
        tot, cnt = ctypes.c_double(), ctypes.c_int64()
        check(LIB.gwl_event_timing_read(lk._event, ctypes.byref(tot), ctypes.byref(cnt)))
        check(LIB.gwl_event_timing(lk._event, 0))
        out[key + "_k_relbin_fp32_ms"] = tot.value / max(1, cnt.value)
        r32 = lk.log_likelihood_ratio_batch(rows, precision="fp32").cpu().numpy()
        r64 = lk.log_likelihood_ratio_batch(rows, precision="fp64").cpu().numpy()
        out[key + "_fp32_vs_fp64_max"] = float(np.max(np.abs(r32 - r64)))
        dets = [OracleDetector(i.name, i.frequency_array, i.frequency_domain_strain,
                               i.power_spectral_density_array, i.frequency_mask) for i in ifos]
        ev = OracleEvent(dets, ifos.duration, ifos.sampling_frequency, ifos.start_time)
        olk = OracleRelativeBinning(ev, "taylorf2", inj,
                                    waveform_arguments=dict(s["waveform_generator"].waveform_arguments),
                                    distance_marginalization=True, phase_marginalization=True,
                                    priors=dict(s["priors"]), epsilon=eps,
                                    distance_table=lk._dist_margd_loglikelihood_array)
        ref = np.array([olk.log_likelihood_ratio(p) for p in draws_as_dicts(d, range(16))])
        out[key + "_fp64_vs_oracle_16rows"] = float(np.max(np.abs(r64[:16] - ref)))
        out[key + "_fp32_vs_oracle_16rows"] = float(np.max(np.abs(r32[:16] - ref)))
    full = GravitationalWaveTransient(ifos, s["waveform_generator"], distance_marginalization=True,
                                      phase_marginalization=True, priors=s["priors"].copy(),
                                      precision="fp32")
    rows = torch.from_numpy(full.rows_from_parameters(d)).cuda()
    check(LIB.gwl_event_reserve(full._event, n))
    dt = timed(lambda: full.log_likelihood_ratio_batch(rows), steps)
    out["full_likelihood_fp32_evals_per_s"] = n / dt
    fr = full.log_likelihood_ratio_batch(rows).cpu().numpy()
    out["eps0.05_vs_full_fp32_median_abs"] = float(np.median(np.abs(r32 - fr)))
    return out


def bench_calibration(steps, batch=2048, n_curves=1000):
    """Calibration marginalisation on the headline event (TaylorF2 H1+L1+V1,
    8 s @ 4096 Hz, 3 x 16225 masked bins), distance + phase marginalised,
    1000 cubic-spline response curves per detector (5 nodes, 10 %% / 0.1 rad),
    a batch of prior draws: prep, two MFMA contractions, epilogue."""
    import ctypes
    import tempfile
    import numpy as np
    import torch
    from bilby_amd import calibration as cal
    from bilby_amd._lib import LIB, check
    from bilby_amd.configs import make_taylorf2_marginalised
    from bilby_amd.likelihood import GravitationalWaveTransient
    from oracle.calibration import OracleCalibrated
    from oracle.likelihood import OracleDetector, OracleEvent
    from tests.helpers import draws_as_dicts, fixed_marginalised
    s = make_taylorf2_marginalised(n_draws=batch, seed=4321, time=False)
    priors = s["priors"]
    ifos = s["interferometers"]
    for ifo in ifos:
        ifo.calibration_model = cal.CubicSpline("recalib_%s_" % ifo.name, 20.0, 2048.0, 5)
        priors.update(cal.constant_uncertainty_spline_priors(0.1, 0.1, 20.0, 2048.0, 5, ifo.name))
    tmp = tempfile.mkdtemp(prefix="gwl-cal-")
    files = {ifo.name: os.path.join(tmp, ifo.name + ".npz") for ifo in ifos}
    opri = dict(priors)
    lk_priors = priors.copy()
    lk = GravitationalWaveTransient(ifos, s["waveform_generator"], calibration_marginalization=True,
                                    distance_marginalization=True, phase_marginalization=True,
                                    priors=lk_priors, number_of_response_curves=n_curves,
                                    calibration_lookup_table=files, precision="fp32")
    draws = fixed_marginalised(s["draws"], lk_priors)
    rows = torch.from_numpy(lk.rows_from_parameters(draws)).cuda()
    K = sum(int(i.frequency_mask.sum()) for i in ifos)
    useful = 10.0 * K * n_curves   # 8 K C (complex <d|h>) + 2 K C (<h|h>) per row
    out = {"rows": batch, "curves": n_curves, "masked_bins": K,
           "useful_gflop_per_eval": useful / 1e9}
    check(LIB.gwl_calib_reserve(lk._calib, batch))
    res = {}
    for mode in ("fp32", "bf16x3", "fp64"):
        st = steps if mode != "fp64" else 1
        dt = timed(lambda: lk.log_likelihood_ratio_batch(rows, precision=mode), st, warmup=1)
        out[mode + "_evals_per_s"] = batch / dt
        out[mode + "_ms_per_batch"] = dt * 1e3
        out[mode + "_tflops_useful"] = batch * useful / dt / 1e12
        res[mode] = lk.log_likelihood_ratio_batch(rows, precision=mode).cpu().numpy()
    for mode in ("fp32", "bf16x3"):
        out[mode + "_vs_fp64_max"] = float(np.max(np.abs(res[mode] - res["fp64"])))
    dets = [OracleDetector(i.name, i.frequency_array, i.frequency_domain_strain,
                           i.power_spectral_density_array, i.frequency_mask) for i in ifos]
    ev = OracleEvent(dets, ifos.duration, ifos.sampling_frequency, ifos.start_time)
    olk = OracleCalibrated(ev, "taylorf2", lk.calibration_draws,
                           waveform_arguments=dict(s["waveform_generator"].waveform_arguments),
                           distance_marginalization=True, phase_marginalization=True, priors=opri,
                           distance_table=lk._dist_margd_loglikelihood_array)
    t0 = time.perf_counter()
    ref = np.array([olk.log_likelihood_ratio(p) for p in draws_as_dicts(draws, range(8))])
    out["oracle_evals_per_s_1core"] = 8 / (time.perf_counter() - t0)
    for mode in res:
        out[mode + "_vs_oracle_8rows"] = float(np.max(np.abs(res[mode][:8] - ref)))
    del lk
    # calibration together with time (+ distance + phase) marginalisation: one
    # window transform per (row, curve), the same event and curves
    tbatch = 256
    s2 = make_taylorf2_marginalised(n_draws=tbatch, seed=4321)
    priors2 = s2["priors"]
    ifos2 = s2["interferometers"]
    for ifo in ifos2:
        ifo.calibration_model = cal.CubicSpline("recalib_%s_" % ifo.name, 20.0, 2048.0, 5)
        priors2.update(cal.constant_uncertainty_spline_priors(0.1, 0.1, 20.0, 2048.0, 5, ifo.name))
    opri2 = dict(priors2)
    lk2_priors = priors2.copy()
    lk2 = GravitationalWaveTransient(ifos2, s2["waveform_generator"],
                                     calibration_marginalization=True, time_marginalization=True,
                                     distance_marginalization=True, phase_marginalization=True,
                                     priors=lk2_priors, number_of_response_curves=n_curves,
                                     calibration_lookup_table=files, precision="fp32")
    draws2 = fixed_marginalised(s2["draws"], lk2_priors)
    draws2["time_jitter"] = np.random.default_rng(5).uniform(-1, 1, tbatch) / 4096.0
    rows2 = torch.from_numpy(lk2.rows_from_parameters(draws2)).cuda()
    check(LIB.gwl_calib_reserve(lk2._calib, tbatch))
    dt = timed(lambda: lk2.log_likelihood_ratio_batch(rows2), steps, warmup=1)
    out["time_tdp_rows"] = tbatch
    out["time_tdp_fp32_evals_per_s"] = tbatch / dt
    out["time_tdp_ms_per_batch"] = dt * 1e3
    out["time_tdp_curve_transforms_per_s"] = tbatch * n_curves / dt
    dt64 = timed(lambda: lk2.log_likelihood_ratio_batch(rows2, precision="fp64"), 1, warmup=1)
    out["time_tdp_fp64_evals_per_s"] = tbatch / dt64
    out["time_tdp_fp64_ms_per_batch"] = dt64 * 1e3
    r2 = lk2.log_likelihood_ratio_batch(rows2).cpu().numpy()
    r64 = lk2.log_likelihood_ratio_batch(rows2, precision="fp64").cpu().numpy()
    ev2 = OracleEvent([OracleDetector(i.name, i.frequency_array, i.frequency_domain_strain,
                                      i.power_spectral_density_array, i.frequency_mask)
                       for i in ifos2], ifos2.duration, ifos2.sampling_frequency,
                      ifos2.start_time)
    olk2 = OracleCalibrated(ev2, "taylorf2", lk2.calibration_draws,
                            waveform_arguments=dict(s2["waveform_generator"].waveform_arguments),
                            time_marginalization=True, distance_marginalization=True,
                            phase_marginalization=True, priors=opri2,
                            distance_table=lk2._dist_margd_loglikelihood_array)
    t0 = time.perf_counter()
    ref2 = np.array([olk2.log_likelihood_ratio(p) for p in draws_as_dicts(draws2, range(2))])
    out["time_tdp_oracle_evals_per_s_1core"] = 2 / (time.perf_counter() - t0)
    out["time_tdp_fp32_vs_oracle_2rows"] = float(np.max(np.abs(r2[:2] - ref2)))
    out["time_tdp_fp64_vs_oracle_2rows"] = float(np.max(np.abs(r64[:2] - ref2)))
    return out


def bench_multiband(steps, batch=4096, duration=128.0):
    """Multibanding on a long segment: a light binary (1.5 + 1.3 Msun,
    TaylorF2, H1+L1+V1), 128 s @ 2048 Hz (131,073 bins per detector before
    the mask), distance + phase marginalised, a batch of draws around the
    injection; the full (unbanded) likelihood timed beside it."""
    import numpy as np
    import torch
    from bilby_amd import InterferometerList, WaveformGenerator, taylorf2
    from bilby_amd._lib import LIB, check
    from bilby_amd.likelihood import GravitationalWaveTransient, inject_signal
    from bilby_amd.multiband import MBGravitationalWaveTransient
    from bilby_amd.prior import PriorDict, Uniform, UniformComovingVolume
    fs = 2048.0
    tc = 1187008882.0
    start = tc - duration + 2.0
    ifos = InterferometerList(["H1", "L1", "V1"])
    ifos.set_strain_data_from_power_spectral_densities(fs, duration, start,
                                                       random_state=np.random.default_rng(11))
    wfg = WaveformGenerator(duration=duration, sampling_frequency=fs,
                            frequency_domain_source_model=taylorf2,
                            waveform_arguments=dict(reference_frequency=20.0,
                                                    minimum_frequency=20.0))
    inj = dict(mass_1=1.5, mass_2=1.3, chi_1=0.0, chi_2=0.0, luminosity_distance=100.0,
               theta_jn=0.4, phase=1.3, ra=1.3, dec=-1.2, psi=0.659, geocent_time=tc)
    inject_signal(ifos, wfg, inj)
    priors = PriorDict(geocent_time=Uniform(tc - 0.1, tc + 0.1),
                       luminosity_distance=UniformComovingVolume(10.0, 300.0),
                       phase=Uniform(0.0, 2 * np.pi, boundary="periodic"))
    mc = (inj["mass_1"] * inj["mass_2"]) ** 0.6 / (inj["mass_1"] + inj["mass_2"]) ** 0.2
    t0 = time.perf_counter()
    lk = MBGravitationalWaveTransient(ifos, wfg, reference_chirp_mass=mc, priors=priors.copy(),
                                      distance_marginalization=True, phase_marginalization=True,
                                      precision="fp32")
    setup_s = time.perf_counter() - t0
    full = GravitationalWaveTransient(ifos, wfg, priors=priors.copy(),
                                      distance_marginalization=True, phase_marginalization=True,
                                      precision="fp32")
    rng = np.random.default_rng(12)
    d = {k: np.full(batch, float(v)) for k, v in inj.items()}
    d["mass_1"] = d["mass_1"] * (1 + rng.normal(0, 2e-6, batch))
    d["mass_2"] = d["mass_2"] * (1 + rng.normal(0, 2e-6, batch))
    for k, sd in (("ra", 0.05), ("dec", 0.05), ("psi", 0.2), ("geocent_time", 2e-4)):
        d[k] = d[k] + rng.normal(0, sd, batch)
    d["luminosity_distance"] = np.full(batch, lk._ref_dist)
    d["phase"] = np.zeros(batch)
    rows = torch.from_numpy(lk.rows_from_parameters(d)).cuda()
    check(LIB.gwl_event_reserve(full._event, batch))
    out = {"rows": batch, "duration_s": duration, "bins_per_detector": len(ifos.frequency_array),
           "banded_points": int(len(lk.setup.banded_frequency_points)),
           "bands": int(lk.setup.number_of_bands), "setup_s": setup_s}
    for prec in ("fp32", "fp64"):
        dt = timed(lambda: lk.log_likelihood_ratio_batch(rows, precision=prec), steps)
        out["mb_%s_evals_per_s" % prec] = batch / dt
    dt = timed(lambda: full.log_likelihood_ratio_batch(rows), steps)
    out["full_fp32_evals_per_s"] = batch / dt
    mb64 = lk.log_likelihood_ratio_batch(rows, precision="fp64").cpu().numpy()
    mb32 = lk.log_likelihood_ratio_batch(rows, precision="fp32").cpu().numpy()
    f64 = full.log_likelihood_ratio_batch(rows, precision="fp64").cpu().numpy()
    out["mb_fp32_vs_mb_fp64_max"] = float(np.max(np.abs(mb32 - mb64)))
    out["mb_vs_full_fp64_max"] = float(np.max(np.abs(mb64 - f64)))
    out["lnL_max"] = float(np.max(f64))
    # the IFFT-FFT <h|h> (linear_interpolation=False) on the same bands
    t0 = time.perf_counter()
    lki = MBGravitationalWaveTransient(ifos, wfg, reference_chirp_mass=mc, priors=priors.copy(),
                                       distance_marginalization=True, phase_marginalization=True,
                                       linear_interpolation=False, precision="fp32")
    out["ifft_setup_s"] = time.perf_counter() - t0
    out["ifft_half_lengths"] = [int(m) // 2 for m in lki.setup.Mbs[1:]]
    for prec in ("fp32", "fp64"):
        dt = timed(lambda: lki.log_likelihood_ratio_batch(rows, precision=prec), steps)
        out["mb_ifft_%s_evals_per_s" % prec] = batch / dt
    i64 = lki.log_likelihood_ratio_batch(rows, precision="fp64").cpu().numpy()
    out["mb_ifft_vs_full_fp64_max"] = float(np.max(np.abs(i64 - f64)))
    return out


def bench_binned_time(steps):
    """Time + distance + phase marginalised binned likelihoods: relative
    binning on the headline event (TaylorF2 H1+L1+V1, 8 s @ 4096 Hz,
    epsilon 0.5 and 0.05, 65536 rows with time jitter; one GEMM of the
    rows' edge ratios against the per-edge window basis per tile of 8192)
    and multibanding on the 128 s @ 2048 Hz light binary (4096 rows; GEMM
    against the banded-point exponentials); the full T+D+P likelihood on
    the same events timed beside them; a few rows against the FP64 oracle."""
    import numpy as np
    import torch
    from bilby_amd import InterferometerList, WaveformGenerator, taylorf2
    from bilby_amd.configs import make_taylorf2_marginalised
    from bilby_amd.likelihood import GravitationalWaveTransient, inject_signal
    from bilby_amd.multiband import MBGravitationalWaveTransient
    from bilby_amd.prior import PriorDict, Uniform, UniformComovingVolume
    from bilby_amd.relative import RelativeBinningGravitationalWaveTransient
    from oracle.likelihood import OracleDetector, OracleEvent
    from oracle.multiband import OracleMultiband
    from oracle.relative import OracleRelativeBinning
    from tests.helpers import draws_as_dicts

    def oracle_event(ifos):
        dets = [OracleDetector(i.name, i.frequency_array, i.frequency_domain_strain,
                               i.power_spectral_density_array, i.frequency_mask) for i in ifos]
        return OracleEvent(dets, ifos.duration, ifos.sampling_frequency, ifos.start_time)

    def draws_near(inj, n, rng, ifos, lk, mass_sd):
        d = {k: np.full(n, float(v)) for k, v in inj.items()}
        d["mass_1"] = d["mass_1"] * (1 + rng.normal(0, mass_sd, n))
        d["mass_2"] = d["mass_2"] * (1 + rng.normal(0, mass_sd, n))
        for k, sd in (("ra", 0.05), ("dec", 0.05), ("psi", 0.2), ("theta_jn", 0.05)):
            d[k] = d[k] + rng.normal(0, sd, n)
        d["geocent_time"] = np.full(n, float(ifos.start_time))
        d["time_jitter"] = rng.uniform(-1, 1, n) / ifos.sampling_frequency
        d["luminosity_distance"] = np.full(n, lk._ref_dist)
        d["phase"] = np.zeros(n)
        return d

    out = {}
    s = make_taylorf2_marginalised(n_draws=8, seed=4321)
    ifos, wfg = s["interferometers"], s["waveform_generator"]
    inj = dict(s["injection"])
    n = 65536
    full = GravitationalWaveTransient(ifos, wfg, time_marginalization=True,
                                      distance_marginalization=True, phase_marginalization=True,
                                      priors=s["priors"].copy(), precision="fp32")
    for eps in (0.5, 0.05):
        lk = RelativeBinningGravitationalWaveTransient(
            ifos, wfg, fiducial_parameters=inj, time_marginalization=True,
            distance_marginalization=True, phase_marginalization=True,
            priors=s["priors"].copy(), epsilon=eps, precision="fp32")
        d = draws_near(inj, n, np.random.default_rng(78), ifos, lk, 0.003)
        rows = torch.from_numpy(lk.rows_from_parameters(d)).cuda()
        key = "relbin_eps%g" % eps
        out[key + "_bins"] = lk.number_of_bins
        for prec in ("fp32", "fp64"):
            dt = timed(lambda: lk.log_likelihood_ratio_batch(rows, precision=prec), steps)
            out["%s_%s_evals_per_s" % (key, prec)] = n / dt
        r32 = lk.log_likelihood_ratio_batch(rows, precision="fp32").cpu().numpy()
        r64 = lk.log_likelihood_ratio_batch(rows, precision="fp64").cpu().numpy()
        out[key + "_fp32_vs_fp64_max"] = float(np.max(np.abs(r32 - r64)))
        olk = OracleRelativeBinning(oracle_event(ifos), "taylorf2", inj,
                                    waveform_arguments=dict(wfg.waveform_arguments),
                                    time_marginalization=True, distance_marginalization=True,
                                    phase_marginalization=True, priors=dict(s["priors"]),
                                    epsilon=eps, distance_table=lk._dist_margd_loglikelihood_array)
        ref = np.array([olk.log_likelihood_ratio(p) for p in draws_as_dicts(d, range(8))])
        out[key + "_fp64_vs_oracle_8rows"] = float(np.max(np.abs(r64[:8] - ref)))
        out[key + "_fp32_vs_oracle_8rows"] = float(np.max(np.abs(r32[:8] - ref)))
        del lk
    rows = torch.from_numpy(full.rows_from_parameters(d)).cuda()
    dt = timed(lambda: full.log_likelihood_ratio_batch(rows), steps)
    out["relbin_event_full_tdp_fp32_evals_per_s"] = n / dt
    fr = full.log_likelihood_ratio_batch(rows).cpu().numpy()
    out["relbin_eps0.05_vs_full_tdp_median_abs"] = float(np.median(np.abs(r32 - fr)))

    # multibanding, 128 s @ 2048 Hz
    duration, fs, batch = 128.0, 2048.0, 4096
    tc = 1187008882.0
    start = tc - duration + 2.0
    ifos = InterferometerList(["H1", "L1", "V1"])
    ifos.set_strain_data_from_power_spectral_densities(fs, duration, start,
                                                       random_state=np.random.default_rng(11))
    wfg = WaveformGenerator(duration=duration, sampling_frequency=fs,
                            frequency_domain_source_model=taylorf2,
                            waveform_arguments=dict(reference_frequency=20.0,
                                                    minimum_frequency=20.0))
    inj = dict(mass_1=1.5, mass_2=1.3, chi_1=0.0, chi_2=0.0, luminosity_distance=100.0,
               theta_jn=0.4, phase=1.3, ra=1.3, dec=-1.2, psi=0.659, geocent_time=tc)
    inject_signal(ifos, wfg, inj)
    priors = PriorDict(geocent_time=Uniform(tc - 0.1, tc + 0.1),
                       luminosity_distance=UniformComovingVolume(10.0, 300.0),
                       phase=Uniform(0.0, 2 * np.pi, boundary="periodic"))
    mc = (inj["mass_1"] * inj["mass_2"]) ** 0.6 / (inj["mass_1"] + inj["mass_2"]) ** 0.2
    t0 = time.perf_counter()
    lk = MBGravitationalWaveTransient(ifos, wfg, reference_chirp_mass=mc, priors=priors.copy(),
                                      time_marginalization=True, distance_marginalization=True,
                                      phase_marginalization=True, precision="fp32")
    out["mb_128s_setup_s"] = time.perf_counter() - t0
    d = draws_near(inj, batch, np.random.default_rng(13), ifos, lk, 2e-6)
    rows = torch.from_numpy(lk.rows_from_parameters(d)).cuda()
    N, j_first, n_window = lk._time_window
    out.update({"mb_128s_rows": batch, "mb_128s_banded_points": int(len(lk.setup.banded_frequency_points)),
                "mb_128s_transform_length": N, "mb_128s_window_bins": n_window})
    for prec in ("fp32", "fp64"):
        dt = timed(lambda: lk.log_likelihood_ratio_batch(rows, precision=prec), steps)
        out["mb_128s_%s_evals_per_s" % prec] = batch / dt
    m32 = lk.log_likelihood_ratio_batch(rows, precision="fp32").cpu().numpy()
    m64 = lk.log_likelihood_ratio_batch(rows, precision="fp64").cpu().numpy()
    out["mb_128s_fp32_vs_fp64_max"] = float(np.max(np.abs(m32 - m64)))
    olk = OracleMultiband(oracle_event(ifos), "taylorf2", lk.setup,
                          waveform_arguments=dict(wfg.waveform_arguments),
                          time_marginalization=True, distance_marginalization=True,
                          phase_marginalization=True, priors=dict(priors),
                          distance_table=lk._dist_margd_loglikelihood_array)
    ref = np.array([olk.log_likelihood_ratio(p) for p in draws_as_dicts(d, range(4))])
    out["mb_128s_fp64_vs_oracle_4rows"] = float(np.max(np.abs(m64[:4] - ref)))
    out["mb_128s_fp32_vs_oracle_4rows"] = float(np.max(np.abs(m32[:4] - ref)))
    del lk
    full = GravitationalWaveTransient(ifos, wfg, time_marginalization=True,
                                      distance_marginalization=True, phase_marginalization=True,
                                      priors=priors.copy(), precision="fp32")
    rows = torch.from_numpy(full.rows_from_parameters(d)).cuda()
    dt = timed(lambda: full.log_likelihood_ratio_batch(rows), steps)
    out["mb_128s_full_tdp_fp32_evals_per_s"] = batch / dt
    return out


def bench_roq_weights(n_basis=400, duration=128.0, fs=4096.0):
    """Building ROQ linear weights (roq.py:849-918) for a 400-element basis
    on a 128 s, 4096 Hz segment, two detectors: the GPU GEMM form against
    bilby's form (one numpy inverse FFT of the segment per basis element and
    detector), the latter timed on a few basis elements and scaled."""
    import types
    import numpy as np
    from bilby_amd import roq_weights as rw
    from oracle import core
    from oracle import roq_weights as orw
    rng = np.random.default_rng(31)
    f = core.frequency_series(fs, duration)
    dets = []
    for name in ("H1", "L1"):
        psd = core.analytic_aligo_psd(f)
        strain, _ = core.noise_realisation(lambda x: core.analytic_aligo_psd(x), 0, np.inf,
                                           len(f), duration, rng)
        mask = core.frequency_mask(f, 20, fs / 2, fs)
        dets.append(types.SimpleNamespace(name=name, frequency_array=f, frequency_mask=mask,
                                          frequency_domain_strain=strain * mask,
                                          power_spectral_density_array=psd,
                                          minimum_frequency=20.0))

    class Ifos(list):
        pass
    ifos = Ifos(dets)
    ifos.duration, ifos.start_time, ifos.frequency_array = duration, 1000.0, f
    n_m = int(dets[0].frequency_mask.sum())
    lin = (rng.normal(size=(n_basis, n_m)) + 1j * rng.normal(size=(n_basis, n_m))) / np.sqrt(n_m)
    quad = rng.normal(size=(8, n_m))
    tmid = 1000.0 + duration - 2.0
    t0 = time.perf_counter()
    w = rw.build_roq_weights(ifos, lin, quad, np.arange(n_basis, dtype=float), np.arange(8.0),
                             tmid - 0.1, tmid + 0.1)
    gpu_s = time.perf_counter() - t0
    times = w["time_samples"]
    n_seg = int(round(duration / (times[1] - times[0])))
    k = 3
    t0 = time.perf_counter()
    ref = orw.linear_weights(f, dets[0].frequency_mask, dets[0].frequency_domain_strain,
                             dets[0].power_spectral_density_array, 20.0, duration, lin[:k], times)
    cpu_s = (time.perf_counter() - t0) * n_basis / k * len(dets)
    err = np.max(np.abs(w["H1_linear"][:, :k] - ref)) / np.max(np.abs(ref))
    return {"basis": n_basis, "masked_bins": n_m, "segment_samples": n_seg,
            "roq_times": len(times), "gpu_build_s": gpu_s,
            "numpy_ifft_build_s_estimated": cpu_s, "cpu_sample": "%d of %d basis elements, "
            "1 of 2 detectors, scaled" % (k, n_basis), "relative_error_vs_ifft": float(err)}


def bench_roq(steps):
    """ROQ, H1+L1, 400 linear / 120 quadratic nodes, 1000 ROQ times,
    time + phase marginalised, 65536 TaylorF2 draws: the contraction
    (65536 x 400) . (400 x 1000) per detector on MFMA."""
    import ctypes
    import numpy as np
    import torch
    from bilby_amd._lib import LIB, check
    from bilby_amd.configs import make_roq
    from bilby_amd.likelihood import get_context
    from bilby_amd.roq import ROQGravitationalWaveTransient
    from oracle.roq import OracleROQ
    from tests.helpers import draws_as_dicts, fixed_marginalised
    s = make_roq(n_draws=65536)
    w = s["weights"]
    opri = dict(s["priors"])
    priors = s["priors"].copy()
    lk = ROQGravitationalWaveTransient(s["interferometers"], s["waveform_generator"], priors,
                                       weights=w, time_marginalization=True,
                                       phase_marginalization=True, delta_tc=s["delta_tc"])
    draws = fixed_marginalised(s["draws"], priors)
    rows = torch.from_numpy(lk.rows_from_parameters(draws)).cuda()
    out = {}
    res = {}
    for mode in ("fp32", "bf16x3", "fp64"):
        st = steps if mode != "fp64" else 1
        dt = timed(lambda: lk.log_likelihood_ratio_batch(rows, precision=mode), st, warmup=1)
        out[mode + "_evals_per_s"] = rows.shape[0] / dt
        res[mode] = lk.log_likelihood_ratio_batch(rows, precision=mode).cpu().numpy()
    o = OracleROQ(["H1", "L1"], s["interferometers"].start_time, w, w["frequency_nodes_linear"],
                  w["frequency_nodes_quadratic"], opri, time_marginalization=True,
                  phase_marginalization=True, delta_tc=s["delta_tc"],
                  waveform_arguments=s["waveform_generator"].waveform_arguments)
    ref = np.array([o.log_likelihood_ratio(p) for p in draws_as_dicts(draws, range(64))])
    for mode in res:
        out[mode + "_max_abs_dlnL_vs_oracle_64rows"] = float(np.max(np.abs(res[mode][:64] - ref)))
        out[mode + "_vs_fp64_gpu_max_all_rows"] = float(np.max(np.abs(res[mode] - res["fp64"])))
    # the GEMM alone at the workload's shape: M = 65536, N = 4096 (2 detectors x 1000
    # complex times, padded), K = 800 (400 complex nodes)
    ctx = get_context(0)
    M, N, K = 65536, 4096, 800
    useful = 8.0 * 65536 * 400 * 1000 * 2
    a = torch.randn(M, K, device="cuda")
    b = torch.randn(N, K, device="cuda")
    c = torch.empty(M, N, device="cuda")
    ah, bh = a.to(torch.bfloat16), b.to(torch.bfloat16)
    al, bl = (a - ah.float()).to(torch.bfloat16), (b - bh.float()).to(torch.bfloat16)
    def g32():
        check(LIB.gwl_gemm(ctx, 1, ctypes.c_void_p(a.data_ptr()), None, ctypes.c_void_p(b.data_ptr()),
                           None, ctypes.c_void_p(c.data_ptr()), M, N, K, None))
    def g16():
        check(LIB.gwl_gemm(ctx, 2, ctypes.c_void_p(ah.data_ptr()), ctypes.c_void_p(al.data_ptr()),
                           ctypes.c_void_p(bh.data_ptr()), ctypes.c_void_p(bl.data_ptr()),
                           ctypes.c_void_p(c.data_ptr()), M, N, K, None))
    t32 = timed(g32, steps)
    t16 = timed(g16, steps)
    out["gemm_fp32_ms"] = t32 * 1e3
    out["gemm_fp32_tflops_useful"] = useful / t32 / 1e12
    out["gemm_fp32_mfma_frac"] = (2.0 * M * N * K) / t32 / 157.3e12
    out["gemm_bf16x3_ms"] = t16 * 1e3
    out["gemm_bf16x3_tflops_useful"] = useful / t16 / 1e12
    out["gemm_bf16x3_mfma_frac"] = (3 * 2.0 * M * N * K) / t16 / 2.5e15
    return out


def bench_sampler(lk, draws, seconds, workers=16, queue=1024):
    """bilby-style callers of the headline event: (i) a dynesty-style queue of
    ``queue`` proposals through BatchedPool.map (one batched call per queue,
    host conversion and copies included); (ii) ``workers`` spawned sampler
    processes, each calling RemoteLikelihood.log_likelihood_ratio one dict at
    a time (bilby's pool workers, base_sampler.py:772-800) against a
    LikelihoodService that owns the GPU."""
    import multiprocessing
    import numpy as np
    from bilby_amd.sampling import (BatchedPool, LikelihoodService, client_load,
                                    log_likelihood_wrapper)
    keys = ["mass_1", "mass_2", "ra", "dec", "theta_jn", "psi"]
    fixed = {k: float(np.asarray(v)[0]) for k, v in draws.items() if k not in keys}
    thetas = np.stack([np.asarray(draws[k], float) for k in keys], axis=1)
    pool = BatchedPool(lk, None, keys, True, fixed)
    pool.map(log_likelihood_wrapper, thetas[:queue])
    n_calls, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < 2.0:
        k = (n_calls * queue) % (len(thetas) - queue + 1)
        pool.map(log_likelihood_wrapper, thetas[k:k + queue])
        n_calls += 1
    t_pool = time.perf_counter() - t0
    out = {"batched_pool_queue": queue,
           "batched_pool_evals_per_s": n_calls * queue / t_pool,
           "batched_pool_ms_per_queue": 1e3 * t_pool / n_calls}
    # the same queue size as device-resident rows (no host conversion or copies)
    import torch
    rows_t = torch.from_numpy(lk.rows_from_parameters(
        {k: np.asarray(v)[:queue] for k, v in draws.items()})).cuda()
    dt = timed(lambda: lk.log_likelihood_ratio_batch(rows_t), 50)
    out["device_rows_same_queue_evals_per_s"] = queue / dt
    out["batched_pool_over_device_rows"] = out["batched_pool_evals_per_s"] * dt / queue
    one = {k: np.asarray(v)[:1] for k, v in draws.items()}
    with LikelihoodService({"c2": lk}, hbm_budget=200e9, max_batch_rows=4096) as svc:
        client = svc.client("c2")
        client.log_likelihood_ratio({k: float(v[0]) for k, v in one.items()})
        ctx = multiprocessing.get_context("spawn")
        jobs = [(client, draws, seconds, 97 * w) for w in range(workers)]
        with ctx.Pool(workers) as wp:
            res = wp.starmap(client_load, jobs, chunksize=1)
    lat = np.concatenate([r[1] for r in res])
    n = sum(r[0] for r in res)
    idx = sorted({i for r in res for i in r[2]})
    direct = lk.log_likelihood_ratio_batch({k: np.asarray(v)[idx] for k, v in draws.items()})
    served = {i: v for r in res for i, v in r[2].items()}
    out.update(workers=workers, worker_seconds=seconds, worker_calls=n,
               worker_evals_per_s=sum(r[0] for r in res) / seconds,
               worker_p50_ms=1e3 * float(np.percentile(lat, 50)),
               worker_p99_ms=1e3 * float(np.percentile(lat, 99)),
               worker_vs_direct_max=float(np.max(np.abs(
                   np.array([served[i] for i in idx]) - direct))))
    return out


def bench_serving(main_lk, main_rows, duration, main_setup, main_draws):
    """32 samplers, seeded Poisson arrivals at 2000 requests/s, 1-1024 rows per
    request from a 60/30/10 mix of the TaylorF2 T+D+P event, uneven events and
    the ROQ event; everything resident; HBM budget 200 GB."""
    import numpy as np
    from bilby_amd.configs import make_roq, make_uneven_events
    from bilby_amd.likelihood import GravitationalWaveTransient
    from bilby_amd.roq import ROQGravitationalWaveTransient
    from bilby_amd.serving import LikelihoodServer, latency_summary, run_load
    from tests.helpers import fixed_marginalised
    server = LikelihoodServer(hbm_budget=200e9, max_batch_rows=4096)
    try:
        server.register("c2", main_lk)
        events = make_uneven_events(n_events=64, seed=97531, rows_range=(256, 4096))
        names, lks, pools, ev_draws = [], [], [], []
        for k, ev in enumerate(events):
            pri = ev["priors"].copy()
            lk = GravitationalWaveTransient(ev["interferometers"], ev["waveform_generator"],
                                            time_marginalization=True,
                                            distance_marginalization=True,
                                            phase_marginalization=True, priors=pri)
            names.append("c3_%d" % k)
            lks.append(lk)
            ev_draws.append(fixed_marginalised(ev["draws"], pri))
            pools.append(lk.rows_from_parameters(ev_draws[-1]))
        server.register_multi(names, lks)
        s = make_roq(n_draws=8192)
        priors = s["priors"].copy()
        roq = ROQGravitationalWaveTransient(s["interferometers"], s["waveform_generator"], priors,
                                            weights=s["weights"], time_marginalization=True,
                                            phase_marginalization=True, delta_tc=s["delta_tc"])
        server.register("c4", roq)
        roq_rows = roq.rows_from_parameters(fixed_marginalised(s["draws"], priors))
        mix = [(0.6, [("c2", main_rows)]), (0.3, list(zip(names, pools))), (0.1, [("c4", roq_rows)])]
        # warm the paths once
        for name, pool in (("c2", main_rows), (names[0], pools[0]), ("c4", roq_rows)):
            server.submit(name, pool[:8]).result()
        t_start = time.perf_counter()
        issued, wall = run_load(server, mix, n_samplers=32, rate=2000.0, duration=duration,
                                rows_range=(1, 1024), seed=2024)
        summ = latency_summary(server, since=t_start)
        for prefix in ("c2", "c3", "c4"):
            sub = latency_summary(server, since=t_start, prefix=prefix)
            if sub:
                summ["%s_p50_ms" % prefix] = sub["p50_ms"]
                summ["%s_p99_ms" % prefix] = sub["p99_ms"]
        summ["batches"], _ = server.stats()
        summ["evals_per_s"] = summ["rows"] / wall
        summ["requests_per_s"] = summ["requests"] / wall
        summ["resident_gb"] = server.resident_bytes / 1e9
        # audit 1% of the requests (at least 1) against the FP64 oracle: the
        # first and last row of each audited request
        from oracle.roq import OracleROQ
        from tests.helpers import draws_as_dicts, oracle_from_setup
        rng = np.random.default_rng(7)
        pick = rng.choice(len(issued), size=max(1, -(-len(issued) // 100)), replace=False)
        w = s["weights"]
        oracles = {"c4": OracleROQ(["H1", "L1"], s["interferometers"].start_time, w,
                                   w["frequency_nodes_linear"], w["frequency_nodes_quadratic"],
                                   dict(s["priors"]), time_marginalization=True,
                                   phase_marginalization=True, delta_tc=s["delta_tc"],
                                   waveform_arguments=s["waveform_generator"].waveform_arguments)}
        draws_of = {"c4": fixed_marginalised(s["draws"], priors), "c2": main_draws}
        draws_of.update(zip(names, ev_draws))
        setup_of = dict(zip(names, events))
        setup_of["c2"] = main_setup
        # one distance table for the uneven events (same distance prior and
        # phase flag); the headline event builds its own
        c3_table = None
        dmax = 0.0
        n_rows_audited = 0
        for i in pick:
            name, idx, fut = issued[i]
            if name not in oracles:
                c3 = name.startswith("c3_")
                olk, _ = oracle_from_setup(setup_of[name], distance_table=c3_table if c3 else None)
                if c3:
                    c3_table = olk._dist_margd_loglikelihood_array
                oracles[name] = olk
            # the request's first and last row, in request order
            pos = [0] if len(idx) == 1 else [0, len(idx) - 1]
            sel = [int(idx[k]) for k in pos]
            ref = np.array([oracles[name].log_likelihood_ratio(p)
                            for p in draws_as_dicts(draws_of[name], sel)])
            got = fut.result()
            dmax = max(dmax, float(np.max(np.abs(got[pos] - ref))))
            n_rows_audited += len(sel)
        summ["audit_requests"] = int(len(pick))
        summ["audit_fraction"] = len(pick) / max(1, len(issued))
        summ["audit_rows"] = n_rows_audited
        summ["audit_max_abs_dlnL_vs_fp64_oracle"] = dmax
        return summ
    finally:
        server.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--per-sample-calls", type=int, default=300)
    ap.add_argument("--audit", type=int, default=64, help="rows checked against the FP64 oracle")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-pool-seconds", type=float, default=8.0,
                    help="seconds of the process-pool CPU baseline (0 = skip)")
    ap.add_argument("--no-extra", action="store_true", help="headline workload only")
    ap.add_argument("--serve-seconds", type=float, default=5.0)
    ap.add_argument("--sampler-seconds", type=float, default=3.0,
                    help="seconds each of the 16 single-dict client processes runs")
    ap.add_argument("--only", default="",
                    help="comma-separated substrings: run only the extra workloads whose name "
                         "contains one of them")
    ap.add_argument("--profile", action="store_true",
                    help="counter runs: warmup + steps launches of the headline batch at --tile, "
                         "then 3 copies of 512 MiB (a plain copy kernel of known size)")
    ap.add_argument("--tile", default="0,0",
                    help="rows per workgroup, workgroups (or wavefronts) per CU for --profile; "
                         "0,0 = the default kernel choice")
    ap.add_argument("--transform", type=int, default=0,
                    help="time transform of the headline event (include/gwl.h gwl_transform): "
                         "0 automatic, 1 tiled, 2 wave per row")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only; --gpus %d is not supported (parameter rows "
              "would shard across GPUs, each holding the event data, with only the 8-byte "
              "results gathered)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import ctypes
    import numpy as np
    import torch
    from bilby_amd._lib import LIB, check
    from bilby_amd.configs import make_taylorf2_marginalised
    from bilby_amd.likelihood import GravitationalWaveTransient
    from tests.helpers import draws_as_dicts, fixed_marginalised

    dev = torch.device("cuda:0")
    log("setup: TaylorF2 H1+L1+V1 8 s @ 4096 Hz, T+D+P marginalised, batch", args.batch)
    setup = make_taylorf2_marginalised(n_draws=args.batch, seed=4321)
    priors = setup["priors"].copy()
    lk = GravitationalWaveTransient(setup["interferometers"], setup["waveform_generator"],
                                    time_marginalization=True, distance_marginalization=True,
                                    phase_marginalization=True, priors=priors, precision="fp32")
    draws = fixed_marginalised(setup["draws"], priors)
    rows = torch.from_numpy(lk.rows_from_parameters(draws)).to(dev)
    check(LIB.gwl_event_reserve(lk._event, args.batch))
    from bilby_amd.autotune import autotune, set_tile, set_transform
    set_transform(lk, args.transform)
    if args.profile:
        set_tile(lk, *[int(v) for v in args.tile.split(",")])
        for _ in range(args.warmup + args.steps):
            lk.log_likelihood_ratio_batch(rows)
        torch.cuda.synchronize()
        src = torch.empty(512 << 20, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("profile: %d headline launches (%d rows, %d bytes/row algorithmic), 3 copies of %d bytes"
            % (args.warmup + args.steps, args.batch, bytes_per_eval(lk), src.numel()))
        return
    tile, sweep = autotune(lk, rows, use_cache=False)
    log("autotune (rows/workgroup, workgroups/CU) -> ms:", sweep, "chosen", tile)

    # accuracy: float32 bulk vs float64 on every row
    out32 = lk.log_likelihood_ratio_batch(rows, precision="fp32")
    out64 = lk.log_likelihood_ratio_batch(rows, precision="fp64")
    torch.cuda.synchronize()
    d_all = (out32 - out64).abs().cpu().numpy()
    log("fp32 vs fp64 (all rows): max %.3g p99 %.3g" % (d_all.max(), np.percentile(d_all, 99)))

    # timed region
    stream = torch.cuda.current_stream(dev)
    for _ in range(args.warmup):
        lk.log_likelihood_ratio_batch(rows)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        lk.log_likelihood_ratio_batch(rows)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    evals_per_s = args.batch * args.steps / wall
    log("fp32 batched: %.3f ms/step, %.4g evals/s (events %.3f ms/step)"
        % (ms_per_step, evals_per_s, ev0.elapsed_time(ev1) / args.steps))

    # dominant kernel duration with HIP events on the launch stream
    check(LIB.gwl_event_timing(lk._event, 1))
    for _ in range(args.steps):
        lk.log_likelihood_ratio_batch(rows)
    torch.cuda.synchronize()
    tot = ctypes.c_double()
    cnt = ctypes.c_int64()
    check(LIB.gwl_event_timing_read(lk._event, ctypes.byref(tot), ctypes.byref(cnt)))
    check(LIB.gwl_event_timing(lk._event, 0))
    k_ms = tot.value / max(1, cnt.value)
    bpe = bytes_per_eval(lk)
    achieved = args.batch * bpe / (k_ms * 1e-3) / 1e9
    # the transform that ran: the tile choice (include/gwl.h)
    wave = args.transform == 2 or (args.transform == 0 and tile[0] == 0)
    kname = "k_time_w" if wave else "k_time"
    traffic_key = ("k_time_w" if wave else "k_time_tiled") + "_taylorf2_tdp_batch%d" % args.batch
    traffic = measured_traffic(traffic_key)
    log("%s: %.3f ms/launch, %.1f GB/s algorithmic, measured traffic %s B/launch"
        % (kname, k_ms, achieved, traffic))

    # per-sample float64 GPU path (one row per launch), device-resident row
    set_tile(lk, 0, 0)
    one = rows[:1].contiguous()
    for _ in range(10):
        lk.log_likelihood_ratio_batch(one, precision="fp64")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.per_sample_calls):
        lk.log_likelihood_ratio_batch(rows[i % args.batch:i % args.batch + 1], precision="fp64")
        torch.cuda.synchronize()
    per_sample_rate = args.per_sample_calls / (time.perf_counter() - t0)
    # the same through the bilby dict API (host dict -> float)
    lk64 = lk
    dicts = draws_as_dicts(draws, range(64))
    lk64.precision = 0
    t0 = time.perf_counter()
    for i in range(args.per_sample_calls):
        lk64.log_likelihood_ratio(dicts[i % 64])
    dict_rate = args.per_sample_calls / (time.perf_counter() - t0)
    lk64.precision = 1
    log("fp64 per-sample path: %.1f evals/s (device rows), %.1f evals/s (dict API)"
        % (per_sample_rate, dict_rate))

    result = {
        "metric": METRIC,
        "value": evals_per_s,
        "unit": "lnL evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded coloured Gaussian noise + injected TaylorF2; analytic aLIGO PSD)",
        "config": {"workload": "TaylorF2 non-spinning, H1+L1+V1, 8 s @ 4096 Hz (16385 bins), "
                               "time+distance+phase marginalised, batch of prior draws",
                   "batch": args.batch, "n_freq": 16385, "detectors": 3},
        "roofline": roofline_block(kname, wave, k_ms, achieved, bpe, traffic, traffic_key),
        "autotune": {"chosen": list(tile), "ms_by_tile": sweep},
        "accuracy": {"fp32_vs_fp64_gpu_max": float(d_all.max()),
                     "fp32_vs_fp64_gpu_p99": float(np.percentile(d_all, 99))},
        "per_sample_fp64": {"device_rows_evals_per_s": per_sample_rate,
                            "dict_api_evals_per_s": dict_rate,
                            "batched_fp32_over_per_sample_fp64": evals_per_s / per_sample_rate},
    }

    extra = {}
    if not args.no_extra:
        for name, fn in (("sine_gaussian_plain_4s_2048Hz_10k", lambda: bench_sine_gaussian(10)),
                         ("uneven_64_events_persistent_graph", lambda: bench_uneven(5)),
                         ("roq_400_120_nodes_1000_times_65536", lambda: bench_roq(3)),
                         ("relative_binning_dp_event_65536", lambda: bench_relbin(5)),
                         ("calibration_1000_curves_dp_event", lambda: bench_calibration(3)),
                         ("multiband_128s_light_binary", lambda: bench_multiband(5)),
                         ("binned_time_marginalised_tdp", lambda: bench_binned_time(5)),
                         ("roq_weights_400_basis_128s", lambda: bench_roq_weights()),
                         ("serving_32_samplers_2000_rps", lambda: bench_serving(
                             lk, rows.cpu().numpy(), args.serve_seconds, setup, draws)),
                         ("sampler_callers_pool_and_16_processes", lambda: bench_sampler(
                             lk, draws, args.sampler_seconds))):
            if args.only and not any(k in name for k in args.only.split(",")):
                continue
            log("extra:", name)
            try:
                extra[name] = fn()
            except Exception as exc:  # report, keep the headline line
                extra[name] = {"error": repr(exc)}
            log(name, json.dumps(extra[name])[:400])
    result_extra = extra

    if not args.no_cpu_baseline:
        log("cpu baseline (oracle, 1 core, %.0f s)" % args.cpu_seconds)
        olk, cb = cpu_baseline(setup, priors, args.cpu_seconds)
        cb["cpu"] = cpu_model()
        result["cpu_baseline"] = cb
        if args.cpu_pool_seconds > 0:
            workers = max(1, min(16, os.cpu_count() or 1))
            log("cpu baseline (oracle, %d processes, %.0f s)" % (workers, args.cpu_pool_seconds))
            from tests.helpers import fixed_marginalised
            pdicts = draws_as_dicts(fixed_marginalised(setup["draws"], priors), range(512))
            result["cpu_baseline_pool"] = cpu_baseline_pool(olk, pdicts, args.cpu_pool_seconds,
                                                            workers)
            log("cpu pool: %.1f evals/s" % result["cpu_baseline_pool"]["value"])
        # audit sample vs the FP64 oracle
        n_a = min(args.audit, args.batch)
        ref = np.array([olk.log_likelihood_ratio(p) for p in draws_as_dicts(draws, range(n_a))])
        o32 = out32[:n_a].cpu().numpy()
        o64 = out64[:n_a].cpu().numpy()
        result["accuracy"].update(
            audit_rows=n_a,
            fp32_vs_oracle_max=float(np.max(np.abs(o32 - ref))),
            fp64_vs_oracle_max=float(np.max(np.abs(o64 - ref))))
        log("audit vs oracle: fp32 max %.3g, fp64 max %.3g"
            % (result["accuracy"]["fp32_vs_oracle_max"], result["accuracy"]["fp64_vs_oracle_max"]))
    else:
        result["cpu_baseline"] = None
    result["configs"] = result_extra
    srv = result_extra.get("serving_32_samplers_2000_rps", {})
    if "p50_ms" in srv:
        result["latency_under_load"] = {"p50_ms": srv["p50_ms"], "p99_ms": srv["p99_ms"],
                                        "evals_per_s": srv["evals_per_s"]}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
