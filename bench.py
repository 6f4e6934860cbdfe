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
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only; --gpus %d is not supported in this build "
              "(rows would shard across GPUs with no exchange, see DESIGN.md)" % args.gpus,
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
    from bilby_amd.autotune import autotune, set_tile
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
    log("k_time: %.3f ms/launch, %.1f GB/s algorithmic" % (k_ms, achieved))

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
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": 8000.0, "unit": "GB/s",
                     "frac": achieved / 8000.0, "traffic": None,
                     "kernel": "k_time (fused waveform + projection + decimated LDS FFT)",
                     "kernel_ms": k_ms, "bytes_per_eval": bpe},
        "autotune": {"chosen": list(tile), "ms_by_tile": sweep},
        "accuracy": {"fp32_vs_fp64_gpu_max": float(d_all.max()),
                     "fp32_vs_fp64_gpu_p99": float(np.percentile(d_all, 99))},
        "per_sample_fp64": {"device_rows_evals_per_s": per_sample_rate,
                            "dict_api_evals_per_s": dict_rate,
                            "batched_fp32_over_per_sample_fp64": evals_per_s / per_sample_rate},
    }

    if not args.no_cpu_baseline:
        log("cpu baseline (oracle, 1 core, %.0f s)" % args.cpu_seconds)
        olk, cb = cpu_baseline(setup, priors, args.cpu_seconds)
        result["cpu_baseline"] = cb
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
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
