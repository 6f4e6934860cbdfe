"""Tile autotuner for the likelihood kernels.

For an event and a batch shape it times the candidate tiles (rows of the
parameter batch per workgroup, target resident workgroups per CU; rows 0
selects the wave-per-row TaylorF2 time transform where the event allows it,
and the second number is then the target wavefronts per CU) on the
device with HIP events, keeps the fastest, and caches the choice per
(device, source model, marginalisation flags, detectors, bin-count bucket,
batch bucket) in a JSON file so later processes reuse it.
"""
import ctypes
import json
import math
import os

import torch

from ._lib import LIB, check

DEFAULT_CANDIDATES_TIME = [(0, 8), (0, 16), (0, 32), (0, 64), (8, 2), (8, 4), (8, 8), (4, 4),
                           (4, 8), (2, 8), (1, 8), (8, 16)]
DEFAULT_CANDIDATES_PLAIN = [(4, 0), (8, 0), (16, 0)]
CACHE_PATH = os.environ.get("BILBY_AMD_AUTOTUNE_CACHE",
                            os.path.join(os.path.expanduser("~"), ".cache", "bilby_amd",
                                         "autotune.json"))


def bucket_key(likelihood, n_rows):
    ifos = likelihood.interferometers
    n_freq = len(ifos.frequency_array)
    dev = torch.cuda.get_device_name(likelihood.device)
    return "|".join(str(x) for x in (
        dev, likelihood.waveform_generator.approximant, likelihood._flags(), len(ifos),
        int(math.log2(max(2, n_freq - 1))), int(math.log2(max(1, n_rows))), likelihood.precision))


def _load_cache():
    try:
        with open(CACHE_PATH) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def _save_cache(cache):
    try:
        os.makedirs(os.path.dirname(CACHE_PATH), exist_ok=True)
        tmp = CACHE_PATH + ".tmp"
        with open(tmp, "w") as fh:
            json.dump(cache, fh, indent=1, sort_keys=True)
        os.replace(tmp, CACHE_PATH)
    except OSError:
        pass


def set_tile(likelihood, rows_per_workgroup, workgroups_per_cu):
    check(LIB.gwl_event_set_tile(likelihood._event, int(rows_per_workgroup),
                                 int(workgroups_per_cu)), "gwl_event_set_tile")


def set_transform(likelihood, transform):
    """0 automatic, 1 tiled, 2 wave per row (bilby_amd.abi)."""
    check(LIB.gwl_event_set_transform(likelihood._event, int(transform)), "gwl_event_set_transform")


def time_tile(likelihood, rows_t, tile, repeats=5, precision=None):
    """Average milliseconds of one batched evaluation with ``tile``."""
    set_tile(likelihood, *tile)
    likelihood.log_likelihood_ratio_batch(rows_t, precision=precision)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(repeats):
        likelihood.log_likelihood_ratio_batch(rows_t, precision=precision)
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / repeats


def autotune(likelihood, rows_t, candidates=None, repeats=5, use_cache=True, precision=None):
    """Pick and apply the fastest tile for this likelihood and batch shape.
    Returns (tile, {tile: ms})."""
    key = bucket_key(likelihood, rows_t.shape[0])
    cache = _load_cache() if use_cache else {}
    if use_cache and key in cache:
        tile = tuple(cache[key]["tile"])
        set_tile(likelihood, *tile)
        return tile, cache[key].get("timings", {})
    if candidates is None:
        candidates = (DEFAULT_CANDIDATES_TIME if likelihood.time_marginalization
                      else DEFAULT_CANDIDATES_PLAIN)
    timings = {}
    for tile in candidates:
        timings[tile] = time_tile(likelihood, rows_t, tile, repeats, precision)
    best = min(timings, key=timings.get)
    set_tile(likelihood, *best)
    if use_cache:
        cache[key] = {"tile": list(best),
                      "timings": {"%d,%d" % t: ms for t, ms in timings.items()}}
        _save_cache(cache)
    return best, {"%d,%d" % t: ms for t, ms in timings.items()}
