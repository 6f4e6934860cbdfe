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


MULTI_CANDIDATES = (8, 16, 32, 64)


def multi_bucket_key(likelihoods, rows):
    """Cache key of a multi-event set: the device and every event's
    (source, flags, detectors, log2 bins, log2 rows), sorted."""
    dev = torch.cuda.get_device_name(likelihoods[0].device)
    shapes = sorted("%d:%d:%d:%d:%d" % (lk.waveform_generator.approximant, lk._flags(),
                                          len(lk.interferometers),
                                          int(math.log2(max(2, len(lk.interferometers.frequency_array) - 1))),
                                          int(math.log2(max(1, n))))
                    for lk, n in zip(likelihoods, rows))
    return dev + "|multi|" + ",".join(shapes)


def set_multi_steps(likelihood, steps):
    check(LIB.gwl_event_set_multi_steps(likelihood._event, int(steps)),
          "gwl_event_set_multi_steps")


def autotune_multi(multi, candidates=MULTI_CANDIDATES, repeats=2, use_cache=True):
    """Tile of a multi-event launch (bilby_amd.multi.MultiEventLikelihood
    whose rows are set): the decimated sub-sequence steps per work item,
    chosen per bin-count bucket by coordinate descent over the buckets
    (each bucket's candidates timed with the others held), cached per event
    set, and stored on every event (gwl_event_set_multi_steps) so that later
    multi-event sets and server groups of these events start from it.
    Returns ({bins: steps}, {"bins:steps": ms})."""
    lks = multi.likelihoods
    key = multi_bucket_key(lks, multi.rows)
    bins = [len(lk.interferometers.frequency_array) - 1 for lk in lks]
    cache = _load_cache() if use_cache else {}
    sweep = {}
    if use_cache and key in cache:
        steps_of = {int(b): int(s) for b, s in cache[key]["steps"].items()}
        sweep = cache[key].get("timings", {})
    else:
        steps_of = {b: 32 for b in sorted(set(bins))}
        for b in sorted(steps_of):
            best = None
            for cand in candidates:
                steps_of[b] = cand
                multi.set_tile([steps_of[x] for x in bins])
                multi.run()
                torch.cuda.synchronize()
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(multi.stream)
                for _ in range(repeats):
                    multi.run(synchronize=False)
                e1.record(multi.stream)
                torch.cuda.synchronize()
                ms = e0.elapsed_time(e1) / repeats
                sweep["%d:%d" % (b, cand)] = ms
                if best is None or ms < best[0]:
                    best = (ms, cand)
            steps_of[b] = best[1]
        if use_cache:
            cache[key] = {"steps": {str(b): s for b, s in steps_of.items()}, "timings": sweep}
            _save_cache(cache)
    multi.set_tile([steps_of[x] for x in bins])
    for lk, b in zip(lks, bins):
        set_multi_steps(lk, steps_of[b])
    return steps_of, sweep


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
