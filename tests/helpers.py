"""Builders shared by the tests: the FP64 oracle on the same host arrays as
the GPU likelihood."""
import numpy as np

from oracle.likelihood import OracleDetector, OracleEvent, OracleGWTransient

_TABLES = {}


def oracle_from_setup(setup, distance_table=None, **overrides):
    """The FP64 oracle of a configs.make_* setup.  distance_table: an
    oracle-built lookup table for the same distance prior (the table depends
    only on the prior and the phase flag), to skip rebuilding it."""
    ifos = setup["interferometers"]
    dets = [OracleDetector(ifo.name, ifo.frequency_array, ifo.frequency_domain_strain,
                           ifo.power_spectral_density_array, ifo.frequency_mask)
            for ifo in ifos]
    ev = OracleEvent(dets, ifos.duration, ifos.sampling_frequency, ifos.start_time)
    wfg = setup["waveform_generator"]
    source = "taylorf2" if wfg.approximant == 2 else "sinegaussian"
    kw = dict(time_marginalization=setup["time_marginalization"],
              distance_marginalization=setup["distance_marginalization"],
              phase_marginalization=setup["phase_marginalization"])
    kw.update(overrides)
    priors = dict(setup["priors"])
    table = distance_table
    if kw["distance_marginalization"] and table is None:
        key = (kw["phase_marginalization"], id(setup["priors"]["luminosity_distance"]))
        table = _TABLES.get(key)
    lk = OracleGWTransient(ev, source, waveform_arguments=dict(wfg.waveform_arguments),
                           priors=priors, distance_table=table, **kw)
    if kw["distance_marginalization"] and distance_table is None:
        _TABLES[key] = lk._dist_margd_loglikelihood_array
    return lk, priors


def draws_as_dicts(draws, idx):
    return [{k: float(v[i]) for k, v in draws.items()} for i in idx]


def fixed_marginalised(draws, priors):
    d = dict(draws)
    n = len(next(iter(d.values())))
    for key in ("geocent_time", "phase", "luminosity_distance"):
        if isinstance(priors.get(key), float):
            d[key] = np.full(n, priors[key])
    return d
