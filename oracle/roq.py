"""FP64 restatement of bilby's ROQGravitationalWaveTransient
(bilby/gw/likelihood/roq.py) for the TaylorF2 node waveform.

Test infrastructure only (see oracle/__init__.py).
"""
import copy

import numpy as np
from scipy.special import logsumexp

from . import core


class OracleROQ:
    """roq.py:99-300 (constructor, weights dict as roq.py uses it:
    weights['time_samples'], weights[ifo + '_linear'] [T x L],
    weights[ifo + '_quadratic'] [Q]), :320-331 (time grid),
    :463-549 (calculate_snrs), :551-651 (five-sample interpolation)."""

    def __init__(self, detector_names, start_time, weights, linear_nodes, quadratic_nodes,
                 priors, time_marginalization=True, phase_marginalization=False,
                 jitter_time=True, delta_tc=None, waveform_arguments=None):
        self.names = list(detector_names)
        self.geom = {n: core.detector_geometry(n) for n in self.names}
        self.start_time = float(start_time)
        self.weights = weights
        self.fl = np.asarray(linear_nodes, float)
        self.fq = np.asarray(quadratic_nodes, float)
        self.priors = dict(priors)
        self.time_marginalization = time_marginalization
        self.phase_marginalization = phase_marginalization
        self.jitter_time = jitter_time if time_marginalization else False
        self.waveform_arguments = dict(waveform_arguments or {})
        if time_marginalization:
            self._delta_tc = delta_tc
            tcmin = self.priors["geocent_time"].minimum
            tcmax = self.priors["geocent_time"].maximum
            n = int(np.ceil((tcmax - tcmin) / self._delta_tc))
            self._delta_tc = (tcmax - tcmin) / n
            self._times = tcmin + self._delta_tc / 2. + np.arange(n) * self._delta_tc
            self._beam_pattern_reference_time = (tcmin + tcmax) / 2.
            priors["geocent_time"] = float(self.start_time)
        if phase_marginalization:
            priors["phase"] = 0.0

    def _node_waveform(self, p):
        """source.py:802-898: the model at the linear and quadratic nodes
        (lalsimulation's frequency-sequence TaylorF2 evaluates every node,
        source.py:1125-1128; with taylorf2_termination='isco' the nodes
        outside [minimum_frequency, f_isco] are zero)."""
        isco = self.waveform_arguments.get("taylorf2_termination") == "isco"
        kw = dict(minimum_frequency=(self.waveform_arguments.get("minimum_frequency", 20.0)
                                     if isco else 0.0),
                  reference_frequency=self.waveform_arguments.get("reference_frequency", 20.0),
                  maximum_frequency=np.inf,
                  taylorf2_termination="isco" if isco else "maximum_frequency")
        nodes = np.concatenate([self.fl, self.fq])
        h = core.taylorf2(nodes, p["mass_1"], p["mass_2"], p["luminosity_distance"], p["theta_jn"],
                          p["phase"], chi_1=p.get("chi_1", 0.0), chi_2=p.get("chi_2", 0.0), **kw)
        L = len(self.fl)
        return ({"plus": h["plus"][:L], "cross": h["cross"][:L]},
                {"plus": h["plus"][L:], "cross": h["cross"][L:]})

    @staticmethod
    def _interp_five_samples(time_samples, values, time):
        """roq.py:576-602."""
        r1 = (-values[0] + 8. * values[1] - 14. * values[2] + 8. * values[3] - values[4]) / 4.
        r2 = values[2] - 2. * values[3] + values[4]
        a = (time_samples[3] - time) / max(time_samples[1] - time_samples[0], 1e-12)
        b = 1. - a
        c = (a ** 3. - a) / 6.
        d = (b ** 3. - b) / 6.
        return a * values[2] + b * values[3] + c * r1 + d * r2

    def _d_inner_h_array(self, times, h_linear, name):
        """roq.py:604-651 (full matrix product branch)."""
        ts = self.weights["time_samples"]
        roq_dt = ts[1] - ts[0]
        tpr = (times - ts[0]) / roq_dt
        closest = np.floor(tpr).astype(int)
        W = self.weights[name + "_linear"]
        full = W @ h_linear.conj()
        m2, m1, z0, p1, p2 = (full[closest - 2], full[closest - 1], full[closest],
                              full[closest + 1], full[closest + 2])
        b = tpr - closest
        a = 1. - b
        c = (a ** 3. - a) / 6.
        d = (b ** 3. - b) / 6.
        r1 = (-m2 + 8. * m1 - 14. * z0 + 8. * p1 - p2) / 4.
        r2 = z0 - 2. * p1 + p2
        return a * z0 + b * p1 + c * r1 + d * r2

    def calculate_snrs(self, lin, quad, name, params):
        """roq.py:467-549."""
        tensor, vertex = self.geom[name]
        time_ref = self._beam_pattern_reference_time if self.time_marginalization \
            else params["geocent_time"]
        fp, fc = core.antenna_response(tensor, params["ra"], params["dec"], time_ref, params["psi"])
        h_linear = lin["plus"] * fp + lin["cross"] * fc
        h_quadratic = quad["plus"] * fp + quad["cross"] * fc
        optimal = np.vdot(np.abs(h_quadratic) ** 2, self.weights[name + "_quadratic"]).real
        dt = core.time_delay_from_geocenter(vertex, params["ra"], params["dec"], time_ref)
        dt_geocent = params["geocent_time"] - self.start_time
        ifo_time = dt_geocent + dt
        ts = self.weights["time_samples"]
        closest = int(np.floor((ifo_time - ts[0]) / (ts[1] - ts[0])))
        idx = np.clip(np.array([closest + k for k in (-2, -1, 0, 1, 2)]), 0, len(ts) - 1)
        in_bounds = (closest - 2 >= 0) & (closest + 2 < len(ts))
        d_tc = np.einsum("i,ji->j", h_linear.conj(), self.weights[name + "_linear"][idx])
        d_inner_h = self._interp_five_samples(ts[idx], d_tc, ifo_time)
        with np.errstate(divide="ignore"):
            d_inner_h = d_inner_h + np.log(in_bounds)
        arr = None
        if self.time_marginalization:
            ifo_times = self._times - self.start_time
            ifo_times = ifo_times + dt
            if self.jitter_time:
                ifo_times = ifo_times + params["time_jitter"]
            arr = self._d_inner_h_array(ifo_times, h_linear, name)
        return d_inner_h, optimal, arr

    def log_likelihood_ratio(self, parameters):
        """base.py:419-446 via roq.py:463-465."""
        params = copy.deepcopy(parameters)
        lin, quad = self._node_waveform(params)
        if self.time_marginalization and self.jitter_time:
            params["geocent_time"] += params["time_jitter"]
        d_inner_h, h_inner_h, arr = 0j, 0.0, None
        for name in self.names:
            dh, hh, a = self.calculate_snrs(lin, quad, name, params)
            d_inner_h += dh
            h_inner_h += hh
            if a is not None:
                arr = a if arr is None else arr + a
        if self.time_marginalization:
            times = self._times
            if self.jitter_time:
                times = self._times + params["time_jitter"]
            prior = self.priors["geocent_time"]
            mask = (times >= prior.minimum) & (times <= prior.maximum)
            times = times[mask]
            arr = arr[mask]
            tp = prior.prob(times) * self._delta_tc
            if self.phase_marginalization:
                ll = core.ln_i0(np.abs(arr)) - h_inner_h / 2
            else:
                ll = arr.real - h_inner_h / 2
            return float(logsumexp(ll, b=tp))
        if self.phase_marginalization:
            return float(core.ln_i0(abs(d_inner_h)) - h_inner_h / 2)
        return float(np.real(d_inner_h) - h_inner_h / 2)

    def inner_products(self, parameters):
        """base.py:686-694 via roq.py:467-549: <d|h>, <h|h> at the
        parameters' own time (no time jitter added), summed over detectors;
        what the distance and phase reconstruction use."""
        params = copy.deepcopy(parameters)
        lin, quad = self._node_waveform(params)
        d_inner_h, h_inner_h = 0j, 0.0
        for name in self.names:
            dh, hh, _ = self.calculate_snrs(lin, quad, name, params)
            d_inner_h += dh
            h_inner_h += hh
        return d_inner_h, h_inner_h

    def time_posterior_on_grid(self, parameters):
        """roq.py:1238-1281: the time-marginalisation grid (plus the
        jitter), the time-conditioned lnL from d_inner_h_array of the
        waveform at the parameters' geocent_time, times the prior,
        normalised (the reference then draws with rng.choice)."""
        params = copy.deepcopy(parameters)
        lin, quad = self._node_waveform(params)
        h_inner_h, arr = 0.0, None
        for name in self.names:
            _, hh, a = self.calculate_snrs(lin, quad, name, params)
            h_inner_h += hh
            arr = a if arr is None else arr + a
        if self.phase_marginalization:
            log_like = core.ln_i0(np.abs(arr)) - h_inner_h / 2
        else:
            log_like = arr.real - h_inner_h / 2
        times = self._times
        if self.jitter_time:
            times = times + params["time_jitter"]
        post = np.exp(log_like - np.max(log_like)) * self.priors["geocent_time"].prob(times)
        return times, post / np.sum(post)
