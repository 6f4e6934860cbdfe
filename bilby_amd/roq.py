"""Drop-in ROQGravitationalWaveTransient on the MI355X.

bilby/gw/likelihood/roq.py:14-300 constructor semantics for precomputed
weights (a dict with 'time_samples', '<ifo>_linear' [T x L],
'<ifo>_quadratic' [Q], 'frequency_nodes_linear', 'frequency_nodes_quadratic'),
the TaylorF2 source model at the nodes, and time / phase / distance
marginalisation.  The linear sums over the ROQ time samples are one complex
GEMM on MFMA (float32, or split-bfloat16 x3), float64 for the reference.
"""
import copy
import ctypes

import numpy as np
import torch

from . import _lib
from . import abi as _abi
from ._lib import LIB, check, dptr
from .detector import InterferometerList
from .likelihood import (GravitationalWaveTransient, _batch_rows, _is_uniform, density_table,
                         distance_interpolant, get_context, grid_time_posterior)
from .prior import Uniform
from .waveform_generator import taylorf2_flags

MODES = {"fp64": 0, "fp32": 1, "bf16x3": 2}


def load_weights(filename, interferometers, waveform_generator, format=None):
    """roq.py:1107-1160 for the npz format (read without pickles): the time
    samples and each detector's weights, with the frequency nodes from the
    waveform arguments when the file does not hold them.  HDF5 needs h5py,
    which is not available."""
    if format is None:
        format = filename.split(".")[-1]
    if format not in ("npz", "hdf5"):
        raise IOError("Format %s not recognized." % format)
    if format == "hdf5":
        raise NotImplementedError("h5py is not available; save the weights as npz")
    weights = dict(np.load(filename))
    wfa = waveform_generator.waveform_arguments
    for kind in ("linear", "quadratic"):
        key = "frequency_nodes_" + kind
        if key not in weights:
            if key not in wfa:
                raise ValueError("%s is neither in %s nor in the waveform arguments"
                                 % (key, filename))
            weights[key] = np.asarray(wfa[key], dtype=float)
        for ifo in interferometers:
            w = np.asarray(weights["%s_%s" % (ifo.name, kind)])
            if kind == "linear" and w.ndim == 3:    # a one-basis list as bilby stores it
                w = w[0]
            if kind == "quadratic" and w.ndim == 2:
                w = w[0]
            weights["%s_%s" % (ifo.name, kind)] = w
    return weights


class ROQGravitationalWaveTransient:
    def __init__(self, interferometers, waveform_generator, priors, weights=None,
                 linear_matrix=None, quadratic_matrix=None, roq_params=None,
                 roq_params_check=True, roq_scale_factor=1, distance_marginalization=False,
                 phase_marginalization=False, time_marginalization=False, jitter_time=True,
                 delta_tc=None, distance_marginalization_lookup_table=None,
                 reference_frame="sky", time_reference="geocenter", parameter_conversion=None,
                 precision="fp32", device=0):
        if isinstance(weights, str):
            weights = load_weights(weights, InterferometerList(interferometers),
                                   waveform_generator)
        if weights is None:
            if linear_matrix is None or quadratic_matrix is None:
                raise ValueError("pass weights, or linear_matrix and quadratic_matrix")
            if roq_params is not None:
                raise NotImplementedError(
                    "bases restricted by roq_params (intersection with the ROQ frequencies) "
                    "are not supported; give bases on each detector's masked frequencies")
            weights = self._weights_from_bases(interferometers, waveform_generator, priors,
                                               linear_matrix, quadratic_matrix, time_reference,
                                               device)
        from .frames import sky_frame
        # base.py:170-181: a detector time reference cannot be marginalised over
        self._frame = sky_frame(reference_frame, time_reference)
        self.reference_frame = reference_frame
        self.time_reference = "geocent" if "geocent" in time_reference else time_reference
        if "geocent" not in time_reference and time_marginalization:
            time_marginalization = False
            jitter_time = False
        if roq_scale_factor != 1:
            raise NotImplementedError("roq_scale_factor != 1 is not supported")
        self.interferometers = InterferometerList(interferometers)
        self.waveform_generator = waveform_generator
        self.weights = weights
        self.priors = priors.copy()
        self.time_marginalization = time_marginalization
        self.phase_marginalization = phase_marginalization
        self.distance_marginalization = distance_marginalization
        self.jitter_time = jitter_time if time_marginalization else False
        self.precision = MODES[precision] if isinstance(precision, str) else int(precision)
        self.device = device
        self._marginalized_parameters = []
        ts = np.asarray(weights["time_samples"], float)
        self.frequency_nodes_linear = np.asarray(weights["frequency_nodes_linear"], float)
        self.frequency_nodes_quadratic = np.asarray(weights["frequency_nodes_quadratic"], float)
        if time_marginalization:
            if delta_tc is None:
                delta_tc = ts[1] - ts[0]
            tp = self.priors["geocent_time"]
            if not (np.isfinite(tp.minimum) and np.isfinite(tp.maximum)):
                raise ValueError("ROQ time marginalisation needs a geocent_time prior with finite "
                                 "bounds (its grid spans them, roq.py:320-331)")
            n = int(np.ceil((tp.maximum - tp.minimum) / delta_tc))
            self._delta_tc = (tp.maximum - tp.minimum) / n
            self._times = tp.minimum + self._delta_tc / 2. + np.arange(n) * self._delta_tc
            self._check_time_coverage(ts)
            priors["geocent_time"] = float(self.interferometers.start_time)
            if self.jitter_time:
                priors["time_jitter"] = Uniform(-self._delta_tc / 2, self._delta_tc / 2,
                                                name="time_jitter", boundary="periodic")
            self._marginalized_parameters.append("geocent_time")
        if phase_marginalization:
            priors["phase"] = 0.0
            self._marginalized_parameters.append("phase")
        self._ref_dist = 1.0
        if distance_marginalization:
            prior = self.priors["luminosity_distance"]
            self._distance_array = np.linspace(prior.minimum, prior.maximum, int(1e4))
            self.distance_prior_array = np.array([prior.prob(d) for d in self._distance_array])
            self._ref_dist = float(prior.rescale(0.5))
            priors["luminosity_distance"] = float(self._ref_dist)
            self._marginalized_parameters.append("luminosity_distance")
        self._build(ts, delta_tc, distance_marginalization_lookup_table)

    def save_weights(self, filename, format="hdf5"):
        """roq.py:1055-1105: the weights in one file; "npz" holds the time
        samples and each detector's linear and quadratic weights (the
        frequency nodes stay in the waveform arguments, as in the
        reference).  HDF5 needs h5py, which is not available."""
        if format not in ("hdf5", "npz"):
            raise IOError("Format %s not recognized." % format)
        if format not in filename:
            filename += "." + format
        if format == "hdf5":
            raise NotImplementedError("h5py is not available; use format='npz'")
        out = {"time_samples": np.asarray(self.weights["time_samples"])}
        for ifo in self.interferometers:
            for kind in ("linear", "quadratic"):
                key = "%s_%s" % (ifo.name, kind)
                out[key] = np.asarray(self.weights[key])
        np.savez(filename, **out)

    def load_weights(self, filename, format=None):
        """roq.py:1107-1160 (npz; see the module-level load_weights)."""
        return load_weights(filename, self.interferometers, self.waveform_generator, format)

    @staticmethod
    def _parse_basis(basis, basis_type):
        """roq.py:334-366: an array [n_freq, n_basis] (stored transposed) or a
        .npy file holding [n_basis, n_freq]; HDF5 basis files need h5py,
        which is not available."""
        if isinstance(basis, str):
            if basis.endswith(".npy"):
                return np.load(basis)
            raise NotImplementedError("%s basis file %s: only .npy files are read"
                                      % (basis_type, basis))
        if isinstance(basis, np.ndarray):
            return basis.T
        raise TypeError("basis needs to be str or np.ndarray")

    @classmethod
    def _weights_from_bases(cls, interferometers, waveform_generator, priors, linear_matrix,
                            quadratic_matrix, time_reference, device):
        """roq.py:736-847 (_set_weights) for one linear and one quadratic
        basis on each detector's masked frequencies: the time resolution of
        roq.py:1165-1229, the time samples covering the time prior, and the
        weights from bilby_amd.roq_weights (linear weights on the GPU).  The
        frequency nodes come from the waveform arguments
        (frequency_nodes_linear / frequency_nodes_quadratic), where the
        reference's ROQ source model reads them."""
        from . import roq_weights as rw
        wfa = waveform_generator.waveform_arguments
        if "frequency_nodes_linear" not in wfa or "frequency_nodes_quadratic" not in wfa:
            raise ValueError("the waveform arguments must hold frequency_nodes_linear and "
                             "frequency_nodes_quadratic")
        ifos = InterferometerList(interferometers)
        key = "geocent_time" if "geocent" in time_reference else "%s_time" % time_reference
        tp = priors[key]
        return rw.build_roq_weights(ifos, cls._parse_basis(linear_matrix, "linear"),
                                    cls._parse_basis(quadratic_matrix, "quadratic"),
                                    wfa["frequency_nodes_linear"], wfa["frequency_nodes_quadratic"],
                                    float(tp.minimum), float(tp.maximum), device=device)

    def _check_time_coverage(self, ts):
        """Every time the interpolation reads (prior window, geocentre delay,
        jitter, +-2 samples) must lie on the weights' time grid (the reference
        checks the basis against the prior in perform_roq_params_check)."""
        start = self.interferometers.start_time
        rel = self._times - start
        margin = 6371e3 * 1.01 / 299792458.0 + self._delta_tc
        lo, hi = rel.min() - margin, rel.max() + margin
        dts = ts[1] - ts[0]
        if lo - 2 * dts < ts[0] or hi + 3 * dts > ts[-1]:
            raise ValueError("ROQ time samples [%g, %g] do not cover the prior window [%g, %g]"
                             % (ts[0], ts[-1], lo, hi))

    def _build(self, ts, delta_tc, lookup_table):
        ifos = self.interferometers
        names = [ifo.name for ifo in ifos]
        lin = np.stack([np.asarray(self.weights[n + "_linear"], complex) for n in names])
        quad = np.stack([np.asarray(self.weights[n + "_quadratic"], complex) for n in names])
        lin_f = np.ascontiguousarray(lin).view(np.float64).copy()
        quad_f = np.ascontiguousarray(quad).view(np.float64).copy()
        tens = np.ascontiguousarray(np.stack([i.geometry.detector_tensor.ravel() for i in ifos]))
        vert = np.ascontiguousarray(np.stack([i.geometry.vertex for i in ifos]))
        fl = np.ascontiguousarray(self.frequency_nodes_linear)
        fq = np.ascontiguousarray(self.frequency_nodes_quadratic)
        wfa = self.waveform_generator.waveform_arguments
        d = _lib.RoqDesc()
        d.n_det = len(ifos)
        d.n_linear = len(fl)
        d.n_quadratic = len(fq)
        d.n_time_samples = len(ts)
        d.start_time = float(ifos.start_time)
        d.time_samples_start = float(ts[0])
        d.time_samples_step = float(ts[1] - ts[0])
        self._ts = np.ascontiguousarray(ts)
        d.time_samples = dptr(self._ts)
        d.linear_nodes = dptr(fl)
        d.quadratic_nodes = dptr(fq)
        d.linear_weights = dptr(lin_f)
        d.quadratic_weights = dptr(quad_f)
        d.detector_tensor = dptr(tens)
        d.vertex = dptr(vert)
        d.reference_frequency = float(wfa.get("reference_frequency", 20.0))
        d.wf_minimum_frequency = float(wfa.get("minimum_frequency", 20.0))
        flags = 0
        if self.time_marginalization:
            flags |= _lib.TIME_MARG | (_lib.JITTER_TIME if self.jitter_time else 0)
            tp = self.priors["geocent_time"]
            d.time_prior_minimum = float(tp.minimum)
            d.time_prior_maximum = float(tp.maximum)
            d.delta_tc = float(delta_tc)
        if self.phase_marginalization:
            flags |= _lib.PHASE_MARG
        if self.distance_marginalization:
            flags |= _lib.DISTANCE_MARG
        d.flags = flags | taylorf2_flags(wfa)
        d.reference_distance = float(self._ref_dist)
        ctx = get_context(self.device)
        h = ctypes.c_void_p()
        nbytes = ctypes.c_int64()
        check(LIB.gwl_roq_create(ctx, ctypes.byref(d), ctypes.byref(h), ctypes.byref(nbytes)),
              "gwl_roq_create")
        self._roq = h
        if self.time_marginalization and not _is_uniform(self.priors["geocent_time"]):
            # the prior's own support bounds the grid; its density weighs each time
            tp = self.priors["geocent_time"]
            t0, dt, dens = density_table(tp, float(tp.minimum), float(tp.maximum), self._delta_tc)
            check(LIB.gwl_roq_set_time_prior(self._roq, len(dens), t0, dt, dptr(dens)),
                  "gwl_roq_set_time_prior")
        self.hbm_bytes = int(nbytes.value)
        if self.distance_marginalization:
            table, (tx, ty, c, bbox) = distance_interpolant(
                self.device, self._distance_array, self.distance_prior_array, self._ref_dist,
                self.phase_marginalization, lookup_table)
            self._spline = (tx, ty, c, bbox)
            self._dist_margd_loglikelihood_array = table
            check(LIB.gwl_roq_set_distance_spline(self._roq, len(tx) - 4, len(ty) - 4, dptr(tx),
                                                  dptr(ty), dptr(c), dptr(bbox)),
                  "gwl_roq_set_distance_spline")

    def __del__(self):
        try:
            if getattr(self, "_roq", None) is not None:
                LIB.gwl_roq_destroy(self._roq)
                self._roq = None
        except Exception:
            pass

    @property
    def marginalized_parameters(self):
        return self._marginalized_parameters

    def _defaults(self):
        """Values the marginalised keys take when a dict omits them."""
        d = {}
        if self.time_marginalization:
            d["geocent_time"] = float(self.interferometers.start_time)
            d["time_jitter"] = 0.0
        if self.phase_marginalization:
            d["phase"] = 0.0
        if self.distance_marginalization:
            d["luminosity_distance"] = self._ref_dist
        return d

    # ------------------------------------------------------------------ reconstruction
    # base.py:502-742 with the ROQ's own time reconstruction (roq.py:1238-1281)
    calibration_marginalization = False
    generate_posterior_sample_from_marginalized_likelihood = (
        GravitationalWaveTransient.generate_posterior_sample_from_marginalized_likelihood)
    _distance_marginalized_values = GravitationalWaveTransient._distance_marginalized_values
    _distance_posterior = GravitationalWaveTransient._distance_posterior
    _phase_posterior = GravitationalWaveTransient._phase_posterior
    generate_distance_sample_from_marginalized_likelihood = (
        GravitationalWaveTransient.generate_distance_sample_from_marginalized_likelihood)
    generate_phase_sample_from_marginalized_likelihood = (
        GravitationalWaveTransient.generate_phase_sample_from_marginalized_likelihood)

    def _inner_products(self, parameters, n_times):
        """gwl_roq_inner_products for one parameter dict (float64 GEMM)."""
        q = dict(parameters)
        for key, value in self._defaults().items():
            q.setdefault(key, value)
        rows = torch.from_numpy(self.rows_from_parameters(q)[:1]).to("cuda:%d" % self.device)
        dh = torch.empty((1, max(1, n_times), 2), dtype=torch.float64, device=rows.device)
        hh = torch.empty(1, dtype=torch.float64, device=rows.device)
        st = torch.cuda.current_stream(rows.device).cuda_stream
        check(LIB.gwl_roq_inner_products(self._roq, ctypes.c_void_p(rows.data_ptr()), 1,
                                         _lib.NPARAM, MODES["fp64"], int(n_times),
                                         ctypes.c_void_p(dh.data_ptr()),
                                         ctypes.c_void_p(hh.data_ptr()), ctypes.c_void_p(st)),
              "gwl_roq_inner_products")
        dh = dh.cpu().numpy()[0]
        return dh[:, 0] + 1j * dh[:, 1], float(hh.cpu().numpy()[0]), q

    def _plain_inner_products(self, parameters):
        """<d|h>, <h|h> at the parameters' own time (roq.py:467-549 as
        base.py:686-694 calls it; the jitter is not added)."""
        q = dict(parameters)
        q["time_jitter"] = 0.0
        dh, hh, _ = self._inner_products(q, 0)
        if not np.isfinite(dh[0]):
            raise ValueError("the parameters' time lies outside the ROQ time samples, or the "
                             "waveform failed")
        return complex(dh[0]), hh

    def _time_posterior(self, parameters):
        """roq.py:1238-1281: d_inner_h_array on the time-marginalisation grid
        shifted by the jitter, the time-conditioned lnL, times the prior at
        _times plus the jitter, normalised."""
        if not self.time_marginalization:
            raise ValueError("the likelihood does not marginalise over time")
        dh, hh, q = self._inner_products(parameters, len(self._times))
        jitter = float(q.get("time_jitter", 0.0)) if self.jitter_time else 0.0
        times = self._times + jitter
        return times, grid_time_posterior(self, dh, hh, self._frame(q), times)

    def generate_time_sample_from_marginalized_likelihood(self, parameters, rng=None):
        """roq.py:1238-1281: one of the grid times, drawn with the
        posterior's probabilities."""
        times, post = self._time_posterior(parameters)
        return float((np.random.default_rng() if rng is None else rng).choice(times, p=post))

    def row_builder(self):
        """The host-side dict -> rows conversion of this likelihood
        (picklable; what worker-process clients use)."""
        return _abi.RowBuilder(self.waveform_generator.parameter_conversion, self._defaults(),
                               frame=self._frame)

    def rows_from_parameters(self, parameters):
        """Parameter dict (scalars or 1-D arrays) -> float64 rows [n, 16] in
        the column order of include/gwl.h (enum gwl_param)."""
        return self.row_builder()(parameters)

    def log_likelihood_ratio_batch(self, parameters, precision=None, stream=None):
        mode = self.precision if precision is None else (
            MODES[precision] if isinstance(precision, str) else int(precision))
        on_device = isinstance(parameters, torch.Tensor) and parameters.is_cuda
        if _batch_rows(parameters) == 0:
            return (torch.empty(0, dtype=torch.float64, device=parameters.device) if on_device
                    else np.zeros(0))
        if on_device:
            rows_t = parameters.to(torch.float64).contiguous()
        else:
            rows = (self.rows_from_parameters(parameters) if isinstance(parameters, dict)
                    else np.ascontiguousarray(parameters, dtype=np.float64))
            rows_t = torch.from_numpy(rows).to("cuda:%d" % self.device)
        n = rows_t.shape[0]
        out = torch.empty(n, dtype=torch.float64, device=rows_t.device)
        st = torch.cuda.current_stream(rows_t.device).cuda_stream if stream is None else stream
        check(LIB.gwl_roq_eval(self._roq, ctypes.c_void_p(rows_t.data_ptr()), n, _lib.NPARAM, mode,
                               ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(st)),
              "gwl_roq_eval")
        return out if on_device else out.cpu().numpy()

    def log_likelihood_ratio(self, parameters):
        return float(self.log_likelihood_ratio_batch(copy.copy(parameters))[0])
