"""Drop-in ROQGravitationalWaveTransient on the MI355X.

bilby/gw/likelihood/roq.py:14-300 constructor semantics for precomputed
weights (a dict with 'time_samples', '<ifo>_linear' [T x L],
'<ifo>_quadratic' [Q], 'frequency_nodes_linear', 'frequency_nodes_quadratic';
lists of these with 'prior_range_linear' / 'prior_range_quadratic' for
several bases chosen per sample by prior range) or for bases (on the masked
frequencies, or on the ROQ grid of roq_params with roq_scale_factor, checked
against the data and a CBC prior), the TaylorF2 source model at the nodes,
and time / phase / distance marginalisation.  The linear sums over the ROQ
time samples are one complex GEMM on MFMA (float32, or split-bfloat16 x3),
float64 for the reference.
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


class BilbyROQParamsRangeError(Exception):
    """roq.py:1274-1275: the prior or the data lie outside a ROQ basis."""


def _check_roq_params(roq_params):
    """roq.py:86-107: None or a scalar structured array holding at least
    flow, fhigh and seglen."""
    if roq_params is None:
        return None
    if not isinstance(roq_params, np.ndarray) or roq_params.dtype.names is None:
        raise TypeError("roq_params must be None or a scalar structured numpy array")
    if roq_params.ndim != 0:
        raise ValueError("roq_params must be a scalar structured array; received shape %s"
                         % (roq_params.shape,))
    missing = {"flow", "fhigh", "seglen"}.difference(roq_params.dtype.names)
    if missing:
        raise ValueError("roq_params is missing required fields: " + ", ".join(sorted(missing)))
    return roq_params


def _as_basis_list(value):
    """One basis (an array) or several (a list / tuple of arrays, or an
    array with one more axis, as bilby stores them) -> list of arrays."""
    if isinstance(value, (list, tuple)):
        return [np.asarray(v) for v in value]
    return [np.asarray(value)]


def _params_from_rows(rows):
    """The mass and spin parameters a basis' prior ranges may name, from
    parameter rows (detector-frame component masses, aligned spins)."""
    m1 = rows[:, _abi.P_MASS_1]
    m2 = rows[:, _abi.P_MASS_2]
    mt = m1 + m2
    return {"mass_1": m1, "mass_2": m2, "total_mass": mt, "mass_ratio": m2 / m1,
            "chirp_mass": (m1 * m2) ** 0.6 / mt ** 0.2, "symmetric_mass_ratio": m1 * m2 / mt ** 2,
            "chi_1": rows[:, _abi.P_CHI_1], "chi_2": rows[:, _abi.P_CHI_2],
            "luminosity_distance": rows[:, _abi.P_LUMINOSITY_DISTANCE],
            "theta_jn": rows[:, _abi.P_THETA_JN]}


class ROQGravitationalWaveTransient:
    """roq.py:14-300.  Weights hold one basis per kind (arrays) or several
    (lists, with ``prior_range_linear`` / ``prior_range_quadratic``: for each
    parameter an array [n_bases, 2]); every row is evaluated with the first
    basis of each kind whose prior ranges hold its parameters
    (roq.py:402-435), rows of each basis pair in one batched launch."""

    @property
    def roq_params(self):
        return self._roq_params

    @roq_params.setter
    def roq_params(self, roq_params):
        self._roq_params = _check_roq_params(roq_params)

    def __init__(self, interferometers, waveform_generator, priors, weights=None,
                 linear_matrix=None, quadratic_matrix=None, roq_params=None,
                 roq_params_check=True, roq_scale_factor=1, distance_marginalization=False,
                 phase_marginalization=False, time_marginalization=False, jitter_time=True,
                 delta_tc=None, distance_marginalization_lookup_table=None,
                 reference_frame="sky", time_reference="geocenter", parameter_conversion=None,
                 precision="fp32", device=0):
        from .frames import sky_frame
        # base.py:170-181: a detector time reference cannot be marginalised over
        self._frame = sky_frame(reference_frame, time_reference)
        self.reference_frame = reference_frame
        self.time_reference = "geocent" if "geocent" in time_reference else time_reference
        if "geocent" not in time_reference and time_marginalization:
            time_marginalization = False
            jitter_time = False
        self.interferometers = InterferometerList(interferometers)
        self.waveform_generator = waveform_generator
        self.priors = priors.copy()
        self.roq_params_check = roq_params_check
        self.roq_scale_factor = roq_scale_factor
        self._roq_params = None
        if isinstance(roq_params, np.ndarray) or roq_params is None:
            self.roq_params = roq_params
        elif isinstance(roq_params, str):
            self.roq_params_file = roq_params
            self.roq_params = np.genfromtxt(roq_params, names=True)
        else:
            raise TypeError("roq_params should be array or str")
        if isinstance(weights, str):
            weights = load_weights(weights, self.interferometers, waveform_generator)
        if weights is None:
            if linear_matrix is None or quadratic_matrix is None:
                raise ValueError("pass weights, or linear_matrix and quadratic_matrix")
            if self.roq_params is not None:
                for ifo in self.interferometers:
                    self.perform_roq_params_check(ifo)
            weights = self._weights_from_bases(self.interferometers, waveform_generator, priors,
                                               linear_matrix, quadratic_matrix, time_reference,
                                               device, self.roq_params, roq_scale_factor)
        self.weights = weights
        self.parameter_conversion = parameter_conversion
        self._setup_bases()
        self.time_marginalization = time_marginalization
        self.phase_marginalization = phase_marginalization
        self.distance_marginalization = distance_marginalization
        self.jitter_time = jitter_time if time_marginalization else False
        self.precision = MODES[precision] if isinstance(precision, str) else int(precision)
        self.device = device
        self._marginalized_parameters = []
        ts = np.asarray(weights["time_samples"], float)
        self.frequency_nodes_linear = self._nodes["linear"][0]
        self.frequency_nodes_quadratic = self._nodes["quadratic"][0]
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
        self._ts = np.ascontiguousarray(ts)
        self._delta_tc_arg = delta_tc
        self._handles = {}
        self._cache = dict(parameters=None, basis_number_linear=None, basis_number_quadratic=None)
        self._prepare_shared(distance_marginalization_lookup_table)
        self._roq = self._handle(0, 0)

    # ------------------------------------------------------------------ bases
    def _setup_bases(self):
        """roq.py:192-217: the number of bases of each kind, their frequency
        nodes and prior ranges, checked as the reference checks them."""
        w = self.weights
        wfa = self.waveform_generator.waveform_arguments
        names = [ifo.name for ifo in self.interferometers]
        self._nodes, self._per_basis = {}, {}
        for kind in ("linear", "quadratic"):
            per_ifo = {n: _as_basis_list(w["%s_%s" % (n, kind)]) for n in names}
            counts = {len(v) for v in per_ifo.values()}
            if len(counts) != 1:
                raise ValueError("the detectors hold different numbers of %s bases" % kind)
            nb = counts.pop()
            setattr(self, "number_of_bases_" + kind, nb)
            self._per_basis[kind] = per_ifo
            key = "frequency_nodes_" + kind
            if nb > 1:
                # roq.py:219-253
                rkey = "prior_range_" + kind
                if rkey not in w:
                    raise AttributeError('For the use of multiple %s ROQ bases, weights should '
                                         'contain "%s".' % (kind, rkey))
                for name, ranges in w[rkey].items():
                    if len(ranges) != nb:
                        raise ValueError('The number of prior ranges for "%s" does not match the '
                                         'number of %s bases' % (name, kind))
                if key not in w:
                    raise AttributeError('For the use of multiple %s ROQ bases, weights should '
                                         'contain "%s".' % (kind, key))
                nodes = _as_basis_list(w[key])
                if len(nodes) != nb:
                    raise ValueError("The number of arrays of frequency nodes does not match the "
                                     "number of %s bases" % kind)
            else:
                # roq.py:282-302
                if key not in w and key not in wfa:
                    raise AttributeError("%s should be contained in weights or waveform "
                                         "arguments." % key)
                nodes = [np.asarray(w[key] if key in w else wfa[key])]
                if nodes[0].ndim == 2:
                    nodes = [nodes[0][0]]
            self._nodes[kind] = [np.ascontiguousarray(np.asarray(x, float)) for x in nodes]
            self._verify_prior_ranges(kind)

    def _verify_prior_ranges(self, basis_type):
        """roq.py:255-280: the union of a kind's prior ranges must hold the
        prior of each parameter they name."""
        key = "prior_range_" + basis_type
        if key not in self.weights:
            return
        for name, ranges in self.weights[key].items():
            ranges = np.asarray(ranges, float)
            if name not in self.priors:
                continue
            prior = self.priors[name]
            if prior.minimum < np.min(ranges[:, 0]):
                raise BilbyROQParamsRangeError(
                    "Prior minimum of %s %s less than ROQ basis bound %s"
                    % (name, prior.minimum, np.min(ranges[:, 0])))
            if prior.maximum > np.max(ranges[:, 1]):
                raise BilbyROQParamsRangeError(
                    "Prior maximum of %s %s greater than ROQ basis bound %s"
                    % (name, prior.maximum, np.max(ranges[:, 1])))

    def perform_roq_params_check(self, ifo=None):
        """roq.py:653-734: the detector's band and duration against roq_params
        (flow, fhigh scaled by roq_scale_factor, seglen divided by it), and for
        a CBC prior its chirp-mass and component-mass bounds against the
        basis' (chirpmassmin, chirpmassmax, compmin, when roq_params holds
        them)."""
        from .prior import CBCPriorDict
        if self.roq_params_check is False:
            return
        p = self.roq_params
        s = float(self.roq_scale_factor)
        names = p.dtype.names

        def field(name):
            return float(p[name]) if name in names else None

        if float(ifo.maximum_frequency) > field("fhigh") * s:
            raise BilbyROQParamsRangeError(
                "Requested maximum frequency %s larger than ROQ basis fhigh %s"
                % (ifo.maximum_frequency, field("fhigh") * s))
        if float(ifo.minimum_frequency) < field("flow") * s:
            raise BilbyROQParamsRangeError(
                "Requested minimum frequency %s lower than ROQ basis flow %s"
                % (ifo.minimum_frequency, field("flow") * s))
        if float(ifo.duration) != field("seglen") / s:
            raise BilbyROQParamsRangeError("Requested duration differs from ROQ basis seglen")
        if not isinstance(self.priors, CBCPriorDict):
            return
        mc_min = field("chirpmassmin")
        mc_max = field("chirpmassmax")
        comp_min = field("compmin")
        if mc_min is not None and self.priors.minimum_chirp_mass is not None \
                and self.priors.minimum_chirp_mass < mc_min / s:
            raise BilbyROQParamsRangeError(
                "Prior minimum chirp mass %s less than ROQ basis bound %s"
                % (self.priors.minimum_chirp_mass, mc_min / s))
        if mc_max is not None and self.priors.maximum_chirp_mass is not None \
                and self.priors.maximum_chirp_mass > mc_max / s:
            raise BilbyROQParamsRangeError(
                "Prior maximum chirp mass %s greater than ROQ basis bound %s"
                % (self.priors.maximum_chirp_mass, mc_max / s))
        if comp_min is not None and self.priors.minimum_component_mass is not None \
                and self.priors.minimum_component_mass < comp_min / s:
            raise BilbyROQParamsRangeError(
                "Prior minimum component mass %s less than ROQ basis bound %s"
                % (self.priors.minimum_component_mass, comp_min / s))

    def _basis_numbers(self, params, n):
        """roq.py:402-435 vectorised: per row the first linear and the first
        quadratic basis whose prior ranges hold the row's parameters
        (parameters a range names but the row lacks are not checked)."""
        out = []
        for kind in ("linear", "quadratic"):
            nb = getattr(self, "number_of_bases_" + kind)
            if nb == 1:
                out.append(np.zeros(n, dtype=np.int64))
                continue
            ok = np.ones((nb, n), dtype=bool)
            for name, ranges in self.weights["prior_range_" + kind].items():
                if name not in params:
                    continue
                ranges = np.asarray(ranges, float)
                v = np.broadcast_to(np.asarray(params[name], float), (n,))[None, :]
                ok &= (ranges[:, 0:1] <= v) & (ranges[:, 1:2] >= v)
            if not np.all(ok.any(axis=0)):
                raise BilbyROQParamsRangeError(
                    "no %s ROQ basis covers the parameters of %d row(s)"
                    % (kind, int(np.sum(~ok.any(axis=0)))))
            out.append(np.argmax(ok, axis=0))
        return out[0], out[1]

    def _update_basis(self, parameters):
        """roq.py:402-435 for one parameter dict: the basis numbers of the
        last evaluated sample (basis_number_linear / _quadratic)."""
        q = dict(parameters)
        if self.parameter_conversion is not None:
            q = self.parameter_conversion(q)
        bl, bq = self._basis_numbers(q, 1)
        self._cache = dict(parameters=q, basis_number_linear=int(bl[0]),
                           basis_number_quadratic=int(bq[0]))
        return int(bl[0]), int(bq[0])

    @property
    def basis_number_linear(self):
        return self._cache["basis_number_linear"] if self.number_of_bases_linear > 1 else 0

    @property
    def basis_number_quadratic(self):
        return self._cache["basis_number_quadratic"] if self.number_of_bases_quadratic > 1 else 0

    @property
    def _multi_basis(self):
        return self.number_of_bases_linear > 1 or self.number_of_bases_quadratic > 1

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
        if self._multi_basis:
            raise ValueError("Format npz not compatible with multiple bases")
        out = {"time_samples": np.asarray(self.weights["time_samples"])}
        for ifo in self.interferometers:
            for kind in ("linear", "quadratic"):
                out["%s_%s" % (ifo.name, kind)] = self._per_basis[kind][ifo.name][0]
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
                            quadratic_matrix, time_reference, device, roq_params=None,
                            roq_scale_factor=1.0):
        """roq.py:736-847 (_set_weights) for one linear and one quadratic
        basis on each detector's masked frequencies, or with roq_params on
        the ROQ frequency grid intersected with them: the time resolution of
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
                                    float(tp.minimum), float(tp.maximum), device=device,
                                    roq_params=roq_params, roq_scale_factor=roq_scale_factor)

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

    def _prepare_shared(self, lookup_table):
        """What every basis pair's GPU likelihood shares: the distance
        lookup table and its spline."""
        self._spline = None
        if self.distance_marginalization:
            table, (tx, ty, c, bbox) = distance_interpolant(
                self.device, self._distance_array, self.distance_prior_array, self._ref_dist,
                self.phase_marginalization, lookup_table)
            self._spline = (tx, ty, c, bbox)
            self._dist_margd_loglikelihood_array = table

    def _handle(self, bl, bq):
        """The GPU likelihood (gwl_roq) of linear basis bl and quadratic
        basis bq, built on first use."""
        key = (int(bl), int(bq))
        if key in self._handles:
            return self._handles[key]
        ifos = self.interferometers
        names = [ifo.name for ifo in ifos]
        ts = self._ts
        lin = np.stack([np.asarray(self._per_basis["linear"][n][key[0]], complex) for n in names])
        quad = np.stack([np.asarray(self._per_basis["quadratic"][n][key[1]], complex)
                         for n in names])
        lin_f = np.ascontiguousarray(lin).view(np.float64).copy()
        quad_f = np.ascontiguousarray(quad).view(np.float64).copy()
        tens = np.ascontiguousarray(np.stack([i.geometry.detector_tensor.ravel() for i in ifos]))
        vert = np.ascontiguousarray(np.stack([i.geometry.vertex for i in ifos]))
        fl = self._nodes["linear"][key[0]]
        fq = self._nodes["quadratic"][key[1]]
        if lin.shape[2] != len(fl) or quad.shape[1] != len(fq):
            raise ValueError("ROQ basis pair %s: weights and frequency nodes differ in length"
                             % (key,))
        wfa = self.waveform_generator.waveform_arguments
        d = _lib.RoqDesc()
        d.n_det = len(ifos)
        d.n_linear = len(fl)
        d.n_quadratic = len(fq)
        d.n_time_samples = len(ts)
        d.start_time = float(ifos.start_time)
        d.time_samples_start = float(ts[0])
        d.time_samples_step = float(ts[1] - ts[0])
        d.time_samples = dptr(ts)
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
            d.delta_tc = float(self._delta_tc_arg)
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
        self._handles[key] = h
        if self.time_marginalization and not _is_uniform(self.priors["geocent_time"]):
            # the prior's own support bounds the grid; its density weighs each time
            tp = self.priors["geocent_time"]
            t0, dt, dens = density_table(tp, float(tp.minimum), float(tp.maximum), self._delta_tc)
            check(LIB.gwl_roq_set_time_prior(h, len(dens), t0, dt, dptr(dens)),
                  "gwl_roq_set_time_prior")
        self.hbm_bytes = getattr(self, "hbm_bytes", 0) + int(nbytes.value)
        if self._spline is not None:
            tx, ty, c, bbox = self._spline
            check(LIB.gwl_roq_set_distance_spline(h, len(tx) - 4, len(ty) - 4, dptr(tx),
                                                  dptr(ty), dptr(c), dptr(bbox)),
                  "gwl_roq_set_distance_spline")
        return h

    def __del__(self):
        try:
            for h in getattr(self, "_handles", {}).values():
                LIB.gwl_roq_destroy(h)
            self._handles = {}
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
        rows_h = self.rows_from_parameters(q)[:1]
        bl, bq = self._update_basis(q) if self._multi_basis else (0, 0)
        rows = torch.from_numpy(rows_h).to("cuda:%d" % self.device)
        dh = torch.empty((1, max(1, n_times), 2), dtype=torch.float64, device=rows.device)
        hh = torch.empty(1, dtype=torch.float64, device=rows.device)
        st = torch.cuda.current_stream(rows.device).cuda_stream
        check(LIB.gwl_roq_inner_products(self._handle(bl, bq), ctypes.c_void_p(rows.data_ptr()), 1,
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
        rows = None
        if on_device:
            rows_t = parameters.to(torch.float64).contiguous()
        else:
            rows = (self.rows_from_parameters(parameters) if isinstance(parameters, dict)
                    else np.ascontiguousarray(parameters, dtype=np.float64))
            rows_t = torch.from_numpy(rows).to("cuda:%d" % self.device)
        n = rows_t.shape[0]
        out = torch.empty(n, dtype=torch.float64, device=rows_t.device)
        st = torch.cuda.current_stream(rows_t.device).cuda_stream if stream is None else stream
        if not self._multi_basis:
            check(LIB.gwl_roq_eval(self._roq, ctypes.c_void_p(rows_t.data_ptr()), n, _lib.NPARAM,
                                   mode, ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(st)),
                  "gwl_roq_eval")
            return out if on_device else out.cpu().numpy()
        # several bases: each row's basis pair from its parameters (a dict as
        # given, after parameter_conversion; rows through their masses and
        # spins), then one launch per pair over its rows
        if isinstance(parameters, dict):
            q = dict(parameters)
            if self.parameter_conversion is not None:
                q = self.parameter_conversion(q)
        else:
            q = _params_from_rows(rows if rows is not None else rows_t.cpu().numpy())
        bl, bq = self._basis_numbers(q, n)
        pair = bl * self.number_of_bases_quadratic + bq
        for code in np.unique(pair):
            idx = np.nonzero(pair == code)[0]
            sel = torch.from_numpy(idx).to(rows_t.device)
            sub = rows_t.index_select(0, sel).contiguous()
            res = torch.empty(len(idx), dtype=torch.float64, device=rows_t.device)
            h = self._handle(int(code) // self.number_of_bases_quadratic,
                             int(code) % self.number_of_bases_quadratic)
            check(LIB.gwl_roq_eval(h, ctypes.c_void_p(sub.data_ptr()), len(idx), _lib.NPARAM, mode,
                                   ctypes.c_void_p(res.data_ptr()), ctypes.c_void_p(st)),
                  "gwl_roq_eval")
            out.index_copy_(0, sel, res)
        return out if on_device else out.cpu().numpy()

    def log_likelihood_ratio(self, parameters):
        """roq.py:437-439: the basis of the sample, then its lnL."""
        if self._multi_basis:
            self._update_basis(parameters)
        return float(self.log_likelihood_ratio_batch(copy.copy(parameters))[0])
