"""Calibration response models and the response-curve lookup that
calibration marginalisation contracts with (bilby/gw/detector/calibration.py,
bilby/gw/prior.py:1339-1399).  Host-side setup in float64: the curves are
built once per event; the per-row contraction runs on the GPU
(csrc/gwl_calib.hip).

bilby stores lookups as HDF5 files; h5py is not part of this stack, so a
lookup file here is an ``.npz`` archive with ``curves`` [n_curves, n_freq]
(complex), ``frequencies`` and one array per calibration parameter.
"""
import os

import numpy as np

from .prior import DeltaFunction, Gaussian, PriorDict


class Recalibrate:
    """calibration.py Recalibrate: the identity response."""

    name = "none"

    def __init__(self, prefix="recalib_"):
        self.prefix = prefix
        self.params = {}

    def set_calibration_parameters(self, **params):
        self.params.update({k[len(self.prefix):]: v for k, v in params.items()
                            if k.startswith(self.prefix)})

    def get_calibration_factor(self, frequency_array, **params):
        return np.ones_like(np.asarray(frequency_array, float), dtype=complex)


class CubicSpline(Recalibrate):
    """calibration.py:257-383: amplitude and phase corrections as cubic
    splines in log10 f through ``n_points`` nodes equally spaced in log10 f
    between ``minimum_frequency`` and ``maximum_frequency``, with the
    not-a-knot end conditions of LIGO-T2300140 Eq. 9; the response is
    (1 + dA) (2 + i dphi) / (2 - i dphi)."""

    name = "cubic_spline"

    def __init__(self, prefix, minimum_frequency, maximum_frequency, n_points):
        super().__init__(prefix=prefix)
        if n_points < 4:
            raise ValueError("Cubic spline calibration requires at least 4 spline nodes.")
        self.n_points = n_points
        self.minimum_frequency = minimum_frequency
        self.maximum_frequency = maximum_frequency
        self.log_spline_points = np.linspace(np.log10(minimum_frequency),
                                             np.log10(maximum_frequency), n_points)
        self.delta_log_spline_points = self.log_spline_points[1] - self.log_spline_points[0]
        self.nodes_to_spline_coefficients = self._spline_matrix(n_points)

    @staticmethod
    def _spline_matrix(n):
        """Second-derivative coefficients from node values: the interior rows
        are the uniform cubic-spline continuity equations, the first and last
        rows ask for a continuous third derivative at the second and
        second-to-last nodes."""
        lhs = np.zeros((n, n))
        rhs = np.zeros((n, n))
        lhs[0, :3] = (-1.0, 2.0, -1.0)
        lhs[-1, -3:] = (-1.0, 2.0, -1.0)
        for i in range(1, n - 1):
            lhs[i, i - 1:i + 2] = (1 / 6, 2 / 3, 1 / 6)
            rhs[i, i - 1:i + 2] = (1.0, -2.0, 1.0)
        return np.linalg.solve(lhs, rhs)

    def _spline(self, kind, a, b, c, d, lower):
        values = np.asarray([self.params["%s_%d" % (kind, i)] for i in range(self.n_points)],
                            dtype=float)
        second = self.nodes_to_spline_coefficients.dot(values)
        upper = lower + 1
        return a * values[lower] + b * values[upper] + c * second[lower] + d * second[upper]

    def get_calibration_factor(self, frequency_array, **params):
        self.set_calibration_parameters(**params)
        with np.errstate(divide="ignore"):
            x = np.nan_to_num(np.log10(np.asarray(frequency_array, float))
                              - self.log_spline_points[0], neginf=0.0)
        x = x / self.delta_log_spline_points
        lower = np.clip(x.astype(int), 0, self.n_points - 2)
        b = x - lower
        a = 1 - b
        c = (a ** 3 - a) / 6
        d = (b ** 3 - b) / 6
        amplitude = self._spline("amplitude", a, b, c, d, lower)
        phase = self._spline("phase", a, b, c, d, lower)
        return np.nan_to_num((1 + amplitude) * (2 + 1j * phase) / (2 - 1j * phase))


class Precomputed(Recalibrate):
    """calibration.py:386-458: a fixed set of response curves."""

    name = "precomputed"

    def __init__(self, label, curves, frequency_array, parameters=None):
        super().__init__(prefix="recalib_index_%s" % label)
        self.label = label
        self.curves = np.asarray(curves)
        self.frequency_array = np.asarray(frequency_array)
        self.parameters = dict(parameters or {})

    @classmethod
    def constant_uncertainty_spline(cls, amplitude_sigma, phase_sigma, frequency_array, n_nodes,
                                    label, n_curves, rng=None):
        priors = constant_uncertainty_spline_priors(amplitude_sigma, phase_sigma,
                                                    frequency_array[0], frequency_array[-1],
                                                    n_nodes, label)
        parameters = priors.sample(n_curves, rng=rng)
        curves = curves_from_spline_and_prior(parameters, label, n_nodes, frequency_array,
                                              n_curves)
        return cls(label, curves, frequency_array, parameters)


def constant_uncertainty_spline_priors(amplitude_sigma, phase_sigma, minimum_frequency,
                                       maximum_frequency, n_nodes, label, boundary="reflective"):
    """prior.py:1339-1399: Gaussian(0, sigma) amplitude and phase at each node
    and the node frequencies as fixed values."""
    nodes = np.logspace(np.log10(minimum_frequency), np.log10(maximum_frequency), n_nodes)
    prior = PriorDict()
    for i in range(n_nodes):
        name = "recalib_%s_amplitude_%d" % (label, i)
        prior[name] = Gaussian(0.0, amplitude_sigma, name=name, boundary=boundary)
    for i in range(n_nodes):
        name = "recalib_%s_phase_%d" % (label, i)
        prior[name] = Gaussian(0.0, phase_sigma, name=name, boundary=boundary)
    for i in range(n_nodes):
        name = "recalib_%s_frequency_%d" % (label, i)
        prior[name] = DeltaFunction(nodes[i], name=name)
    return prior


def curves_from_spline_and_prior(parameters, label, n_points, frequency_array, n_curves):
    """calibration.py:581-594: one response curve per parameter draw."""
    spline = CubicSpline("recalib_%s_" % label, frequency_array[0], frequency_array[-1], n_points)
    curves = []
    for i in range(n_curves):
        draw = {k: float(np.asarray(v)[i]) for k, v in parameters.items()}
        curves.append(spline.get_calibration_factor(frequency_array, **draw))
    return np.array(curves)


def _read_lookup(filename, frequencies, n_curves, starting_index):
    with np.load(filename, allow_pickle=False) as f:
        if not np.allclose(f["frequencies"], frequencies):
            raise ValueError("%s: lookup frequencies do not match the data" % filename)
        stop = starting_index + n_curves
        if f["curves"].shape[0] < stop:
            raise ValueError("%s holds %d curves, %d requested from index %d"
                             % (filename, f["curves"].shape[0], n_curves, starting_index))
        curves = np.array(f["curves"][starting_index:stop])
        params = {k: np.array(f[k][starting_index:stop]) for k in f.files
                  if k not in ("curves", "frequencies")}
    return curves, params


def write_calibration_file(filename, frequencies, curves, parameters):
    np.savez(filename, curves=np.asarray(curves), frequencies=np.asarray(frequencies),
             **{k: np.asarray(v) for k, v in parameters.items()})


def build_calibration_lookup(interferometers, lookup_files=None, priors=None,
                             number_of_response_curves=1000, starting_index=0, rng=None):
    """calibration.py:503-554: per detector, curves on the masked frequencies
    from (in order) an existing lookup file, the detector's Precomputed
    model, or draws from the ``recalib_<ifo>_*`` priors through its
    CubicSpline model (then written to the lookup file)."""
    if lookup_files is None and priors is None:
        raise ValueError("One of calibration_lookup_table or priors must be specified for "
                         "building calibration marginalization lookup table.")
    lookup_files = dict(lookup_files or {})
    draws, parameters = {}, {}
    for ifo in interferometers:
        name = ifo.name
        frequencies = ifo.frequency_array[ifo.frequency_mask]
        filename = lookup_files.get(name, "%s_calibration_file.npz" % name)
        model = getattr(ifo, "calibration_model", None)
        if os.path.exists(filename):
            draws[name], parameters[name] = _read_lookup(filename, frequencies,
                                                         number_of_response_curves, starting_index)
        elif isinstance(model, Precomputed):
            idx = np.arange(number_of_response_curves) + starting_index
            draws[name] = model.curves[idx]
            parameters[name] = {k: np.asarray(v)[idx] for k, v in model.parameters.items()}
            parameters[name][model.prefix] = idx
        else:
            if priors is None:
                raise ValueError("Priors must be passed to generate calibration response curves "
                                 "for cubic spline.")
            if not isinstance(model, CubicSpline):
                raise ValueError("%s has no cubic-spline calibration model to draw curves from"
                                 % name)
            sub = PriorDict({k: v for k, v in priors.items()
                             if "recalib" in k and name in k})
            params = sub.sample(number_of_response_curves, rng=rng)
            draws[name] = curves_from_spline_and_prior(params, name, model.n_points, frequencies,
                                                       number_of_response_curves)
            parameters[name] = params
            write_calibration_file(filename, frequencies, draws[name], params)
        ifo.calibration_model = Recalibrate()
    return draws, parameters
