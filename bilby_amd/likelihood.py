"""Drop-in GravitationalWaveTransient on the MI355X.

Same constructor, same ``log_likelihood_ratio(parameters: dict) -> float``,
``log_likelihood``, ``noise_log_likelihood``, ``marginalized_parameters`` and
prior side effects as bilby.gw.likelihood.GravitationalWaveTransient
(bilby/gw/likelihood/base.py:150-229, 392-446), plus
``log_likelihood_ratio_batch`` over N parameter rows.  Every evaluation runs
in libgwl's HIP kernels; there is no host fallback.
"""
import copy
import ctypes
import threading
import os

import numpy as np
import torch

from . import _lib
from . import abi as _abi
from ._lib import LIB, check, dptr, iptr
from .detector import InterferometerList
from .prior import Cosmological, Uniform
from .waveform_generator import taylorf2_flags

_CTX = {}
_INTERPOLANTS = {}  # distance interpolants by (grid and prior digest, phase flag)


def get_context(device=0):
    """One libgwl context per HIP device (per process)."""
    if device not in _CTX:
        if not torch.cuda.is_available():
            raise RuntimeError("no HIP device visible: the MI355X likelihood needs a GPU")
        h = ctypes.c_void_p()
        check(LIB.gwl_ctx_create(int(device), ctypes.byref(h)), "gwl_ctx_create")
        _CTX[device] = h
    return _CTX[device]


def _batch_rows(parameters):
    """Rows in a batch given as a dict of arrays, an [N, 16] array or tensor."""
    if isinstance(parameters, dict):
        for v in parameters.values():
            if np.ndim(v) > 0:
                return len(v)
        return 1
    shape = getattr(parameters, "shape", None)
    if shape is None:
        shape = np.shape(parameters)
    return int(shape[0]) if len(shape) > 0 else 1


def _precision_code(precision):
    if precision in ("fp32", "float32", _lib.FP32):
        return _lib.FP32
    if precision in ("fp64", "float64", _lib.FP64):
        return _lib.FP64
    raise ValueError("precision must be 'fp32' or 'fp64'")


def d_inner_h_ref_axis(phase_marginalization, n=800):
    """base.py:902-914."""
    if phase_marginalization:
        return np.logspace(-5, 10, n)
    n_negative = n // 2
    n_positive = n - n_negative
    return np.hstack((-np.logspace(3, -3, n_negative), np.logspace(-3, 10, n_positive)))


def optimal_snr_squared_ref_axis(n=400):
    """base.py:897-900."""
    return np.logspace(-5, 10, n)


def ln_i0_host(value):
    """gw/utils.py:1006-1022 on the host: ln I0(|x|) = ln(i0e(|x|)) + |x|."""
    from scipy.special import i0e
    value = np.abs(np.asarray(value, dtype=float))
    return np.log(i0e(value)) + value


def grid_time_posterior(likelihood, d_inner_h, h_inner_h, parameters, times):
    """The time posterior on a likelihood's own time grid (roq.py:1255-1281,
    multiband.py:828-848): the time-conditioned lnL (distance / phase
    marginalised as configured) times the time prior, normalised.  NaN
    inner products (times the interpolation cannot reach) count as
    lnL = -inf."""
    if likelihood.distance_marginalization:
        log_like = likelihood._distance_marginalized_values(d_inner_h, h_inner_h, parameters)
    elif likelihood.phase_marginalization:
        log_like = ln_i0_host(np.abs(d_inner_h)) - h_inner_h / 2
    else:
        log_like = d_inner_h.real - h_inner_h / 2
    log_like = np.where(np.isnan(log_like), -np.inf, log_like)
    if not np.any(np.isfinite(log_like)):
        raise ValueError("the time-conditioned likelihood is not finite anywhere on the grid "
                         "(did the waveform fail at these parameters?)")
    post = np.exp(log_like - np.max(log_like)) * likelihood.priors["geocent_time"].prob(times)
    total = np.sum(post)
    if not total > 0:
        raise ValueError("the time posterior vanishes on the grid")
    return post / total


def interped_sample(xx, yy, rng):
    """One draw of bilby's Interped(xx, yy).sample()
    (core/prior/interpolated.py:12-175): the density re-gridded linearly to
    linspace(min, max, len(xx)), normalised by the trapezoid rule, its
    cumulative trapezoid (last value set to 1) inverted by linear
    interpolation at a uniform variate."""
    from scipy.integrate import cumulative_trapezoid, trapezoid
    xx = np.asarray(xx, dtype=float)
    yy = np.asarray(yy, dtype=float)
    order = np.argsort(xx, kind="stable")
    xx, yy = xx[order], yy[order]
    grid = np.linspace(xx[0], xx[-1], len(xx))
    density = np.interp(grid, xx, yy)
    density = density / trapezoid(density, grid)
    cdf = cumulative_trapezoid(density, grid, initial=0)
    cdf[-1] = 1
    return float(np.interp(rng.uniform(0, 1), cdf, grid))


def _test_cached_lookup_table(loaded, distance_array, prior_array, ref_dist, phase):
    """base.py:979-992."""
    pairs = dict(distance_array=distance_array, prior_array=prior_array,
                 reference_distance=ref_dist, phase_marginalization=phase)
    for key, val in pairs.items():
        if key not in loaded:
            return False
        if not np.allclose(np.atleast_1d(loaded[key]), np.atleast_1d(val), rtol=1e-15):
            return False
    return True


def create_lookup_table(device, distance_array, prior_array, ref_dist, phase):
    """base.py:994-1018 on the GPU (gwl_distance_table)."""
    ctx = get_context(device)
    x = d_inner_h_ref_axis(phase)
    y = optimal_snr_squared_ref_axis()
    out = np.zeros((len(y), len(x)))
    dist = np.ascontiguousarray(distance_array, dtype=np.float64)
    prior = np.ascontiguousarray(prior_array, dtype=np.float64)
    check(LIB.gwl_distance_table(ctx, dptr(dist), dptr(prior), len(dist), float(ref_dist),
                                 int(phase), dptr(x), len(x), dptr(y), len(y), dptr(out)),
          "gwl_distance_table")
    return out


def distance_interpolant(device, distance_array, prior_array, ref_dist, phase, lookup_table=None):
    """Lookup table (loaded from / cached to an .npz file when a name is
    given, base.py:936-977) and the knots, coefficients and bounding box of
    its bicubic interpolant (base.py:929-934)."""
    import hashlib
    key = None
    if lookup_table is None:
        h = hashlib.sha1()
        h.update(np.ascontiguousarray(distance_array, dtype=np.float64).tobytes())
        h.update(np.ascontiguousarray(prior_array, dtype=np.float64).tobytes())
        h.update(np.float64(ref_dist).tobytes())
        key = (h.hexdigest(), bool(phase))
        if key in _INTERPOLANTS:
            return _INTERPOLANTS[key]
    table = None
    filename = None
    if isinstance(lookup_table, str):
        filename = lookup_table if lookup_table.endswith(".npz") else lookup_table + ".npz"
        if os.path.exists(filename):
            loaded = dict(np.load(filename))
            if _test_cached_lookup_table(loaded, distance_array, prior_array, ref_dist, phase):
                table = loaded["lookup_table"]
    elif isinstance(lookup_table, dict):
        if _test_cached_lookup_table(lookup_table, distance_array, prior_array, ref_dist, phase):
            table = lookup_table["lookup_table"]
    if table is None:
        table = create_lookup_table(device, distance_array, prior_array, ref_dist, phase)
        if filename is not None:
            np.savez(filename, distance_array=distance_array, prior_array=prior_array,
                     lookup_table=table, reference_distance=ref_dist,
                     phase_marginalization=phase)
    x = d_inner_h_ref_axis(phase, table.shape[1])
    y = optimal_snr_squared_ref_axis(table.shape[0])
    z = np.ascontiguousarray(table.T)
    tx = np.zeros(len(x) + 4)
    ty = np.zeros(len(y) + 4)
    c = np.zeros_like(z)
    check(LIB.gwl_spline2d_coefficients(len(x), len(y), dptr(x), dptr(y), dptr(z), dptr(tx),
                                        dptr(ty), dptr(c)), "gwl_spline2d_coefficients")
    bbox = np.array([x.min(), x.max(), y.min(), y.max()])
    if key is not None:
        _INTERPOLANTS[key] = (table, (tx, ty, c, bbox))
    return table, (tx, ty, c, bbox)


def _is_uniform(prior):
    return isinstance(prior, Uniform) or type(prior).__name__ == "Uniform"


def time_prior_table(prior, start_time, duration, delta_tc, points_per_step=64):
    """The geocent_time prior as the kernels weigh it (base.py:794-820:
    ``prior.prob(times) * delta_tc`` at the jittered times inside
    [prior.minimum, prior.maximum]).  A Uniform prior: (minimum, maximum,
    None), its constant weight applied on the device.  Any other prior
    (bilby's Prior API: ``minimum``, ``maximum``, ``prob``): the support
    clipped to the segment and to where the density is non-zero, and the
    density on points_per_step points per delta_tc with nodes on both ends
    (so no interpolation crosses the support's edges), interpolated
    linearly on the device: (lo, hi, (t0, dt, density))."""
    if _is_uniform(prior):
        return float(prior.minimum), float(prior.maximum), None
    lo = max(float(prior.minimum), float(start_time))
    hi = min(float(prior.maximum), float(start_time) + float(duration))
    if not hi > lo:
        raise ValueError("the geocent_time prior has no support inside the segment")
    probe = np.linspace(lo, hi, max(3, int(np.ceil((hi - lo) / delta_tc)) * 4 + 1))
    nz = np.nonzero(np.asarray(prior.prob(probe), dtype=float) > 0)[0]
    if len(nz) == 0:
        raise ValueError("the geocent_time prior density is zero on the whole segment")
    # an end where the density vanishes (a tail, an end outside the segment)
    # moves in to the last probe point before its non-zero part; an end where
    # it does not (a hard edge) stays a table node
    if nz[0] > 0:
        lo = float(probe[nz[0] - 1])
    if nz[-1] < len(probe) - 1:
        hi = float(probe[nz[-1] + 1])
    return lo, hi, density_table(prior, lo, hi, delta_tc, points_per_step)


def density_table(prior, lo, hi, delta_tc, points_per_step=64):
    """The prior density on nodes lo + k dt, k < n, for the device's linear
    interpolation: dt the power of two at or below delta_tc /
    points_per_step, so that every node is an exact float64 (GPS times
    carry 2^-22 s steps; a node rounded to them would put each density
    value up to 1.2e-7 s off its time); the last node at or after hi, with
    the density there taken at hi (the times the kernels weigh stay inside
    [lo, hi], so no interpolation reaches past a hard upper edge).  Returns
    (lo, dt, density)."""
    dt = 2.0 ** np.floor(np.log2(delta_tc / points_per_step))
    n = int(np.ceil((hi - lo) / dt)) + 1
    grid = lo + dt * np.arange(n)
    dens = np.asarray(prior.prob(np.minimum(grid, hi)), dtype=float)
    return float(lo), float(dt), np.ascontiguousarray(np.nan_to_num(dens), dtype=np.float64)


class GravitationalWaveTransient:
    """bilby/gw/likelihood/base.py:150-229 constructor semantics.

    Extra keyword arguments (not in bilby): ``precision`` ('fp32' bulk
    arithmetic with float64 phase arguments and accumulation, or 'fp64'),
    ``device`` (HIP device index).
    """

    def __init__(self, interferometers, waveform_generator, time_marginalization=False,
                 distance_marginalization=False, phase_marginalization=False,
                 calibration_marginalization=False, priors=None,
                 distance_marginalization_lookup_table=None, calibration_lookup_table=None,
                 number_of_response_curves=1000, starting_index=0, jitter_time=True,
                 reference_frame="sky", time_reference="geocenter", precision="fp32", device=0):
        if calibration_marginalization and time_marginalization and precision == "bf16x3":
            raise NotImplementedError("calibration together with time marginalisation is "
                                      "evaluated in float32 or float64")
        from .frames import sky_frame
        # base.py:170-181: a detector time reference cannot be marginalised over
        self._frame = sky_frame(reference_frame, time_reference)
        self.reference_frame = reference_frame
        self.time_reference = "geocent" if "geocent" in time_reference else time_reference
        if "geocent" not in time_reference and time_marginalization:
            time_marginalization = False
            jitter_time = False
        self.waveform_generator = waveform_generator
        if calibration_marginalization and \
                waveform_generator.approximant == _lib.HOST_POLARIZATIONS:
            raise NotImplementedError("calibration marginalisation needs a source model evaluated "
                                      "on the device (sine-Gaussian or TaylorF2)")
        self.interferometers = InterferometerList(interferometers)
        self.time_marginalization = time_marginalization
        self.distance_marginalization = distance_marginalization
        self.phase_marginalization = phase_marginalization
        self.calibration_marginalization = calibration_marginalization
        self._calib = None
        if priors is not None:
            self.priors = priors.copy()
        elif time_marginalization or distance_marginalization or phase_marginalization:
            raise ValueError("You can't use a marginalized likelihood without specifying a priors")
        else:
            self.priors = None
        # calibration marginalisation also takes "bf16x3" (split-bfloat16 MFMA contraction)
        self.calibration_mode = 2 if precision == "bf16x3" else None
        self.precision = _precision_code("fp32" if precision == "bf16x3" else precision)
        if self.calibration_mode is None:
            self.calibration_mode = self.precision
        self.device = device
        self._marginalized_parameters = []
        self._noise_log_likelihood_value = None
        self.jitter_time = jitter_time
        # base.py:239-258
        for attr in ("duration", "sampling_frequency", "start_time"):
            setattr(self.waveform_generator, attr, getattr(self.interferometers, attr))
        self._delta_tc = 2 / self.waveform_generator.sampling_frequency
        if self.time_marginalization:
            self._check_marginalized_prior_is_set("geocent_time")
            priors["geocent_time"] = float(self.interferometers.start_time)
            if self.jitter_time:
                priors["time_jitter"] = Uniform(minimum=-self._delta_tc / 2,
                                                maximum=self._delta_tc / 2,
                                                boundary="periodic", name="time_jitter")
            self._marginalized_parameters.append("geocent_time")
        else:
            self.jitter_time = False
        if self.phase_marginalization:
            self._check_marginalized_prior_is_set("phase")
            priors["phase"] = float(0)
            self._marginalized_parameters.append("phase")
        self._ref_dist = 1.0
        if self.distance_marginalization:
            self._check_marginalized_prior_is_set("luminosity_distance")
            prior = self.priors["luminosity_distance"]
            self._distance_array = np.linspace(prior.minimum, prior.maximum, int(1e4))
            self.distance_prior_array = np.array([prior.prob(d) for d in self._distance_array])
            self._ref_dist = float(prior.rescale(0.5))
            for key in ("redshift", "comoving_distance"):
                if key in priors:
                    del priors[key]
            priors["luminosity_distance"] = float(self._ref_dist)
            self._marginalized_parameters.append("luminosity_distance")
        self._event = None
        self._build_event()
        if self.distance_marginalization:
            self._setup_distance_marginalization(distance_marginalization_lookup_table)
        if self.calibration_marginalization:
            self.number_of_response_curves = number_of_response_curves
            self.starting_index = starting_index
            self._setup_calibration_marginalization(calibration_lookup_table, priors)
            self._marginalized_parameters.append("recalib_index")
        self._one = None
        self._stage = None                 # pinned host / device row buffers (_staged_eval)
        self._stage_lock = threading.Lock()

    # ------------------------------------------------------------------ setup
    def _check_marginalized_prior_is_set(self, key):
        """base.py:356-390: fixed priors are rejected; a missing time prior is
        Uniform over the segment; a missing distance prior comes from a
        Cosmological prior on redshift or comoving distance (which is
        removed); a missing phase prior is the BBH default, Uniform over
        [0, 2 pi] with periodic boundary."""
        if key in self.priors and getattr(self.priors[key], "is_fixed", False):
            raise ValueError("Cannot use marginalized likelihood for {}: prior is fixed".format(key))
        if key not in self.priors or isinstance(self.priors[key], (int, float)):
            if key == "geocent_time":
                self.priors[key] = Uniform(self.interferometers.start_time,
                                           self.interferometers.start_time
                                           + self.interferometers.duration)
            elif key == "luminosity_distance":
                for other in ("redshift", "comoving_distance"):
                    if other in self.priors:
                        if not isinstance(self.priors[other], Cosmological):
                            raise TypeError(
                                "To marginalize over {}, the prior must be specified as a "
                                "subclass of bilby.gw.prior.Cosmological.".format(other))
                        self.priors["luminosity_distance"] = \
                            self.priors[other].get_corresponding_prior("luminosity_distance")
                        del self.priors[other]
                        break
                else:
                    raise ValueError("a prior for luminosity_distance (or a Cosmological prior "
                                     "on redshift or comoving_distance) is required")
            elif key == "phase":
                self.priors[key] = Uniform(0.0, 2 * np.pi, name="phase", boundary="periodic")
            else:
                raise ValueError("a prior for {} is required".format(key))

    def _flags(self):
        f = 0
        if self.time_marginalization:
            f |= _lib.TIME_MARG
            if self.jitter_time:
                f |= _lib.JITTER_TIME
        if self.distance_marginalization:
            f |= _lib.DISTANCE_MARG
        if self.phase_marginalization:
            f |= _lib.PHASE_MARG
        return f | taylorf2_flags(self.waveform_generator.waveform_arguments)

    def _build_event(self):
        self._event, self.hbm_bytes = self._make_event()

    def _make_event(self, n_freq=None, flags=None, time_prior=None):
        """The device event of this likelihood's data (gwl_event_create).
        ``n_freq``: pad the data with zeros and the PSD with +inf (zero
        weight) to that many frequency bins at the same spacing (sampling
        frequency 2 (n_freq - 1) / T); ``flags`` / ``time_prior`` override
        the likelihood's."""
        ifos = self.interferometers
        ctx = get_context(self.device)
        nd = len(ifos)
        freq = ifos.frequency_array
        nf0 = len(freq)
        nf = nf0 if n_freq is None else int(n_freq)
        strain = np.zeros((nd, nf), dtype=np.complex128)
        strain[:, :nf0] = np.stack([ifo.frequency_domain_strain for ifo in ifos])
        strain_f = np.ascontiguousarray(strain.view(np.float64).reshape(nd, nf, 2))
        psd = np.full((nd, nf), np.inf)
        psd[:, :nf0] = np.stack([ifo.power_spectral_density_array for ifo in ifos])
        psd = np.ascontiguousarray(psd, dtype=np.float64)
        lo = np.array([ifo.mask_range[0] for ifo in ifos], np.int32)
        hi = np.array([ifo.mask_range[1] for ifo in ifos], np.int32)
        tens = np.ascontiguousarray(np.stack([ifo.geometry.detector_tensor.ravel() for ifo in ifos]),
                                    dtype=np.float64)
        vert = np.ascontiguousarray(np.stack([ifo.geometry.vertex for ifo in ifos]), dtype=np.float64)
        wfa = self.waveform_generator.waveform_arguments
        desc = _lib.EventDesc()
        desc.n_det = nd
        desc.n_freq = nf
        desc.duration = float(ifos.duration)
        desc.sampling_frequency = (float(ifos.sampling_frequency) if n_freq is None
                                   else 2.0 * (nf - 1) / float(ifos.duration))
        desc.start_time = float(ifos.start_time)
        desc.mask_lo = iptr(lo)
        desc.mask_hi = iptr(hi)
        desc.strain = dptr(strain_f)
        desc.psd = dptr(psd)
        desc.detector_tensor = dptr(tens)
        desc.vertex = dptr(vert)
        desc.approximant = self.waveform_generator.approximant
        desc.wf_minimum_frequency = float(wfa.get("minimum_frequency", 20.0)) \
            if desc.approximant == _lib.TAYLORF2 else 0.0
        desc.wf_maximum_frequency = float(wfa.get("maximum_frequency", freq[-1]))
        desc.reference_frequency = float(wfa.get("reference_frequency", 50.0))
        desc.flags = self._flags() if flags is None else int(flags)
        table = None
        if desc.flags & _lib.TIME_MARG:
            if time_prior is None:
                lo, hi, table = time_prior_table(self.priors["geocent_time"], ifos.start_time,
                                                 ifos.duration, 2.0 / ifos.sampling_frequency)
                time_prior = (lo, hi)
            desc.time_prior_minimum = float(time_prior[0])
            desc.time_prior_maximum = float(time_prior[1])
        desc.reference_distance = float(self._ref_dist)
        ev = ctypes.c_void_p()
        nbytes = ctypes.c_int64()
        check(LIB.gwl_event_create(ctx, ctypes.byref(desc), ctypes.byref(ev), ctypes.byref(nbytes)),
              "gwl_event_create")
        if table is not None:
            t0, dt, dens = table
            check(LIB.gwl_event_set_time_prior(ev, len(dens), t0, dt, dptr(dens)),
                  "gwl_event_set_time_prior")
        return ev, int(nbytes.value)

    def _setup_distance_marginalization(self, lookup_table=None):
        """base.py:916-992: lookup table (cached to an .npz file when a file
        name is given, validated like _test_cached_lookup_table) and its
        bicubic interpolant, attached to the device event."""
        table, spline = distance_interpolant(self.device, self._distance_array,
                                             self.distance_prior_array, self._ref_dist,
                                             self.phase_marginalization, lookup_table)
        self._dist_margd_loglikelihood_array = table
        self._spline = spline
        tx, ty, c, bbox = spline
        check(LIB.gwl_event_set_distance_spline(self._event, len(tx) - 4, len(ty) - 4, dptr(tx),
                                                dptr(ty), dptr(c), dptr(bbox)),
              "gwl_event_set_distance_spline")

    def _setup_calibration_marginalization(self, calibration_lookup_table, priors=None):
        """base.py:1037-1051: response curves per detector on its masked
        frequencies; the sampled calibration parameters become fixed at 0 in
        the caller's priors; the curves go to the device (gwl_calib_create)."""
        from .calibration import build_calibration_lookup
        from .prior import DeltaFunction
        self.calibration_draws, self.calibration_parameter_draws = build_calibration_lookup(
            self.interferometers, lookup_files=calibration_lookup_table, priors=priors,
            number_of_response_curves=self.number_of_response_curves,
            starting_index=self.starting_index)
        if priors is not None:
            for name, params in self.calibration_parameter_draws.items():
                for key in set(params).intersection(priors.keys()):
                    priors[key] = DeltaFunction(0.0)
        self._attach_calibration_curves(self.calibration_draws)

    def _attach_calibration_curves(self, draws):
        """The response curves (detector name -> complex [curves, masked
        bins]) of this likelihood's detectors to the device event."""
        self.calibration_draws = {ifo.name: np.asarray(draws[ifo.name], dtype=complex)
                                  for ifo in self.interferometers}
        self.calibration_abs_draws = {k: np.abs(v) ** 2 for k, v in self.calibration_draws.items()}
        curves = np.concatenate([self.calibration_draws[ifo.name] for ifo in self.interferometers],
                                axis=1)
        self.number_of_response_curves = int(curves.shape[0])
        blob = np.ascontiguousarray(curves.view(np.float64))
        cal = ctypes.c_void_p()
        check(LIB.gwl_calib_create(self._event, int(curves.shape[0]), dptr(blob), ctypes.byref(cal)),
              "gwl_calib_create")
        self._calib = cal

    def get_calibration_log_likelihoods(self, parameters):
        """base.py:822-858: ln L of every response curve for one parameter
        dict (float64 contraction; with time marginalisation each curve's
        time-marginalised value, transforms in the likelihood's calibration
        precision)."""
        rows = self.rows_from_parameters(dict(parameters))
        rows_t = torch.from_numpy(np.ascontiguousarray(rows[:1])).to("cuda:%d" % self.device)
        out = torch.empty(1, dtype=torch.float64, device=rows_t.device)
        per = torch.empty((1, self.number_of_response_curves), dtype=torch.float64,
                          device=rows_t.device)
        st = torch.cuda.current_stream(rows_t.device).cuda_stream
        mode = self.calibration_mode if self.time_marginalization else _lib.FP64
        check(LIB.gwl_calib_eval(self._calib, ctypes.c_void_p(rows_t.data_ptr()), 1, _lib.NPARAM,
                                 mode, ctypes.c_void_p(out.data_ptr()),
                                 ctypes.c_void_p(per.data_ptr()), ctypes.c_void_p(st)),
              "gwl_calib_eval")
        return per.cpu().numpy()[0]

    def generate_calibration_sample_from_marginalized_likelihood(self, parameters, rng=None):
        """base.py:544-578: draw a response-curve index from its posterior."""
        log_like = self.get_calibration_log_likelihoods(parameters)
        post = np.exp(log_like - np.max(log_like))
        post /= np.sum(post)
        rng = rng if rng is not None else np.random.default_rng()
        return int(rng.choice(self.number_of_response_curves, p=post))

    def create_lookup_table(self):
        """base.py:994-1018 on the GPU (gwl_distance_table)."""
        return create_lookup_table(self.device, self._distance_array, self.distance_prior_array,
                                   self._ref_dist, self.phase_marginalization)

    def __del__(self):
        try:
            if getattr(self, "_calib", None) is not None:
                LIB.gwl_calib_destroy(self._calib)
                self._calib = None
            if self._event is not None:
                LIB.gwl_event_destroy(self._event)
                self._event = None
            if getattr(self, "_up_event", None) is not None:
                LIB.gwl_event_destroy(self._up_event)
                self._up_event = None
        except Exception:
            pass

    # ------------------------------------------------------------------ per detector
    def compute_per_detector_log_likelihood(self, parameters):
        """base.py:479-500: each detector's own log-likelihood ratio, with
        the time / distance / phase marginalisation of this likelihood, as
        parameters['<ifo>_log_likelihood'] (the dict is updated, as there,
        and a copy returned).  Evaluated by one single-detector companion
        likelihood per detector on the GPU (float64); under calibration marginalisation
        each companion holds this likelihood's curves of its detector, so a
        detector's value is marginalised over its own curves
        (calculate_snrs per interferometer, base.py:296-346, 860-877)."""
        if getattr(self, "_per_detector", None) is None:
            prec = "fp64"
            self._per_detector = []
            for ifo in self.interferometers:
                lk = GravitationalWaveTransient(
                    InterferometerList([ifo]), self.waveform_generator,
                    time_marginalization=self.time_marginalization,
                    distance_marginalization=self.distance_marginalization,
                    phase_marginalization=self.phase_marginalization,
                    priors=None if self.priors is None else self.priors.copy(),
                    jitter_time=self.jitter_time, reference_frame=self.reference_frame,
                    time_reference=self.time_reference, precision=prec, device=self.device)
                if self.calibration_marginalization:
                    lk.calibration_marginalization = True
                    lk._attach_calibration_curves({ifo.name: self.calibration_draws[ifo.name]})
                    lk.calibration_mode = lk.precision
                self._per_detector.append(lk)
        for ifo, lk in zip(self.interferometers, self._per_detector):
            parameters["%s_log_likelihood" % ifo.name] = lk.log_likelihood_ratio(dict(parameters))
        return parameters.copy()

    @property
    def lal_version(self):
        """base.py:1139-1149: lal is not used (the source models are this
        library's own); "N/A" as the reference reports without lal."""
        return "N/A"

    @property
    def lalsimulation_version(self):
        return "N/A"

    # ------------------------------------------------------------------ reconstruction
    def generate_posterior_sample_from_marginalized_likelihood(self, parameters, rng=None):
        """base.py:502-542: new values of the marginalised parameters drawn
        from their conditional posteriors given the rest of ``parameters``
        (Thrane & Talbot, arXiv:1809.02293, Eq. C29-C32), in bilby's order:
        calibration curve, time, distance, phase.  ``rng``: a numpy Generator
        (bilby uses its global one)."""
        if len(self._marginalized_parameters) == 0:
            return parameters
        rng = np.random.default_rng() if rng is None else rng
        p = dict(parameters)
        if self.calibration_marginalization:
            p["recalib_index"] = self.generate_calibration_sample_from_marginalized_likelihood(
                p, rng=rng)
        if self.time_marginalization:
            p["geocent_time"] = self.generate_time_sample_from_marginalized_likelihood(p, rng=rng)
        if self.distance_marginalization:
            p["luminosity_distance"] = self.generate_distance_sample_from_marginalized_likelihood(
                p, rng=rng)
        if self.phase_marginalization:
            p["phase"] = self.generate_phase_sample_from_marginalized_likelihood(p, rng=rng)
        return p.copy()

    def _plain_inner_products(self, parameters):
        """<d|h> (complex) and <h|h> of one parameter dict in the sky frame,
        from a companion float64 likelihood without marginalisation on the
        same data (base.py:686-694, _calculate_inner_products)."""
        if getattr(self, "_plain", None) is None:
            self._plain = GravitationalWaveTransient(self.interferometers, self.waveform_generator,
                                                     precision="fp64", device=self.device)
        if self.calibration_marginalization and "recalib_index" in parameters:
            return self._calibrated_inner_products(parameters, int(parameters["recalib_index"]))
        if self._host_model:   # the caller's model reads the named parameters
            rows = {k: np.atleast_1d(v) for k, v in parameters.items()}
        else:
            rows = self._plain.rows_from_parameters(dict(parameters))[:1]
        _, aux = self._plain.log_likelihood_ratio_batch(rows, return_inner_products=True)
        return complex(aux[0, 0], aux[0, 1]), float(aux[0, 2])

    def _calibrated_inner_products(self, parameters, index):
        """base.py:286-294: <d|h>, <h|h> with response curve ``index`` applied
        to the signal on each detector's mask (gwl_weighted_inner_products on
        the companion plain event)."""
        row = np.ascontiguousarray(self._plain.rows_from_parameters(dict(parameters))[0])
        nf = len(self.interferometers.frequency_array)
        factor = np.zeros((len(self.interferometers), nf, 2))
        factor[..., 0] = 1.0
        for k, ifo in enumerate(self.interferometers):
            m = np.asarray(ifo.frequency_mask)
            c = np.asarray(self.calibration_draws[ifo.name][index])
            factor[k, m, 0], factor[k, m, 1] = c.real, c.imag
        out = np.zeros(3)
        check(LIB.gwl_weighted_inner_products(self._plain._event, dptr(row), dptr(factor),
                                              dptr(out)), "gwl_weighted_inner_products")
        return complex(out[0], out[1]), float(out[2])

    def _distance_marginalized_values(self, d_inner_h, h_inner_h, parameters):
        """base.py:775-784 on the host for an array of <d|h>: the lookup
        table's bicubic interpolant (-inf outside its axes)."""
        from scipy.interpolate import RectBivariateSpline
        if getattr(self, "_host_interp", None) is None:
            x = d_inner_h_ref_axis(self.phase_marginalization)
            y = optimal_snr_squared_ref_axis()
            self._host_interp = (RectBivariateSpline(x, y, self._dist_margd_loglikelihood_array.T,
                                                     kx=3, ky=3, s=0), x, y)
        spline, x, y = self._host_interp
        scale = parameters["luminosity_distance"] / self._ref_dist
        d = np.asarray(d_inner_h) * scale
        d = np.abs(d) if self.phase_marginalization else d.real
        h = np.broadcast_to(np.asarray(h_inner_h, dtype=float) * scale ** 2, d.shape)
        out = spline(d, h, grid=False)
        bad = (d < x.min()) | (d > x.max()) | (h < y.min()) | (h > y.max())
        return np.where(bad, -np.inf, out)

    def _upsampled_time_event(self):
        """The reconstruction grid of base.py:580-640: the data and PSD
        zero-padded (PSD +inf, zero weight) to 16384 T + 1 frequency bins,
        so that the event's transform (length 16384 T) steps 1/16384 s; a
        time window widened by the time-jitter range and a millisecond, the
        prior itself being applied on the host."""
        if getattr(self, "_up_event", None) is None:
            T = float(self.interferometers.duration)
            n_freq = int(round(16384 * T)) + 1
            tp = self.priors["geocent_time"]
            pad = 2.0 / self.interferometers.sampling_frequency + 1e-3
            flags = _lib.TIME_MARG | (_lib.JITTER_TIME if self.jitter_time else 0)
            self._up_event, _ = self._make_event(n_freq=n_freq, flags=flags,
                                                 time_prior=(tp.minimum - pad, tp.maximum + pad))
        return self._up_event

    def _time_posterior(self, parameters):
        """base.py:580-640: times on the 1/16384 s grid from the row's
        (jittered) geocent time, modulo the segment, inside the prior; the
        time-conditioned lnL (distance / phase marginalised as configured)
        times the prior."""
        if self._host_model:
            raise NotImplementedError("the time reconstruction needs a source model evaluated on "
                                      "the device (sine-Gaussian or TaylorF2)")
        p = self._frame(dict(parameters))
        for key, value in self._defaults().items():
            p.setdefault(key, value)
        tg = float(p["geocent_time"]) + (float(p.get("time_jitter", 0.0)) if self.jitter_time
                                         else 0.0)
        ev = self._upsampled_time_event()
        j_first, n_window, t_step = ctypes.c_int64(), ctypes.c_int32(), ctypes.c_double()
        check(LIB.gwl_event_time_window(ev, ctypes.byref(j_first), ctypes.byref(n_window),
                                        ctypes.byref(t_step)), "gwl_event_time_window")
        rows = torch.from_numpy(self._row_builder_sky()(p)[:1]).to("cuda:%d" % self.device)
        L = int(n_window.value)
        dh = torch.empty((1, L, 2), dtype=torch.float64, device=rows.device)
        hh = torch.empty(1, dtype=torch.float64, device=rows.device)
        st = torch.cuda.current_stream(rows.device).cuda_stream
        check(LIB.gwl_eval_time_series(ev, ctypes.c_void_p(rows.data_ptr()), 1, _lib.NPARAM,
                                       _lib.FP64, ctypes.c_void_p(dh.data_ptr()),
                                       ctypes.c_void_p(hh.data_ptr()), ctypes.c_void_p(st)),
              "gwl_eval_time_series")
        T = float(self.interferometers.duration)
        dh = dh.cpu().numpy()[0]
        # base.py:604-606 transforms h conj(d) / S without the 4/T of the
        # inner product (equal for T = 4 s only); kept for identical results
        d_inner_h = (dh[:, 0] + 1j * dh[:, 1]) * (T / 4.0)
        h_inner_h = float(hh.cpu().numpy()[0])
        start = float(self.interferometers.start_time)
        j = int(j_first.value) + np.arange(L)
        times = ((tg - start) + j / 16384.0) % T + start
        prior = self.priors["geocent_time"]
        inside = (times >= prior.minimum) & (times < prior.maximum)
        # the companion event's window covers the prior for rows at the
        # likelihood's fixed time (the segment start, plus jitter); every grid
        # time inside the prior must come from it
        n_fft = int(round(16384 * T))
        all_times = ((tg - start) + np.arange(n_fft) / 16384.0) % T + start
        js = np.flatnonzero((all_times >= prior.minimum) & (all_times < prior.maximum))
        if len(js) == int(np.count_nonzero(inside)):
            times, d_inner_h = times[inside], d_inner_h[inside]
        else:
            # another geocent_time: the direct transform at the window it needs
            times, d_inner_h, h_inner_h = self._direct_time_window(p, all_times, js, n_fft)
        if self.distance_marginalization:
            log_like = self._distance_marginalized_values(d_inner_h, h_inner_h, p)
        elif self.phase_marginalization:
            log_like = ln_i0_host(np.abs(d_inner_h)) - h_inner_h / 2
        else:
            log_like = d_inner_h.real - h_inner_h / 2
        post = np.exp(log_like - np.max(log_like)) * prior.prob(times)
        return times, post

    def _direct_time_window(self, p, all_times, js, n_fft):
        """base.py:604-629 at the grid times ``js`` (ascending, possibly
        wrapping past the segment end) by gwl_window_time_series on this
        likelihood's event."""
        if len(js) == 0:
            raise ValueError("no reconstruction time lies inside the geocent_time prior")
        gaps = np.flatnonzero(np.diff(js) > 1)
        if len(gaps) == 0:
            j0 = int(js[0])
        elif len(gaps) == 1 and js[0] == 0 and js[-1] == n_fft - 1:
            j0 = int(js[gaps[0] + 1])
        else:
            raise ValueError("the geocent_time prior is not one interval on the grid")
        row = np.ascontiguousarray(self._row_builder_sky()(dict(p))[0])
        out = np.zeros((len(js), 2))
        hh = np.zeros(1)
        check(LIB.gwl_window_time_series(self._event, dptr(row), n_fft, j0, len(js), dptr(out),
                                         dptr(hh)), "gwl_window_time_series")
        out = out[(js - j0) % n_fft]
        T = float(self.interferometers.duration)
        return all_times[js], (out[:, 0] + 1j * out[:, 1]) * (T / 4.0), float(hh[0])

    def _row_builder_sky(self):
        return _abi.RowBuilder(self.waveform_generator.parameter_conversion, {})

    def generate_time_sample_from_marginalized_likelihood(self, parameters, rng=None):
        """base.py:580-640: keep the times above 1/1000 of the posterior's
        peak (at least three around it) and draw from the interpolated
        posterior."""
        times, post = self._time_posterior(parameters)
        keep = post > np.max(post) / 1000
        if np.sum(keep) < 3:
            keep[1:-1] = keep[1:-1] | keep[2:] | keep[:-2]
        return interped_sample(times[keep], post[keep],
                               np.random.default_rng() if rng is None else rng)

    def _distance_posterior(self, parameters):
        """base.py:642-684 (Eq. C31): lnL on the distance grid from <d|h>,
        <h|h> at the parameters' distance, times the prior."""
        p = self._frame(dict(parameters))
        d_inner_h, h_inner_h = self._plain_inner_products(p)
        ratio = p["luminosity_distance"] / self._distance_array
        d = d_inner_h * ratio
        h = h_inner_h * ratio ** 2
        log_like = ln_i0_host(np.abs(d)) - h / 2 if self.phase_marginalization else d.real - h / 2
        return self._distance_array, np.exp(log_like - np.max(log_like)) * self.distance_prior_array

    def generate_distance_sample_from_marginalized_likelihood(self, parameters, rng=None):
        x, post = self._distance_posterior(parameters)
        return interped_sample(x, post, np.random.default_rng() if rng is None else rng)

    def _phase_posterior(self, parameters):
        """base.py:712-742 (mu(phi) ~ exp(-2 i phi)): lnL on 101 phases."""
        p = self._frame(dict(parameters))
        d_inner_h, h_inner_h = self._plain_inner_products(p)
        phases = np.linspace(0, 2 * np.pi, 101)
        log_post = (d_inner_h * np.exp(-2j * phases)).real - h_inner_h / 2
        return phases, np.exp(log_post - np.max(log_post))

    def generate_phase_sample_from_marginalized_likelihood(self, parameters, rng=None):
        x, post = self._phase_posterior(parameters)
        return interped_sample(x, post, np.random.default_rng() if rng is None else rng)

    # ------------------------------------------------------------------ API
    @property
    def marginalized_parameters(self):
        return self._marginalized_parameters

    @property
    def meta_data(self):
        return dict(interferometers=[ifo.name for ifo in self.interferometers],
                    time_marginalization=self.time_marginalization,
                    phase_marginalization=self.phase_marginalization,
                    distance_marginalization=self.distance_marginalization,
                    calibration_marginalization=self.calibration_marginalization,
                    waveform_arguments=self.waveform_generator.waveform_arguments,
                    time_reference=self.time_reference,
                    reference_frame=(self.reference_frame if isinstance(self.reference_frame, str)
                                     else "".join(getattr(i, "name", str(i))
                                                  for i in self.reference_frame)),
                    precision="fp32" if self.precision == _lib.FP32 else "fp64",
                    backend="libgwl (HIP, gfx950)")

    def noise_log_likelihood(self):
        """base.py:402-417 (computed once on the device, then cached)."""
        if self._noise_log_likelihood_value is None:
            out = ctypes.c_double()
            check(LIB.gwl_noise_log_likelihood(self._event, ctypes.byref(out)),
                  "gwl_noise_log_likelihood")
            self._noise_log_likelihood_value = float(out.value)
        return self._noise_log_likelihood_value

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

    def row_builder(self):
        """The host-side dict -> rows conversion of this likelihood
        (picklable; what worker-process clients use)."""
        return _abi.RowBuilder(self.waveform_generator.parameter_conversion, self._defaults(),
                               frame=self._frame)

    def get_sky_frame_parameters(self, parameters):
        """base.py:1091-1137: ra, dec and geocent_time of a parameter dict in
        this likelihood's reference frame and time reference."""
        p = self._frame(dict(parameters))
        return dict(ra=p["ra"], dec=p["dec"], geocent_time=p["geocent_time"])

    def rows_from_parameters(self, parameters):
        """Parameter dict (scalars or 1-D arrays) -> float64 rows [n, 16] in
        the column order of include/gwl.h (enum gwl_param)."""
        return self.row_builder()(parameters)

    def _device_eval(self, rows_t, precision, aux=False, stream=None):
        n = rows_t.shape[0]
        dev = rows_t.device
        out = torch.empty(n, dtype=torch.float64, device=dev)
        aux_t = torch.empty((n, 3), dtype=torch.float64, device=dev) if aux else None
        st = torch.cuda.current_stream(dev).cuda_stream if stream is None else stream
        if self._calib is not None:
            if aux:
                raise NotImplementedError("inner products are not returned under calibration "
                                          "marginalisation")
            check(LIB.gwl_calib_eval(self._calib, ctypes.c_void_p(rows_t.data_ptr()), n,
                                     _lib.NPARAM, int(precision), ctypes.c_void_p(out.data_ptr()), None,
                                     ctypes.c_void_p(st)), "gwl_calib_eval")
            return out, None
        check(LIB.gwl_eval(self._event, ctypes.c_void_p(rows_t.data_ptr()), n, _lib.NPARAM,
                           precision, ctypes.c_void_p(out.data_ptr()),
                           ctypes.c_void_p(aux_t.data_ptr()) if aux else None,
                           ctypes.c_void_p(st)), "gwl_eval")
        return out, aux_t

    def set_precision_escalation(self, coeff=2e-7, tol=8e-4):
        """Float32 evaluations without time marginalisation re-evaluate in
        float64, inside the same gwl_eval call, every row whose float32 error
        bound ``coeff * Q`` exceeds ``tol`` (Q: the row's inner-product
        scale, include/gwl.h gwl_event_set_escalation); ``coeff = 0`` turns
        it off.  No counterpart in the reference (float64 throughout)."""
        check(LIB.gwl_event_set_escalation(self._event, float(coeff), float(tol)),
              "gwl_event_set_escalation")

    def escalated_rows(self):
        """Rows the last escalating float32 evaluation re-evaluated in
        float64 (synchronises the device)."""
        v = ctypes.c_int64()
        check(LIB.gwl_event_escalated_rows(self._event, ctypes.byref(v)),
              "gwl_event_escalated_rows")
        return int(v.value)

    def _mode(self, precision):
        """Precision code of one call: the constructor's choice by default;
        "bf16x3" selects the split-bfloat16 contraction under calibration
        marginalisation."""
        if precision is None:
            return self.calibration_mode if self._calib is not None else self.precision
        if self._calib is not None and self.time_marginalization and precision == "bf16x3":
            raise NotImplementedError("calibration together with time marginalisation is "
                                      "evaluated in float32 or float64")
        if precision == "bf16x3":
            if self._calib is None:
                raise ValueError("bf16x3 applies to calibration marginalisation only")
            return 2
        return _precision_code(precision)

    # ------------------------------------------------------------------ host source models
    @property
    def _host_model(self):
        return self.waveform_generator.approximant == _lib.HOST_POLARIZATIONS

    def _model_dicts(self, parameters, rows):
        """Per-row parameter dicts for a source model the caller evaluates:
        a dict's entries row by row (with the marginalised keys' defaults),
        or the named columns of parameter rows."""
        n = rows.shape[0]
        if isinstance(parameters, dict):
            cols = {k: np.asarray(v) for k, v in parameters.items()}
            out = []
            for i in range(n):
                d = {k: (v.item() if v.ndim == 0 else v[i].item()) for k, v in cols.items()}
                for k, v in self._defaults().items():
                    d.setdefault(k, v)
                out.append(d)
            return out
        return [{k: float(rows[i, c]) for k, c in _lib.PARAM_COLUMNS.items()} for i in range(n)]

    def _polarizations(self, dicts):
        """waveform_generator.py:113-141 per row: the model's h+ and hx on
        the frequency array, [n, 2, n_freq] complex, and the rows whose model
        returned None (base.py:424-425 gives them -inf as nan_to_num)."""
        wfg = self.waveform_generator
        nf = len(self.interferometers.frequency_array)
        pol = np.zeros((len(dicts), 2, nf), dtype=np.complex128)
        invalid = np.zeros(len(dicts), dtype=bool)
        for i, d in enumerate(dicts):
            h = wfg.frequency_domain_strain(d)
            if h is None:
                invalid[i] = True
                continue
            pol[i, 0] = h["plus"]
            pol[i, 1] = h["cross"]
        return pol, invalid

    def _host_batch(self, parameters, prec, return_inner_products, stream, chunk=256):
        """log_likelihood_ratio_batch for a source model the caller
        evaluates: per chunk of rows the model on the host, its polarizations
        to the device (gwl_event_set_polarizations), then gwl_eval."""
        on_device = isinstance(parameters, torch.Tensor) and parameters.is_cuda
        if isinstance(parameters, dict):
            rows = self.rows_from_parameters(parameters)
        elif on_device:
            rows = parameters.detach().to(torch.float64).cpu().numpy()
        else:
            rows = np.ascontiguousarray(parameters, dtype=np.float64)
        dicts = self._model_dicts(parameters, rows)
        dev = "cuda:%d" % self.device
        outs, auxs, keep = [], [], []
        for a in range(0, rows.shape[0], chunk):
            pol, invalid = self._polarizations(dicts[a:a + chunk])
            pol_t = torch.from_numpy(np.ascontiguousarray(pol).view(np.float64)).to(dev)
            rows_t = torch.from_numpy(np.ascontiguousarray(rows[a:a + chunk])).to(dev)
            keep.append(pol_t)
            check(LIB.gwl_event_set_polarizations(self._event, ctypes.c_void_p(pol_t.data_ptr()),
                                                  pol.shape[0]), "gwl_event_set_polarizations")
            out, aux = self._device_eval(rows_t, prec, aux=return_inner_products, stream=stream)
            if invalid.any():
                out[torch.from_numpy(np.nonzero(invalid)[0]).to(dev)] = -np.finfo(float).max
            outs.append(out)
            auxs.append(aux)
        out = torch.cat(outs)
        aux = torch.cat(auxs) if return_inner_products else None
        if on_device:
            return (out, aux) if return_inner_products else out
        out_h = out.cpu().numpy()
        return (out_h, aux.cpu().numpy()) if return_inner_products else out_h

    def log_likelihood_ratio_batch(self, parameters, precision=None, return_inner_products=False,
                                   stream=None):
        """Batched ln L ratio over N rows.

        ``parameters``: a dict of 1-D arrays (bilby parameter names), a
        float64 array [N, 16] in gwl column order, or a torch float64 tensor
        [N, 16] already on the GPU (then a GPU tensor is returned and nothing
        is synchronised).
        """
        prec = self._mode(precision)
        on_device = isinstance(parameters, torch.Tensor) and parameters.is_cuda
        if _batch_rows(parameters) == 0:
            empty = (torch.empty(0, dtype=torch.float64, device=parameters.device) if on_device
                     else np.zeros(0))
            if return_inner_products:
                return empty, (torch.empty((0, 3), dtype=torch.float64, device=parameters.device)
                               if on_device else np.zeros((0, 3)))
            return empty
        if self._host_model:
            return self._host_batch(parameters, prec, return_inner_products, stream)
        if on_device:
            rows_t = parameters.to(torch.float64).contiguous()
        else:
            if isinstance(parameters, dict):
                rows = self.rows_from_parameters(parameters)
            else:
                rows = np.ascontiguousarray(parameters, dtype=np.float64)
            if stream is None and not return_inner_products:
                return self._staged_eval(rows, prec)
            rows_t = torch.from_numpy(rows).to("cuda:%d" % self.device)
        out, aux = self._device_eval(rows_t, prec, aux=return_inner_products, stream=stream)
        if on_device:
            return (out, aux) if return_inner_products else out
        out_h = out.cpu().numpy()
        if return_inner_products:
            return out_h, aux.cpu().numpy()
        return out_h

    def _staged_eval(self, rows, precision):
        """Host rows -> host lnL through reused pinned buffers: one
        asynchronous copy in, the evaluation, one copy out and one stream
        synchronisation (the sampler pools' path: no pageable copies and no
        allocations per call).  Calls from several threads take turns."""
        n = rows.shape[0]
        with self._stage_lock:
            cap = 0 if self._stage is None else self._stage[0].shape[0]
            if n > cap:
                cap = max(n, 2 * cap, 1024)
                dev = "cuda:%d" % self.device
                self._stage = (torch.empty((cap, _lib.NPARAM), dtype=torch.float64).pin_memory(),
                               torch.empty((cap, _lib.NPARAM), dtype=torch.float64, device=dev),
                               torch.empty(cap, dtype=torch.float64).pin_memory())
            h_in, d_in, h_out = self._stage
            h_in.numpy()[:n] = rows
            d_rows = d_in[:n]
            d_rows.copy_(h_in[:n], non_blocking=True)
            out, _ = self._device_eval(d_rows, precision)
            h_out[:n].copy_(out, non_blocking=True)
            torch.cuda.current_stream(d_in.device).synchronize()
            return h_out.numpy()[:n].copy()

    def log_likelihood_ratio(self, parameters):
        """base.py:419-446: one parameter dict -> float (the caller's dict is
        not modified)."""
        if self._host_model:
            q = {k: np.atleast_1d(v) for k, v in parameters.items()}
            return float(self._host_batch(q, self._mode(None), False, None)[0])
        rows = self.rows_from_parameters(copy.copy(parameters))
        if self._one is None:
            self._one = (torch.empty((1, _lib.NPARAM), dtype=torch.float64).pin_memory(),
                         torch.empty((1, _lib.NPARAM), dtype=torch.float64,
                                     device="cuda:%d" % self.device))
        host, dev = self._one
        host.numpy()[:] = rows[:1]
        dev.copy_(host, non_blocking=True)
        out, _ = self._device_eval(dev, self._mode(None))
        return float(out.item())

    def log_likelihood(self, parameters):
        return self.log_likelihood_ratio(parameters) + self.noise_log_likelihood()

    def get_detector_response(self, parameters, detector_index=0, precision="fp64"):
        """interferometer.py:303-368 for this likelihood's source model."""
        rows = self.rows_from_parameters(parameters)
        row = np.ascontiguousarray(rows[0])
        out = np.zeros((len(self.interferometers.frequency_array), 2))
        pol_t = None
        if self._host_model:
            pol, _ = self._polarizations(self._model_dicts(
                {k: np.atleast_1d(v) for k, v in parameters.items()}, rows[:1]))
            pol_t = torch.from_numpy(np.ascontiguousarray(pol).view(np.float64)).to(
                "cuda:%d" % self.device)
            check(LIB.gwl_event_set_polarizations(self._event, ctypes.c_void_p(pol_t.data_ptr()),
                                                  1), "gwl_event_set_polarizations")
        check(LIB.gwl_detector_response(self._event, dptr(row), int(detector_index),
                                        _precision_code(precision), dptr(out)),
              "gwl_detector_response")
        return out[:, 0] + 1j * out[:, 1]

    def __repr__(self):
        return ("GravitationalWaveTransient(interferometers={},\n\twaveform_generator={},\n\t"
                "time_marginalization={}, distance_marginalization={}, phase_marginalization={}, "
                "calibration_marginalization={}, priors={})".format(
                    [i.name for i in self.interferometers], self.waveform_generator,
                    self.time_marginalization, self.distance_marginalization,
                    self.phase_marginalization, self.calibration_marginalization, self.priors))


def inject_signal(interferometers, waveform_generator, parameters, device=0):
    """Add the detector response of ``parameters`` to each detector's data
    (bilby/gw/detector/interferometer.py:412-523), computed in float64 on
    the GPU through a zero-data event."""
    ifos = InterferometerList(interferometers)
    saved = [ifo._strain.copy() for ifo in ifos]
    for ifo in ifos:
        ifo._strain = np.zeros_like(ifo._strain)
    try:
        lk = GravitationalWaveTransient(ifos, waveform_generator, precision="fp64", device=device)
        signals = [lk.get_detector_response(parameters, k) for k in range(len(ifos))]
    finally:
        for ifo, s in zip(ifos, saved):
            ifo._strain = s
    for ifo, sig in zip(ifos, signals):
        ifo.add_to_strain(sig)
    return signals
