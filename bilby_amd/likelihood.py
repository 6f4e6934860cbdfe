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
import os

import numpy as np
import torch

from . import _lib
from . import abi as _abi
from ._lib import LIB, check, dptr, iptr
from .detector import InterferometerList
from .prior import Uniform

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
        if calibration_marginalization and time_marginalization and precision != "fp32":
            raise NotImplementedError("calibration together with time marginalisation is "
                                      "evaluated in float32 (precision='fp32')")
        from .frames import sky_frame
        # base.py:170-181: a detector time reference cannot be marginalised over
        self._frame = sky_frame(reference_frame, time_reference)
        self.reference_frame = reference_frame
        self.time_reference = "geocent" if "geocent" in time_reference else time_reference
        if "geocent" not in time_reference and time_marginalization:
            time_marginalization = False
            jitter_time = False
        self.waveform_generator = waveform_generator
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

    # ------------------------------------------------------------------ setup
    def _check_marginalized_prior_is_set(self, key):
        """base.py:356-390 (fixed priors rejected; missing time prior -> the segment)."""
        if key in self.priors and getattr(self.priors[key], "is_fixed", False):
            raise ValueError("Cannot use marginalized likelihood for {}: prior is fixed".format(key))
        if key not in self.priors or isinstance(self.priors[key], (int, float)):
            if key == "geocent_time":
                self.priors[key] = Uniform(self.interferometers.start_time,
                                           self.interferometers.start_time
                                           + self.interferometers.duration)
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
        return f

    def _build_event(self):
        ifos = self.interferometers
        ctx = get_context(self.device)
        nd = len(ifos)
        freq = ifos.frequency_array
        nf = len(freq)
        strain = np.ascontiguousarray(
            np.stack([ifo.frequency_domain_strain for ifo in ifos]).astype(np.complex128))
        strain_f = strain.view(np.float64).reshape(nd, nf, 2).copy()
        psd = np.ascontiguousarray(np.stack([ifo.power_spectral_density_array for ifo in ifos]),
                                   dtype=np.float64)
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
        desc.sampling_frequency = float(ifos.sampling_frequency)
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
        desc.flags = self._flags()
        if self.time_marginalization:
            tp = self.priors["geocent_time"]
            if not isinstance(tp, Uniform):
                raise NotImplementedError("time marginalisation supports a Uniform geocent_time prior")
            desc.time_prior_minimum = float(tp.minimum)
            desc.time_prior_maximum = float(tp.maximum)
        desc.reference_distance = float(self._ref_dist)
        ev = ctypes.c_void_p()
        nbytes = ctypes.c_int64()
        check(LIB.gwl_event_create(ctx, ctypes.byref(desc), ctypes.byref(ev), ctypes.byref(nbytes)),
              "gwl_event_create")
        self._event = ev
        self.hbm_bytes = int(nbytes.value)

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
        self.calibration_abs_draws = {k: np.abs(v) ** 2 for k, v in self.calibration_draws.items()}
        curves = np.concatenate([np.asarray(self.calibration_draws[ifo.name], dtype=complex)
                                 for ifo in self.interferometers], axis=1)
        blob = np.ascontiguousarray(curves.view(np.float64))
        cal = ctypes.c_void_p()
        check(LIB.gwl_calib_create(self._event, int(curves.shape[0]), dptr(blob), ctypes.byref(cal)),
              "gwl_calib_create")
        self._calib = cal

    def get_calibration_log_likelihoods(self, parameters):
        """base.py:822-858: ln L of every response curve for one parameter
        dict (float64 contraction; with time marginalisation each curve's
        time-marginalised value, float32 transforms)."""
        rows = self.rows_from_parameters(dict(parameters))
        rows_t = torch.from_numpy(np.ascontiguousarray(rows[:1])).to("cuda:%d" % self.device)
        out = torch.empty(1, dtype=torch.float64, device=rows_t.device)
        per = torch.empty((1, self.number_of_response_curves), dtype=torch.float64,
                          device=rows_t.device)
        st = torch.cuda.current_stream(rows_t.device).cuda_stream
        mode = _lib.FP32 if self.time_marginalization else _lib.FP64
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
        except Exception:
            pass

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
                    calibration_marginalization=False,
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

    def _mode(self, precision):
        """Precision code of one call: the constructor's choice by default;
        "bf16x3" selects the split-bfloat16 contraction under calibration
        marginalisation."""
        if precision is None:
            return self.calibration_mode if self._calib is not None else self.precision
        if self._calib is not None and self.time_marginalization and precision != "fp32":
            raise NotImplementedError("calibration together with time marginalisation is "
                                      "evaluated in float32")
        if precision == "bf16x3":
            if self._calib is None:
                raise ValueError("bf16x3 applies to calibration marginalisation only")
            return 2
        return _precision_code(precision)

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
        if on_device:
            rows_t = parameters.to(torch.float64).contiguous()
        else:
            if isinstance(parameters, dict):
                rows = self.rows_from_parameters(parameters)
            else:
                rows = np.ascontiguousarray(parameters, dtype=np.float64)
            rows_t = torch.from_numpy(rows).to("cuda:%d" % self.device)
        out, aux = self._device_eval(rows_t, prec, aux=return_inner_products, stream=stream)
        if on_device:
            return (out, aux) if return_inner_products else out
        out_h = out.cpu().numpy()
        if return_inner_products:
            return out_h, aux.cpu().numpy()
        return out_h

    def log_likelihood_ratio(self, parameters):
        """base.py:419-446: one parameter dict -> float (the caller's dict is
        not modified)."""
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
                    self.phase_marginalization, False, self.priors))


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
