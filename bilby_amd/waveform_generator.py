"""Source models and the waveform generator, as the likelihood sees them.

bilby's WaveformGenerator (bilby/gw/waveform_generator.py:113-286) calls a
Python source function on the host per parameter dict.  Here the source
model is evaluated inside the fused GPU kernels; the generator only names
it and carries its ``waveform_arguments`` and parameter conversion.
"""
import numpy as np

from . import abi as _lib


def sinegaussian(frequency_array, hrss, Q, frequency, **kwargs):
    """Sine-Gaussian burst (bilby/gw/source.py:1153-1209), evaluated on the
    GPU by the likelihood kernels."""
    raise NotImplementedError("evaluated on the device by GravitationalWaveTransient")


sinegaussian._gwl_approximant = _lib.SINE_GAUSSIAN


def taylorf2(frequency_array, mass_1, mass_2, luminosity_distance, theta_jn, phase,
             chi_1=0.0, chi_2=0.0, **kwargs):
    """Aligned-spin 3.5PN TaylorF2 inspiral (the TaylorF2 approximant bilby
    reaches through source.py:269-348, 552-690), evaluated on the GPU."""
    raise NotImplementedError("evaluated on the device by GravitationalWaveTransient")


taylorf2._gwl_approximant = _lib.TAYLORF2


def taylorf2_flags(waveform_arguments):
    """Event flag for the end of TaylorF2.  By default the waveform ends at
    ``maximum_frequency``: bilby hands that frequency to lalsimulation as fEnd
    (source.py:338-343, 633-643) and XLALSimInspiralTaylorF2 ends at the ISCO
    only for fEnd = 0, which bilby never passes.  ``waveform_arguments
    ['taylorf2_termination'] = 'isco'`` (no counterpart in the reference)
    also ends it at the Schwarzschild ISCO, 1 / (6^3/2 pi M)."""
    end = (waveform_arguments or {}).get("taylorf2_termination", "maximum_frequency")
    if end == "maximum_frequency":
        return 0
    if end == "isco":
        return _lib.TF2_END_ISCO
    raise ValueError("taylorf2_termination must be 'maximum_frequency' or 'isco', not %r" % (end,))


def lal_binary_black_hole(frequency_array, mass_1, mass_2, luminosity_distance, a_1, tilt_1,
                          phi_12, a_2, tilt_2, phi_jl, theta_jn, phase, **kwargs):
    """bilby/gw/source.py:269-348 entry point; only
    waveform_approximant='TaylorF2' with aligned spins is supported."""
    raise NotImplementedError("evaluated on the device by GravitationalWaveTransient")


lal_binary_black_hole._gwl_approximant = "lal"


def generate_component_masses(p):
    """bilby/gw/conversion.py generate_component_masses for chirp_mass /
    mass_ratio and total_mass / mass_ratio inputs."""
    if "mass_1" in p and "mass_2" in p:
        return p
    q = p.get("mass_ratio")
    if "chirp_mass" in p and q is not None:
        mc = np.asarray(p["chirp_mass"], float)
        q = np.asarray(q, float)
        total = mc * (1 + q) ** 1.2 / q ** 0.6
        p["mass_1"] = total / (1 + q)
        p["mass_2"] = p["mass_1"] * q
    elif "total_mass" in p and q is not None:
        q = np.asarray(q, float)
        p["mass_1"] = np.asarray(p["total_mass"], float) / (1 + q)
        p["mass_2"] = p["mass_1"] * q
    return p


def convert_to_lal_binary_black_hole_parameters(parameters):
    """The parts of bilby/gw/conversion.py:182-283 that apply to aligned
    spins: component masses, chi_i from a_i cos(tilt_i), cos_theta_jn."""
    p = dict(parameters)
    generate_component_masses(p)
    for idx in ("1", "2"):
        if "chi_" + idx not in p:
            a = p.get("a_" + idx, 0.0)
            if "cos_tilt_" + idx in p:
                p["chi_" + idx] = np.asarray(a) * np.asarray(p["cos_tilt_" + idx])
            else:
                p["chi_" + idx] = np.asarray(a) * np.cos(p.get("tilt_" + idx, 0.0))
    if "theta_jn" not in p and "cos_theta_jn" in p:
        p["theta_jn"] = np.arccos(p["cos_theta_jn"])
    return p


class WaveformGenerator:
    """bilby/gw/waveform_generator.py:14-286 (frequency-domain only)."""

    def __init__(self, duration=None, sampling_frequency=None, start_time=0,
                 frequency_domain_source_model=None, time_domain_source_model=None,
                 parameters=None, parameter_conversion=None, waveform_arguments=None):
        if time_domain_source_model is not None:
            raise NotImplementedError("time-domain source models are not supported")
        self.duration = duration
        self.sampling_frequency = sampling_frequency
        self.start_time = start_time
        self.frequency_domain_source_model = frequency_domain_source_model
        self.waveform_arguments = dict(waveform_arguments or {})
        if parameter_conversion is None:
            parameter_conversion = convert_to_lal_binary_black_hole_parameters
        self.parameter_conversion = parameter_conversion

    @property
    def approximant(self):
        model = self.frequency_domain_source_model
        tag = getattr(model, "_gwl_approximant", None)
        if tag is None and isinstance(model, str):
            tag = {"sinegaussian": _lib.SINE_GAUSSIAN, "taylorf2": _lib.TAYLORF2}.get(model.lower())
        if tag == "lal":
            approx = self.waveform_arguments.get("waveform_approximant", "IMRPhenomPv2")
            if approx != "TaylorF2":
                raise NotImplementedError(
                    "waveform_approximant %r is not available on the device" % approx)
            tag = _lib.TAYLORF2
        if tag not in (_lib.SINE_GAUSSIAN, _lib.TAYLORF2):
            raise NotImplementedError("unsupported source model %r" % (model,))
        return tag

    @property
    def frequency_array(self):
        from .detector import create_frequency_series
        return create_frequency_series(self.sampling_frequency, self.duration)
