"""Source models and the waveform generator, as the likelihood sees them.

bilby's WaveformGenerator (bilby/gw/waveform_generator.py:113-286) calls a
Python source function on the host per parameter dict.  The built-in models
(sine-Gaussian, TaylorF2) are evaluated inside the fused GPU kernels; the
generator then only names them and carries their ``waveform_arguments`` and
parameter conversion.  Any other ``frequency_domain_source_model`` is called
on the host per sample, as bilby calls it, and its polarizations go to the
GPU for the projection, inner products and marginalisations.
"""
import functools
import inspect

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


def _mass_ratio_from_component_and_chirp_mass(m, mc):
    """The other mass over m, x, from a component mass m and the chirp mass:
    the positive root of x^3 - a x - a = 0, a = (mc / m)^5 (from mc =
    m x^(3/5) / (1 + x)^(1/5)); there is exactly one (Descartes).  Cardano's
    formula for a < 27/4, where the cubic has one real root; the
    trigonometric form of its largest root above, where it has three (x > 3,
    reached from the lighter mass of a pair with q < 1/3, where bilby's
    closed form, conversion.py:1072-1095, returns NaN)."""
    a = np.asarray((np.asarray(mc, float) / np.asarray(m, float)) ** 5)
    with np.errstate(invalid="ignore"):
        disc = np.sqrt(np.maximum(a * a / 4 - a ** 3 / 27, 0.0))
        cardano = np.cbrt(a / 2 + disc) + np.cbrt(a / 2 - disc)
        trig = 2 * np.sqrt(a / 3) * np.cos(np.arccos(np.minimum(1.5 * np.sqrt(3 / a), 1.0)) / 3)
    out = np.where(a < 6.75, cardano, trig)
    return out[()] if out.ndim == 0 else out


def _mass_ratio_from_symmetric(eta):
    """q <= 1 with q / (1 + q)^2 = eta."""
    t = 1 / (2 * np.asarray(eta, float)) - 1
    return t - np.sqrt(t * t - 1)


def generate_component_masses(sample, require_add=False, source=False):
    """bilby/gw/conversion.py:1826-1975: add mass_1 and mass_2 (or the
    ``_source`` pair) from whichever complete set of mass parameters the
    sample holds (a component mass with total mass, mass ratio, symmetric
    mass ratio or chirp mass; total or chirp mass with mass ratio or
    symmetric mass ratio; chirp and total mass).  Incomplete sets add
    nothing, or raise KeyError with ``require_add``."""
    out = dict(sample)
    sfx = "_source" if source else ""
    m1k, m2k, mtk, mck = "mass_1" + sfx, "mass_2" + sfx, "total_mass" + sfx, "chirp_mass" + sfx

    def incomplete():
        if require_add:
            raise KeyError("Insufficient mass parameters in input dictionary")
        return dict(sample)

    if m1k in out or m2k in out:
        known, other = (m1k, m2k) if m1k in out else (m2k, m1k)
        if other in out:
            return out
        if mtk in out:
            out[other] = np.asarray(out[mtk], float) - np.asarray(out[known], float)
            return out
        if "mass_ratio" in out:
            q = np.asarray(out["mass_ratio"], float)
        elif "symmetric_mass_ratio" in out:
            q = out["mass_ratio"] = _mass_ratio_from_symmetric(out["symmetric_mass_ratio"])
        elif mck in out:
            q = _mass_ratio_from_component_and_chirp_mass(out[known], out[mck])
            if known == m2k:   # the lighter mass gives 1 / q
                q = 1 / q
            out["mass_ratio"] = q
        else:
            return incomplete()
        m = np.asarray(out[known], float)
        out[other] = m * q if known == m1k else m / q
        return out
    if mtk in out:
        if "mass_ratio" not in out:
            if "symmetric_mass_ratio" in out:
                out["mass_ratio"] = _mass_ratio_from_symmetric(out["symmetric_mass_ratio"])
            elif mck in out:
                out["symmetric_mass_ratio"] = (np.asarray(out[mck], float)
                                               / np.asarray(out[mtk], float)) ** (5 / 3)
                out["mass_ratio"] = _mass_ratio_from_symmetric(out["symmetric_mass_ratio"])
            else:
                return incomplete()
    elif mck in out:
        if "mass_ratio" not in out:
            if "symmetric_mass_ratio" in out:
                out["mass_ratio"] = _mass_ratio_from_symmetric(out["symmetric_mass_ratio"])
            else:
                return incomplete()
        q = np.asarray(out["mass_ratio"], float)
        with np.errstate(invalid="ignore", divide="ignore"):
            out[mtk] = np.asarray(out[mck], float) * (1 + q) ** 1.2 / q ** 0.6
    else:
        return incomplete()
    q = np.asarray(out["mass_ratio"], float)
    out[m1k] = np.asarray(out[mtk], float) / (1 + q)
    out[m2k] = out[m1k] * q
    return out


_COSMOLOGY = None


def _cosmology():
    global _COSMOLOGY
    if _COSMOLOGY is None:
        from .prior import FlatLambdaCDM
        _COSMOLOGY = FlatLambdaCDM()
    return _COSMOLOGY


def redshift_to_luminosity_distance(redshift):
    """bilby/gw/conversion.py redshift_to_luminosity_distance, with this
    package's flat Lambda-CDM cosmology (H0 = 67.74, Om0 = 0.3075) in place
    of astropy's Planck15."""
    return _cosmology().luminosity_distance(redshift)


def comoving_distance_to_luminosity_distance(comoving_distance):
    c = _cosmology()
    z = np.interp(comoving_distance, c._dc, c._z)
    return (1 + z) * np.asarray(comoving_distance, float)


def luminosity_distance_to_redshift(distance):
    return _cosmology().z_at_luminosity_distance(distance)


def convert_to_lal_binary_black_hole_parameters(parameters, aligned=True):
    """bilby/gw/conversion.py:182-283 for the aligned-spin TaylorF2 the
    device evaluates: luminosity distance from redshift or comoving distance;
    ``*_source`` masses times (1 + z); component masses from any complete
    mass set; spins from chi_i (a_i = |chi_i|, cos tilt_i = sign chi_i, or
    chi_i / a_i when a_i is given), from a_i and tilt_i / cos_tilt_i, and
    chi_i_in_plane; cos_theta_jn -> theta_jn; delta_phase -> phase =
    (delta_phase - sign(cos theta_jn) psi) mod 2 pi.  Adds chi_1, chi_2 (the
    aligned components the kernels read).  Returns the converted dict (bilby
    also returns the list of added keys).

    TaylorF2 is a non-precessing approximant: lalsimulation refuses non-zero
    in-plane spins for it (XLAL_EINVAL, which bilby re-raises whatever
    catch_waveform_errors says, source.py:644-662), and so does this
    conversion (ValueError) unless ``aligned`` is False (source models
    evaluated by the caller, which take what they take)."""
    p = dict(parameters)
    original = list(p)
    if "luminosity_distance" not in original:
        if "redshift" in p:
            p["luminosity_distance"] = redshift_to_luminosity_distance(p["redshift"])
        elif "comoving_distance" in p:
            p["luminosity_distance"] = comoving_distance_to_luminosity_distance(
                p["comoving_distance"])
    for key in original:
        if key.endswith("_source"):
            if "redshift" not in p:
                p["redshift"] = luminosity_distance_to_redshift(p["luminosity_distance"])
            p[key[:-7]] = np.asarray(p[key], float) * (1 + np.asarray(p["redshift"], float))
    p = generate_component_masses(p, require_add=False)
    for idx in ("1", "2"):
        chi, a, ct = "chi_" + idx, "a_" + idx, "cos_tilt_" + idx
        if chi in original:
            inplane = p.get(chi + "_in_plane", 0.0)
            if aligned and np.any(np.asarray(inplane, float) != 0):
                raise ValueError("TaylorF2 takes aligned spins: %s_in_plane must be zero" % chi)
            if a not in original:
                p[a] = np.abs(np.asarray(p[chi], float))
                p[ct] = np.sign(np.asarray(p[chi], float))
            else:
                with np.errstate(invalid="ignore", divide="ignore"):
                    cos_tilt = np.asarray(p[chi], float) / np.asarray(p[a], float)
                p[ct] = np.where(np.isfinite(cos_tilt), cos_tilt, 1.0)
    for key in ("phi_jl", "phi_12"):
        p.setdefault(key, 0.0)
    for angle in ("tilt_1", "tilt_2", "theta_jn"):
        if "cos_" + angle in p:
            with np.errstate(invalid="ignore"):
                p[angle] = np.arccos(p["cos_" + angle])
    if "delta_phase" in p:
        with np.errstate(invalid="ignore"):
            p["phase"] = np.mod(np.asarray(p["delta_phase"], float)
                                - np.sign(np.cos(p["theta_jn"])) * np.asarray(p["psi"], float),
                                2 * np.pi)
    for idx in ("1", "2"):
        chi = "chi_" + idx
        if chi not in original:
            amp = np.asarray(p.get("a_" + idx, 0.0), float)
            tilt = np.asarray(p.get("tilt_" + idx, 0.0), float)
            # bilby_to_lalsimulation_spins (conversion.py:106-166): the aligned
            # case needs a_i = 0 or tilt_i in {0, pi}; anything else has
            # in-plane spin components
            if aligned and np.any(np.abs(amp * np.sin(tilt)) > 1e-12 * np.maximum(amp, 1.0)):
                raise ValueError("TaylorF2 takes aligned spins: tilt_%s must be 0 or pi" % idx)
            p[chi] = amp * np.cos(tilt)
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
            if self.approximant == _lib.HOST_POLARIZATIONS:
                parameter_conversion = functools.partial(
                    convert_to_lal_binary_black_hole_parameters, aligned=False)
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
        if tag is None and callable(model):
            return _lib.HOST_POLARIZATIONS
        if tag not in (_lib.SINE_GAUSSIAN, _lib.TAYLORF2):
            raise NotImplementedError("unsupported source model %r" % (model,))
        return tag

    @property
    def source_parameter_keys(self):
        """waveform_generator.py:271-286 (core/utils/introspection.py): the
        positional-or-keyword arguments of the source model after the
        frequency array (``*args``, keyword-only and ``**kwargs`` dropped)."""
        sig = inspect.signature(self.frequency_domain_source_model)
        names = [p.name for p in sig.parameters.values() if p.kind == p.POSITIONAL_OR_KEYWORD]
        return set(names[1:])

    def frequency_domain_strain(self, parameters):
        """waveform_generator.py:113-141, 196-212, 254-269 for a source model
        the caller evaluates: the converted parameters the model names, plus
        the waveform arguments, handed to the model with the frequency array;
        its polarizations dict (or None) is returned as the model gives it."""
        if not isinstance(parameters, dict):
            raise TypeError('"parameters" must be a dictionary.')
        p = self.parameter_conversion(dict(parameters))
        if isinstance(p, tuple):  # bilby's conversions also return the added keys
            p = p[0]
        keys = self.source_parameter_keys
        p = {k: v for k, v in p.items() if k in keys}
        p.update(self.waveform_arguments)
        return self.frequency_domain_source_model(self.frequency_array, **p)

    @property
    def frequency_array(self):
        from .detector import create_frequency_series
        return create_frequency_series(self.sampling_frequency, self.duration)
