"""bilby_amd: bilby's compact-binary likelihood hot path on the AMD MI355X.

Drop-in for bilby.gw.likelihood.GravitationalWaveTransient (parameter dict in,
log-likelihood ratio out) plus a batched call over N parameter rows; all
evaluation runs in hand-written HIP kernels (libgwl.so, gfx950).
"""
from . import detector, prior, waveform_generator  # noqa: F401
from .detector import (Interferometer, InterferometerList, PowerSpectralDensity,  # noqa: F401
                       aligo_analytic_psd, get_empty_interferometer)
from .waveform_generator import WaveformGenerator, sinegaussian, taylorf2  # noqa: F401


def __getattr__(name):
    # the likelihood needs the HIP library; import it lazily so that the host
    # setup utilities stay importable without it
    if name in ("GravitationalWaveTransient", "inject_signal", "likelihood"):
        from . import likelihood
        return likelihood if name == "likelihood" else getattr(likelihood, name)
    if name in ("ROQGravitationalWaveTransient", "roq"):
        from . import roq
        return roq if name == "roq" else getattr(roq, name)
    raise AttributeError(name)
