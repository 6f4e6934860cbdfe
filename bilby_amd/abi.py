"""Constants of the libgwl C ABI (include/gwl.h), importable without the
library, and the host-side conversion of parameter dicts to its rows."""
import numpy as np

NPARAM = 16
MAX_DET = 4

# enum gwl_param
P_RA, P_DEC, P_PSI, P_GEOCENT_TIME, P_TIME_JITTER, P_LUMINOSITY_DISTANCE, P_THETA_JN, \
    P_PHASE, P_MASS_1, P_MASS_2, P_CHI_1, P_CHI_2, P_HRSS, P_Q, P_FREQUENCY = range(15)
PARAM_COLUMNS = {
    "ra": P_RA, "dec": P_DEC, "psi": P_PSI, "geocent_time": P_GEOCENT_TIME,
    "time_jitter": P_TIME_JITTER, "luminosity_distance": P_LUMINOSITY_DISTANCE,
    "theta_jn": P_THETA_JN, "phase": P_PHASE, "mass_1": P_MASS_1, "mass_2": P_MASS_2,
    "chi_1": P_CHI_1, "chi_2": P_CHI_2, "hrss": P_HRSS, "Q": P_Q, "frequency": P_FREQUENCY,
}
SINE_GAUSSIAN, TAYLORF2, HOST_POLARIZATIONS = 1, 2, 3
TIME_MARG, DISTANCE_MARG, PHASE_MARG, JITTER_TIME = 1, 2, 4, 8
# TaylorF2 ends at the ISCO as well as at maximum_frequency (opt-in; see
# waveform_generator.taylorf2_flags)
TF2_END_ISCO = 16
FP64, FP32 = 0, 1

# enum gwl_transform (time-marginalised TaylorF2 transform choice)
TRANSFORM_AUTO, TRANSFORM_TILED, TRANSFORM_WAVE = 0, 1, 2

# gwl_server_listen wire magics
WIRE_REQUEST, WIRE_REPLY = 0x51574C47, 0x52574C47


class RowBuilder:
    """Parameter dict (scalars or 1-D arrays, bilby names) -> float64 rows
    [n, NPARAM] in the column order of enum gwl_param: the waveform
    generator's parameter conversion (waveform_generator.py:260-269), then
    the defaults of the marginalised keys, then the likelihood's sky frame
    (``frame``, a bilby_amd.frames.SkyFrame: detector-frame zenith / azimuth
    and <ifo>_time to ra / dec / geocent_time, base.py:1091-1137), then the
    columns.  Host-only and picklable, so a sampler's worker process can
    build rows itself."""

    def __init__(self, parameter_conversion, defaults, frame=None):
        self.parameter_conversion = parameter_conversion
        self.defaults = dict(defaults)
        self.frame = None if frame is None or frame.is_identity else frame

    def __call__(self, parameters):
        p = self.parameter_conversion(dict(parameters))
        for key, value in self.defaults.items():
            p.setdefault(key, value)
        if self.frame is not None:
            p = self.frame(p)
        # Python / numpy float64 scalars skip numpy's array dispatch: one
        # sampler dict is converted per call on the per-sample path
        n = 1
        for v in p.values():
            if isinstance(v, (float, int)):
                continue
            if np.ndim(v) > 0:
                n = len(v)
                break
        rows = np.zeros((n, NPARAM))
        for key, col in PARAM_COLUMNS.items():
            if key in p:
                v = p[key]
                rows[:, col] = v if isinstance(v, (float, int)) else np.asarray(v, dtype=float)
        return rows
