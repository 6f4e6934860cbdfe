"""Constants of the libgwl C ABI (include/gwl.h), importable without the
library."""
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
SINE_GAUSSIAN, TAYLORF2 = 1, 2
TIME_MARG, DISTANCE_MARG, PHASE_MARG, JITTER_TIME = 1, 2, 4, 8
FP64, FP32 = 0, 1

# enum gwl_transform (time-marginalised TaylorF2 transform choice)
TRANSFORM_AUTO, TRANSFORM_TILED, TRANSFORM_WAVE, TRANSFORM_CHIRP = 0, 1, 2, 3
