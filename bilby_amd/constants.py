"""Physical constants, identical to bilby/core/utils/constants.py:3-7."""
speed_of_light = 299792458.0
parsec = 3.085677581491367e16
solar_mass = 1.988409870698050731911960804878414216e30
radius_of_earth = 6378136.6
gravitational_constant = 6.6743e-11
