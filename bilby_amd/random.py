"""Package-wide random generator, as bilby/core/utils/random.py:60-143:
``seed(s)`` replaces the global numpy Generator that noise realisations use
when no ``random_state`` is given."""
import numpy as np

rng = np.random.default_rng()


def seed(value):
    global rng
    rng = np.random.default_rng(value)


def resolve_random_state(random_state):
    """random.py:113-143: None -> the global generator, int -> a new
    default_rng(int), Generator -> itself."""
    if isinstance(random_state, np.random.Generator):
        return random_state
    if random_state is None:
        return rng
    if isinstance(random_state, (int, np.integer)):
        return np.random.default_rng(int(random_state))
    raise ValueError("Invalid random state. Must be None, int, or np.random.Generator.")
