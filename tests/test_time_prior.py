"""The geocent_time prior table (bilby_amd.likelihood.time_prior_table) that
the kernels interpolate for priors other than Uniform (base.py:794-820
weighs each time by prior.prob(t) delta_tc): nodes on the support's ends,
densities equal to prob at the nodes, linear interpolation within 1e-6 of
prob between them, tails where the density vanishes trimmed, Uniform priors
left to the device's constant weight."""
import numpy as np
import pytest

from bilby_amd.likelihood import time_prior_table
from bilby_amd.prior import Gaussian, TruncatedGaussian, Uniform

START, T, DTC = 1126259600.0, 8.0, 2.0 / 4096


def test_uniform_has_no_table():
    lo, hi, table = time_prior_table(Uniform(START + 3.9, START + 4.1), START, T, DTC)
    assert (lo, hi, table) == (START + 3.9, START + 4.1, None)


def test_truncated_gaussian_keeps_its_hard_edges_as_nodes():
    p = TruncatedGaussian(START + 4.0, 0.03, START + 3.95, START + 4.12)
    lo, hi, (t0, dt, dens) = time_prior_table(p, START, T, DTC)
    assert lo == p.minimum and hi == p.maximum and t0 == lo
    # power-of-two spacing: every node lo + k dt is an exact float64
    assert dt <= DTC / 64 and 2 * dt > DTC / 64 and np.log2(dt) == np.round(np.log2(dt))
    k = np.arange(len(dens))
    grid = t0 + dt * k
    assert np.all((grid - t0) / dt == k)
    assert grid[-2] < hi <= grid[-1]
    assert np.allclose(dens, p.prob(np.minimum(grid, hi)), rtol=1e-12, atol=0)
    # linear interpolation between nodes vs the density itself
    t = np.random.default_rng(2).uniform(lo, hi, 5000)
    interp = np.interp(t, grid, dens)
    assert np.max(np.abs(interp - p.prob(t)) / p.prob(t).max()) < 1e-6


def test_gaussian_tails_are_trimmed_to_the_non_zero_density():
    p = Gaussian(START + 4.0, 0.01)
    lo, hi, (t0, dt, dens) = time_prior_table(p, START, T, DTC)
    assert START < lo < START + 4.0 - 0.3 and START + 4.0 + 0.3 < hi < START + T
    assert p.prob(lo - DTC) == 0.0 or p.prob(lo) < 1e-300
    assert np.all(dens >= 0) and dens.max() == pytest.approx(p.prob(START + 4.0), rel=1e-3)


def test_prior_outside_the_segment_is_refused():
    with pytest.raises(ValueError):
        time_prior_table(TruncatedGaussian(START - 5, 0.1, START - 6, START - 4), START, T, DTC)
