"""Host logic of the ROQ basis management (bilby/gw/likelihood/roq.py:86-145,
192-280, 402-435, 653-734, 800-838): roq_params validation, the ROQ
frequency grid intersected with the masked frequencies (against the
oracle's restatement), the CBC prior mass bounds, and the basis choice by
prior range."""
import numpy as np
import pytest

from bilby_amd.detector import InterferometerList
from bilby_amd.prior import CBCPriorDict, Uniform
from bilby_amd.roq import (BilbyROQParamsRangeError, ROQGravitationalWaveTransient,
                           _check_roq_params, _params_from_rows)
from bilby_amd.roq_weights import roq_frequency_indices
from oracle import roq_weights as orw


def _params(flow=10.0, fhigh=256.0, seglen=8.0, **extra):
    names = ["flow", "fhigh", "seglen"] + list(extra)
    return np.array(tuple([flow, fhigh, seglen] + list(extra.values())),
                    dtype=[(n, float) for n in names])


def test_roq_params_validation(tmp_path):
    assert _check_roq_params(None) is None
    p = _params(chirpmassmin=5.0)
    assert _check_roq_params(p) is p
    with pytest.raises(TypeError):
        _check_roq_params(np.zeros(3))
    with pytest.raises(ValueError):
        _check_roq_params(np.array([(1.0, 2.0, 3.0)], dtype=[("flow", float), ("fhigh", float),
                                                              ("seglen", float)]))
    with pytest.raises(ValueError, match="seglen"):
        _check_roq_params(np.array((1.0, 2.0), dtype=[("flow", float), ("fhigh", float)]))
    # the file form bilby reads with genfromtxt(names=True): a header and one row
    f = tmp_path / "params.dat"
    f.write_text("flow fhigh seglen chirpmassmin chirpmassmax compmin\n20 1024 4 1.4 2.6 0.5\n")
    q = _check_roq_params(np.genfromtxt(str(f), names=True))
    assert q.ndim == 0 and float(q["fhigh"]) == 1024.0 and float(q["compmin"]) == 0.5


@pytest.mark.parametrize("scale", [1.0, 2.0])
def test_roq_frequency_grid_intersection_matches_oracle(scale):
    """The ROQ grid (2 fhigh s, seglen / s, from flow s) against each
    detector's masked frequencies: same indices as the oracle; every common
    frequency found, each pair equal."""
    ifos = InterferometerList(["H1", "L1"])
    ifos.set_strain_data_from_zero_noise(1024.0, 4.0, 1e9)
    ifos[1].maximum_frequency = 300.0
    p = _params(flow=8.0 * 2 / scale, fhigh=512.0 / scale, seglen=4.0 * scale)
    roq_idxs, ifo_idxs = roq_frequency_indices(ifos, p, scale)
    for ifo in ifos:
        r, i = orw.roq_indices(ifo.frequency_array, ifo.frequency_mask, float(p["flow"]),
                               float(p["fhigh"]), float(p["seglen"]), scale)
        assert np.array_equal(roq_idxs[ifo.name], r) and np.array_equal(ifo_idxs[ifo.name], i)
        f_ifo = ifo.frequency_array[ifo.frequency_mask]
        f_roq = 16.0 + r * 0.25  # grid from flow s = 16 Hz, spacing 1 / (4 s)
        assert np.array_equal(f_ifo[i], f_roq)
        assert len(i) == int(np.sum(ifo.frequency_mask))  # the grid covers the whole band


def test_cbc_prior_mass_bounds():
    """gw/prior.py:721-782: chirp-mass bounds from a chirp-mass prior or from
    component masses, the lighter component at the chirp-mass and
    mass-ratio minima (m2 = M q / (1 + q), M = Mc (1 + q)^(6/5) / q^(3/5))."""
    p = CBCPriorDict(chirp_mass=Uniform(10, 30), mass_ratio=Uniform(0.25, 1))
    assert p.minimum_chirp_mass == 10 and p.maximum_chirp_mass == 30
    # the defining identity: the pair (m2 / q, m2) has chirp mass 10
    m2 = p.minimum_component_mass
    m1 = m2 / 0.25
    assert abs((m1 * m2) ** 0.6 / (m1 + m2) ** 0.2 - 10.0) < 1e-12
    p2 = CBCPriorDict(mass_1=Uniform(5, 50), mass_2=Uniform(5, 50))
    assert abs(p2.minimum_chirp_mass - 5 * 2 ** -0.2) < 1e-12
    assert abs(p2.maximum_chirp_mass - 50 * 2 ** -0.2) < 1e-12
    assert p2.minimum_component_mass == 5
    assert CBCPriorDict(a_1=Uniform(0, 1)).minimum_chirp_mass is None
    assert isinstance(p2.copy(), CBCPriorDict)


def _bare_roq(n_lin, n_quad, ranges_lin, ranges_quad):
    """An ROQ likelihood object with only the basis bookkeeping set (no GPU)."""
    lk = ROQGravitationalWaveTransient.__new__(ROQGravitationalWaveTransient)
    lk.number_of_bases_linear, lk.number_of_bases_quadratic = n_lin, n_quad
    lk.weights = {"prior_range_linear": ranges_lin, "prior_range_quadratic": ranges_quad}
    lk.parameter_conversion = None
    return lk


def test_basis_choice_by_prior_range():
    """roq.py:418-427: the first basis whose ranges hold the parameters;
    parameters the sample lacks are not checked; none -> error."""
    lk = _bare_roq(3, 2,
                   {"chirp_mass": np.array([[4.0, 20.0], [18.0, 40.0], [35.0, 90.0]]),
                    "mass_ratio": np.array([[0.0, 1.0], [0.0, 1.0], [0.5, 1.0]])},
                   {"chirp_mass": np.array([[4.0, 30.0], [25.0, 90.0]])})
    mc = np.array([5.0, 19.0, 20.0, 21.0, 36.0, 36.0, 60.0])
    q = np.array([0.3, 0.3, 0.3, 0.3, 0.3, 0.9, 0.9])
    bl, bq = lk._basis_numbers({"chirp_mass": mc, "mass_ratio": q}, len(mc))
    assert list(bl) == [0, 0, 0, 1, 1, 1, 2]
    assert list(bq) == [0, 0, 0, 0, 1, 1, 1]
    # without mass_ratio in the sample only the chirp mass decides
    bl2, _ = lk._basis_numbers({"chirp_mass": mc}, len(mc))
    assert list(bl2) == [0, 0, 0, 1, 1, 1, 2]
    with pytest.raises(BilbyROQParamsRangeError):
        lk._basis_numbers({"chirp_mass": np.array([100.0])}, 1)
    lk._cache = {}
    assert lk._update_basis({"chirp_mass": 36.0, "mass_ratio": 0.9}) == (1, 1)
    assert lk.basis_number_linear == 1 and lk.basis_number_quadratic == 1


def test_params_from_rows():
    rows = np.zeros((2, 16))
    rows[:, 8] = [30.0, 10.0]
    rows[:, 9] = [20.0, 10.0]
    p = _params_from_rows(rows)
    assert np.allclose(p["chirp_mass"], [(600.0 ** 0.6) / 50 ** 0.2, 10 * 2 ** -0.2])
    assert np.allclose(p["mass_ratio"], [2 / 3, 1.0]) and np.allclose(p["total_mass"], [50, 20])
