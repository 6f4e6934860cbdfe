"""Host parameter conversion (bilby_amd.waveform_generator, restating
bilby/gw/conversion.py:182-283, 1826-1975) and the marginalised-prior
defaults (base.py:356-390), on the CPU:
* every complete set of mass parameters gives back the component masses it
  was made from (1e-12 relative);
* redshift / comoving distance -> luminosity distance, *_source masses;
* spins: chi_i, a_i with tilt_i or cos_tilt_i, in-plane spins refused
  (TaylorF2 is non-precessing); cos_theta_jn; delta_phase;
* Cosmological priors carried between redshift, comoving and luminosity
  distance; the distance and phase prior defaults of the likelihood."""
import types

import numpy as np
import pytest

from bilby_amd.waveform_generator import (comoving_distance_to_luminosity_distance,
                                          convert_to_lal_binary_black_hole_parameters as conv,
                                          generate_component_masses,
                                          luminosity_distance_to_redshift,
                                          redshift_to_luminosity_distance)

M1 = np.array([30.0, 12.0, 50.0, 5.2])
M2 = np.array([20.0, 11.0, 5.0, 5.1])


def _derived():
    M = M1 + M2
    return dict(chirp_mass=(M1 * M2) ** 0.6 / M ** 0.2, mass_ratio=M2 / M1, total_mass=M,
                symmetric_mass_ratio=M1 * M2 / M ** 2)


@pytest.mark.parametrize("keys", [("chirp_mass", "mass_ratio"), ("total_mass", "mass_ratio"),
                                  ("chirp_mass", "total_mass"), ("mass_1", "chirp_mass"),
                                  ("mass_2", "chirp_mass"), ("mass_1", "symmetric_mass_ratio"),
                                  ("mass_2", "symmetric_mass_ratio"), ("mass_1", "total_mass"),
                                  ("mass_2", "total_mass"), ("chirp_mass", "symmetric_mass_ratio"),
                                  ("total_mass", "symmetric_mass_ratio"), ("mass_1", "mass_ratio"),
                                  ("mass_2", "mass_ratio")])
def test_every_mass_set_gives_the_component_masses(keys):
    d = dict(_derived(), mass_1=M1, mass_2=M2)
    out = generate_component_masses({k: d[k] for k in keys})
    assert np.allclose(out["mass_1"], M1, rtol=1e-12, atol=0)
    assert np.allclose(out["mass_2"], M2, rtol=1e-12, atol=0)


def test_incomplete_mass_set():
    assert "mass_1" not in generate_component_masses({"chirp_mass": 10.0})
    with pytest.raises(KeyError):
        generate_component_masses({"chirp_mass": 10.0}, require_add=True)


def test_distances_and_source_frame_masses():
    z = np.array([0.05, 0.2, 0.5])
    dl = redshift_to_luminosity_distance(z)
    assert np.allclose(luminosity_distance_to_redshift(dl), z, rtol=1e-6)
    from bilby_amd.prior import FlatLambdaCDM
    dc = FlatLambdaCDM().comoving_distance(z)
    assert np.allclose(comoving_distance_to_luminosity_distance(dc), (1 + z) * dc, rtol=1e-6)
    p = conv(dict(redshift=z, chirp_mass_source=np.full(3, 10.0), mass_ratio=np.full(3, 0.5),
                  theta_jn=0.3, phase=0.1))
    assert np.allclose(p["luminosity_distance"], dl)
    assert np.allclose(p["chirp_mass"], 10.0 * (1 + z))
    p2 = conv(dict(luminosity_distance=dl, mass_1_source=np.full(3, 30.0),
                   mass_2_source=np.full(3, 20.0), theta_jn=0.3, phase=0.1))
    assert np.allclose(p2["mass_1"], 30.0 * (1 + z), rtol=1e-6)
    assert np.allclose(p2["mass_2"], 20.0 * (1 + z), rtol=1e-6)


def test_spins_angles_and_delta_phase():
    base = dict(mass_1=30.0, mass_2=20.0, luminosity_distance=400.0, psi=0.7)
    p = conv(dict(base, chi_1=-0.4, chi_2=0.3, theta_jn=0.5, phase=1.0))
    assert p["chi_1"] == -0.4 and p["a_1"] == 0.4 and p["cos_tilt_1"] == -1.0
    p = conv(dict(base, a_1=0.5, tilt_1=np.pi, a_2=0.2, cos_tilt_2=1.0, cos_theta_jn=0.6,
                  phase=1.0))
    assert abs(p["chi_1"] + 0.5) < 1e-15 and abs(p["chi_2"] - 0.2) < 1e-15
    assert abs(p["theta_jn"] - np.arccos(0.6)) < 1e-15
    with pytest.raises(ValueError):
        conv(dict(base, a_1=0.5, tilt_1=0.3, theta_jn=0.5, phase=1.0))
    with pytest.raises(ValueError):
        conv(dict(base, chi_1=0.2, chi_1_in_plane=0.1, theta_jn=0.5, phase=1.0))
    # delta_phase = phase + sign(cos theta_jn) psi (mod 2 pi) round trip
    for theta_jn in (0.4, 2.5):
        phase = 5.9
        dphase = np.mod(phase + np.sign(np.cos(theta_jn)) * base["psi"], 2 * np.pi)
        p = conv(dict(base, chi_1=0.0, chi_2=0.0, theta_jn=theta_jn, delta_phase=dphase))
        assert abs(p["phase"] - phase) < 1e-12


def test_cosmological_prior_between_redshift_and_distances():
    from bilby_amd.prior import UniformComovingVolume
    pz = UniformComovingVolume(0.02, 0.4, name="redshift")
    pdl = pz.get_corresponding_prior("luminosity_distance")
    direct = UniformComovingVolume(pdl.minimum, pdl.maximum, name="luminosity_distance")
    d = np.linspace(pdl.minimum, pdl.maximum, 200)
    assert np.allclose(pdl.prob(d), direct.prob(d), rtol=1e-6)
    # p(d_L) dd_L = p(z) dz between the two priors
    z = np.linspace(0.05, 0.35, 50)
    dl = redshift_to_luminosity_distance(z)
    dldz = (redshift_to_luminosity_distance(z + 1e-4)
            - redshift_to_luminosity_distance(z - 1e-4)) / 2e-4
    assert np.allclose(pdl.prob(dl) * dldz, pz.prob(z), rtol=2e-3)
    pdc = pz.get_corresponding_prior("comoving_distance")
    assert pdc.name == "comoving_distance" and pdc.maximum < pdl.maximum


def test_marginalised_prior_defaults():
    from bilby_amd.likelihood import GravitationalWaveTransient as GWT
    from bilby_amd.prior import PriorDict, Uniform, UniformComovingVolume
    ifos = types.SimpleNamespace(start_time=100.0, duration=8.0)
    stub = types.SimpleNamespace(priors=PriorDict(redshift=UniformComovingVolume(
        0.02, 0.4, name="redshift")), interferometers=ifos)
    GWT._check_marginalized_prior_is_set(stub, "luminosity_distance")
    assert "redshift" not in stub.priors
    assert stub.priors["luminosity_distance"].name == "luminosity_distance"
    GWT._check_marginalized_prior_is_set(stub, "phase")
    assert (stub.priors["phase"].minimum, stub.priors["phase"].maximum) == (0.0, 2 * np.pi)
    GWT._check_marginalized_prior_is_set(stub, "geocent_time")
    assert (stub.priors["geocent_time"].minimum, stub.priors["geocent_time"].maximum) == (100, 108)
    bad = types.SimpleNamespace(priors=PriorDict(redshift=Uniform(0.0, 1.0)), interferometers=ifos)
    with pytest.raises(TypeError):
        GWT._check_marginalized_prior_is_set(bad, "luminosity_distance")
    none = types.SimpleNamespace(priors=PriorDict(), interferometers=ifos)
    with pytest.raises(ValueError):
        GWT._check_marginalized_prior_is_set(none, "luminosity_distance")
