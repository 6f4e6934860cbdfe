"""Priors the likelihood reads (setup only): the marginalised keys need
``minimum``, ``maximum``, ``prob`` and ``rescale`` (bilby/core/prior).

UniformComovingVolume uses an in-package flat Lambda-CDM cosmology
(H0 = 67.74 km/s/Mpc, Omega_m = 0.3075, no radiation), because astropy is not
available; bilby's default Planck15 (astropy) differs slightly in the
distance-redshift relation, so the prior *shape* is not identical to bilby's.
The likelihood only consumes the resulting prior array.
"""
import numpy as np


class Prior:
    def __init__(self, minimum=-np.inf, maximum=np.inf, name=None, latex_label=None, unit=None,
                 boundary=None):
        self.minimum = minimum
        self.maximum = maximum
        self.name = name
        self.latex_label = latex_label
        self.unit = unit
        self.boundary = boundary

    @property
    def is_fixed(self):
        return False

    def sample(self, size=None, rng=None):
        rng = rng if rng is not None else np.random.default_rng()
        return self.rescale(rng.uniform(0, 1, size))


class Uniform(Prior):
    """bilby/core/prior/analytical.py:191-242."""

    def rescale(self, val):
        return self.minimum + np.asarray(val) * (self.maximum - self.minimum)

    def prob(self, val):
        val = np.asarray(val)
        return ((val >= self.minimum) & (val <= self.maximum)) / (self.maximum - self.minimum)


class Gaussian(Prior):
    """bilby/core/prior/analytical.py Gaussian: rescale mu + erfinv(2u - 1)
    sqrt(2) sigma."""

    def __init__(self, mu, sigma, name=None, boundary=None, **kw):
        super().__init__(name=name, boundary=boundary)
        self.mu = mu
        self.sigma = sigma

    def rescale(self, val):
        from scipy.special import erfinv
        return self.mu + erfinv(2 * np.asarray(val, float) - 1) * 2 ** 0.5 * self.sigma

    def prob(self, val):
        val = np.asarray(val, float)
        return np.exp(-(self.mu - val) ** 2 / (2 * self.sigma ** 2)) / (2 * np.pi) ** 0.5 / self.sigma


class TruncatedGaussian(Prior):
    """bilby/core/prior/analytical.py TruncatedGaussian: a Gaussian density
    restricted to [minimum, maximum] and renormalised."""

    def __init__(self, mu, sigma, minimum, maximum, name=None, boundary=None, **kw):
        super().__init__(minimum=minimum, maximum=maximum, name=name, boundary=boundary)
        self.mu = mu
        self.sigma = sigma

    @property
    def normalisation(self):
        from scipy.special import erf
        return (erf((self.maximum - self.mu) / 2 ** 0.5 / self.sigma)
                - erf((self.minimum - self.mu) / 2 ** 0.5 / self.sigma)) / 2

    def rescale(self, val):
        from scipy.special import erf, erfinv
        lo = erf((self.minimum - self.mu) / 2 ** 0.5 / self.sigma)
        return self.mu + erfinv(2 * np.asarray(val, float) * self.normalisation + lo) \
            * 2 ** 0.5 * self.sigma

    def prob(self, val):
        val = np.asarray(val, float)
        inside = (val >= self.minimum) & (val <= self.maximum)
        return (np.exp(-(self.mu - val) ** 2 / (2 * self.sigma ** 2)) / (2 * np.pi) ** 0.5
                / self.sigma / self.normalisation * inside)


class DeltaFunction(Prior):
    def __init__(self, peak, name=None, **kw):
        super().__init__(minimum=peak, maximum=peak, name=name)
        self.peak = peak

    @property
    def is_fixed(self):
        return True

    def rescale(self, val):
        return self.peak * np.ones_like(np.asarray(val, float))

    def prob(self, val):
        return (np.asarray(val) == self.peak) * 1.0


class PowerLaw(Prior):
    """p(x) ~ x^alpha on [minimum, maximum] (bilby/core/prior/analytical.py)."""

    def __init__(self, alpha, minimum, maximum, name=None, **kw):
        super().__init__(minimum=minimum, maximum=maximum, name=name)
        self.alpha = alpha

    def rescale(self, val):
        val = np.asarray(val, float)
        if self.alpha == -1:
            return self.minimum * np.exp(val * np.log(self.maximum / self.minimum))
        a1 = self.alpha + 1
        return (self.minimum ** a1 + val * (self.maximum ** a1 - self.minimum ** a1)) ** (1 / a1)

    def prob(self, val):
        val = np.asarray(val, float)
        if self.alpha == -1:
            norm = 1 / np.log(self.maximum / self.minimum)
        else:
            a1 = self.alpha + 1
            norm = a1 / (self.maximum ** a1 - self.minimum ** a1)
        inside = (val >= self.minimum) & (val <= self.maximum)
        return np.where(inside, norm * np.abs(val) ** self.alpha, 0.0)


def _trapezoid(y, x):
    return float(np.sum((y[1:] + y[:-1]) * np.diff(x)) / 2)


class Interped(Prior):
    """bilby/core/prior/interpolated.py:10-177: linear interpolation of a
    tabulated density, re-gridded on linspace(minimum, maximum, len(xx)),
    normalised by the trapezoid rule; rescale by the inverse of the
    cumulative trapezoid."""

    def __init__(self, xx, yy, minimum=np.nan, maximum=np.nan, name=None, **kw):
        xx = np.asarray(xx, float)
        yy = np.asarray(yy, float)
        minimum = float(np.nanmax(np.array((xx.min(), minimum))))
        maximum = float(np.nanmin(np.array((xx.max(), maximum))))
        super().__init__(minimum=minimum, maximum=maximum, name=name)
        self._xx0, self._yy0 = xx, yy
        self.xx = np.linspace(minimum, maximum, len(xx))
        self.yy = self._interp_all(self.xx)
        self.yy = self.yy / _trapezoid(self.yy, self.xx)
        ydx = (self.yy[1:] + self.yy[:-1]) * np.diff(self.xx) / 2
        self.YY = np.concatenate([[0.0], np.cumsum(ydx)])
        self.YY[-1] = 1

    def _interp_all(self, x):
        out = np.interp(x, self._xx0, self._yy0)
        out[(x < self._xx0[0]) | (x > self._xx0[-1])] = 0.0
        return out

    def prob(self, val):
        val = np.asarray(val, float)
        out = np.interp(val, self.xx, self.yy)
        out = np.where((val < self.xx[0]) | (val > self.xx[-1]), 0.0, out)
        return out[()] if out.ndim == 0 else out

    def cdf(self, val):
        return np.interp(val, self.xx, self.YY, left=0.0, right=1.0)

    def rescale(self, val):
        val = np.asarray(val, float)
        if np.any((val < 0) | (val > 1)):
            raise ValueError("rescale input outside [0, 1]")
        out = np.interp(val, self.YY, self.xx)
        return out[()] if out.ndim == 0 else out


class FlatLambdaCDM:
    """Flat Lambda-CDM distances (Mpc); radiation neglected."""

    def __init__(self, H0=67.74, Om0=0.3075):
        self.H0 = H0
        self.Om0 = Om0
        self.hubble_distance = 299792.458 / H0
        z = np.linspace(0.0, 3.0, 30001)
        ez = np.sqrt(Om0 * (1 + z) ** 3 + (1 - Om0))
        dc = np.concatenate([[0.0], np.cumsum((1 / ez[1:] + 1 / ez[:-1]) / 2 * np.diff(z))])
        self._z = z
        self._dc = dc * self.hubble_distance

    def comoving_distance(self, z):
        return np.interp(z, self._z, self._dc)

    def luminosity_distance(self, z):
        return (1 + np.asarray(z)) * self.comoving_distance(z)

    def z_at_luminosity_distance(self, dl):
        return np.interp(dl, (1 + self._z) * self._dc, self._z)

    def differential_comoving_volume(self, z):
        z = np.asarray(z, float)
        ez = np.sqrt(self.Om0 * (1 + z) ** 3 + (1 - self.Om0))
        return self.hubble_distance * self.comoving_distance(z) ** 2 / ez


class Cosmological(Interped):
    """bilby/gw/prior.py:97-300: a prior on redshift, comoving distance or
    luminosity distance (``name``) defined through a redshift density p(z)
    on [0.99 z_min, 1.01 z_max] (1000 points) and the cosmology's distance
    relations, p(x) = p(z) / (dx/dz) (prior.py:277-290);
    ``get_corresponding_prior`` gives the same prior on another of the three
    (prior.py:224-249), as the likelihood's distance-marginalisation setup
    uses it (base.py:370-380)."""

    NAMES = ("redshift", "comoving_distance", "luminosity_distance")

    def __init__(self, minimum, maximum, name="luminosity_distance", cosmology=None, **kw):
        if name not in self.NAMES:
            raise ValueError("Name {} not recognized.".format(name))
        self.cosmology = cosmology or FlatLambdaCDM()
        self._zmin, self._zmax = self._to_redshift(minimum, name), self._to_redshift(maximum, name)
        zs = np.linspace(self._zmin * 0.99, self._zmax * 1.01, 1000)
        p_dz = self._redshift_density(zs)
        if name == "redshift":
            xx, yy = zs, p_dz
        else:
            xx = (self.cosmology.comoving_distance(zs) if name == "comoving_distance"
                  else self.cosmology.luminosity_distance(zs))
            yy = p_dz / np.gradient(xx, zs)
        super().__init__(xx, yy, minimum=minimum, maximum=maximum, name=name)

    def _to_redshift(self, value, name):
        c = self.cosmology
        if name == "redshift" or value == 0:
            return float(value)
        if name == "luminosity_distance":
            return float(c.z_at_luminosity_distance(value))
        return float(np.interp(value, c._dc, c._z))

    def _from_redshift(self, z, name):
        if name == "redshift":
            return float(z)
        if name == "comoving_distance":
            return float(self.cosmology.comoving_distance(z))
        return float(self.cosmology.luminosity_distance(z))

    def _redshift_density(self, zs):
        raise NotImplementedError

    def get_corresponding_prior(self, name=None, unit=None):
        return type(self)(self._from_redshift(self._zmin, name),
                          self._from_redshift(self._zmax, name), name=name,
                          cosmology=self.cosmology)


class UniformComovingVolume(Cosmological):
    """bilby/gw/prior.py:310-326: p(z) ~ dVc/dz."""

    def _redshift_density(self, zs):
        return self.cosmology.differential_comoving_volume(zs)


class Constraint(Prior):
    """bilby/core/prior/base.py:495-504: the indicator minimum < value < maximum
    (a constraint on a converted parameter, never sampled)."""

    def __init__(self, minimum, maximum, name=None, **kw):
        super().__init__(minimum=minimum, maximum=maximum, name=name)

    @property
    def is_fixed(self):
        return True

    def prob(self, val):
        val = np.asarray(val)
        return (val > self.minimum) & (val < self.maximum)


def _no_conversion(sample):
    """bilby/core/prior/dict.py:75-89 (default_conversion_function)."""
    return sample


class PriorDict(dict):
    """Subset of bilby/core/prior/dict.py: priors by name, an optional
    conversion function for constraints, ``sample``, ``rescale`` and
    ``evaluate_constraints`` (what the samplers' likelihood wrappers call)."""

    def __init__(self, dictionary=None, conversion_function=None, **priors):
        super().__init__()
        if dictionary:
            self.update(dictionary)
        self.update(priors)
        self.conversion_function = conversion_function or _no_conversion

    def copy(self):
        return PriorDict(self, conversion_function=self.conversion_function)

    def sample(self, size, rng=None):
        rng = rng if rng is not None else np.random.default_rng()
        out = {}
        for k, p in self.items():
            if isinstance(p, Constraint):
                continue
            if isinstance(p, Prior):
                out[k] = p.sample(size, rng)
            else:
                out[k] = np.full(size, float(p))
        return out

    def rescale(self, keys, theta):
        """dict.py:647-666: unit-cube values -> prior values, key by key."""
        return np.asarray([self[key].rescale(t) for key, t in zip(keys, list(theta))])

    def evaluate_constraints(self, sample):
        """dict.py:64-73: product of every Constraint's indicator over the
        converted sample (scalars or arrays); True where no constraint applies."""
        out = self.conversion_function(dict(sample))
        prob = np.ones_like(np.asarray(next(iter(out.values())), dtype=float), dtype=bool)
        for key, p in self.items():
            if isinstance(p, Constraint) and key in out:
                prob = prob & p.prob(out[key])
        return prob


class CBCPriorDict(PriorDict):
    """The mass bounds of bilby/gw/prior.py:721-782 that the ROQ checks read
    (perform_roq_params_check); None where the priors do not fix them."""

    def copy(self):
        return CBCPriorDict(self, conversion_function=self.conversion_function)

    def _chirp_mass_bound(self, which):
        if "chirp_mass" in self:
            return getattr(self["chirp_mass"], which)
        m1 = m2 = None
        if "mass_1" in self:
            m1 = getattr(self["mass_1"], which)
            if "mass_2" in self:
                m2 = getattr(self["mass_2"], which)
            elif "mass_ratio" in self:
                m2 = m1 * getattr(self["mass_ratio"], which)
        if m1 is None or m2 is None:
            return None
        return (m1 * m2) ** 0.6 / (m1 + m2) ** 0.2

    @property
    def minimum_chirp_mass(self):
        return self._chirp_mass_bound("minimum")

    @property
    def maximum_chirp_mass(self):
        return self._chirp_mass_bound("maximum")

    @property
    def minimum_component_mass(self):
        """mass_2's minimum, or the lighter mass at the chirp-mass and
        mass-ratio minima."""
        if "mass_2" in self:
            return self["mass_2"].minimum
        if "chirp_mass" in self and "mass_ratio" in self:
            mc, q = self["chirp_mass"].minimum, self["mass_ratio"].minimum
            total = mc * (1 + q) ** 1.2 / q ** 0.6
            return total * q / (1 + q)
        return None
