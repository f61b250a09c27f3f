"""
Polydispersity distributions: the points and weights of one parameter.

A distribution is a *law* with three parts:

* placement -- where its candidate points go for a parameter of value c and
  width s ("span": c + linspace(-n s, +n s, npts), n = nsigmas; "edges":
  linspace(c - s, c + s, npts));
* window -- which candidates it keeps besides the parameter limits (the
  rectangle keeps |x - c| <= sqrt(3) |s|); the lognormal and Schulz laws are
  defined for x > 0 only, so their limits are raised to at least 1e-8;
* density -- the unnormalised weight of each kept point.

The laws are the rows of ``_LAWS``; the classes SasView selects by name
(``DISTRIBUTIONS``, also bound under its older name ``MODELS``) only hold
npts, width and nsigmas and apply their law.  Widths of volume parameters
are relative to the value; orientation jitter has an absolute width and is
centred on zero, whatever the view angle.  The placement and the floating
point order of each density follow sasmodels/weights.py:86-176, so points
and weights agree with the reference's to the last bit; ``get_weights``
normalises the weights to sum one (weights.py:264-291).
"""

from collections import OrderedDict, namedtuple
from math import sqrt

import numpy as np
from scipy.special import gammaln

#: one distribution law (see the module docstring)
Law = namedtuple("Law", "name npts nsigmas placement positive window density")


def _gaussian(x, c, s):
    return np.exp((x - c)**2 / (-2.0 * s * s))


def _lognormal(x, c, s):
    # the width of a lognormal is given on the value; in log space it is s/c
    t = np.fabs(s / c)
    return np.exp(-0.5 * ((np.log(x) - np.log(c)) / t)**2) / (x * t)


def _schulz(x, c, s):
    # z^z R^(z-1) e^(-R z) / (c Gamma(z)) with R = x/c, z = (c/s)^2, in logs
    R = x / c
    z = (c / s)**2
    return np.exp(z * np.log(z) + (z - 1) * np.log(R) - R * z - np.log(c) - gammaln(z))


def _flat(x, c, s):
    return np.ones_like(x)


def _laplace(x, c, s):
    return np.exp(-np.abs(x - c) / np.abs(s))


def _within_sqrt3(x, c, s):
    return np.fabs(x - c) <= np.fabs(s) * sqrt(3.0)


#: name -> law, in the order SasView lists the distributions (weights.py:221-229)
_LAWS = OrderedDict((law.name, law) for law in (
    Law("rectangle", 35, 1.73205, "span", False, _within_sqrt3, _flat),
    Law("uniform", 35, None, "edges", False, None, _flat),
    Law("array", 35, 1, None, False, None, None),
    Law("lognormal", 80, 8, "span", True, None, _lognormal),
    Law("gaussian", 35, 3, "span", False, None, _gaussian),
    Law("schulz", 80, 8, "span", True, None, _schulz),
    Law("boltzmann", 35, 3, "span", False, None, _laplace),
))


def sample(law, npts, nsigmas, center, sigma, lb, ub):
    """Points and unnormalised weights of *law* for value *center* and
    absolute width *sigma* > 0, inside [lb, ub]."""
    if law.positive:
        lb, ub = max(lb, 1e-8), max(ub, 1e-8)
    if law.placement == "edges":
        x = np.linspace(center - sigma, center + sigma, npts)
    else:
        x = center + np.linspace(-nsigmas * sigma, +nsigmas * sigma, npts)
    x = x[(x >= lb) & (x <= ub)]
    if law.window is not None:
        x = x[law.window(x, center, sigma)]
    return x, law.density(x, center, sigma)


class Dispersion:
    """
    A distribution with *npts* points spanning *nsigmas* widths of *width*
    (unset arguments take the law's defaults).  Subclasses name their law.
    """
    law = Law("base disperser", 35, 3, "span", False, None, _flat)

    def __init_subclass__(cls, **kw):
        super().__init_subclass__(**kw)
        cls.type = cls.law.name
        cls.default = dict(npts=cls.law.npts, width=0, nsigmas=cls.law.nsigmas)

    def __init__(self, npts=None, width=None, nsigmas=None):
        self.npts = self.law.npts if npts is None else int(npts)
        self.width = 0 if width is None else width
        self.nsigmas = self.law.nsigmas if nsigmas is None else nsigmas

    def get_pars(self):
        """{'type': name, 'npts', 'width', 'nsigmas'} (plus the arrays of an
        array distribution)."""
        return dict(type=self.type, **vars(self))

    def set_weights(self, values, weights):
        """Explicit points and weights; only an array distribution takes them."""
        raise RuntimeError("%s distributions compute their weights; use 'array' for given ones"
                           % self.type)

    def get_weights(self, center, lb, ub, relative):
        """
        (points, unnormalised weights) for a parameter of value *center*
        limited to [*lb*, *ub*]; *relative* scales the width by the value
        (volume parameters), otherwise the width is absolute and the points
        are centred on zero (orientation jitter).
        """
        sigma = self.width * center if relative else self.width
        if not relative:
            center = 0
        if sigma == 0 or self.npts < 2:
            keep = lb <= center <= ub
            return np.array([center] if keep else [], "d"), np.array([1.0] if keep else [], "d")
        return self._weights(center, sigma, lb, ub)

    def _weights(self, center, sigma, lb, ub):
        return sample(self.law, self.npts, self.nsigmas, center, sigma, lb, ub)


Dispersion.type = Dispersion.law.name
Dispersion.default = dict(npts=35, width=0, nsigmas=3)


class RectangleDispersion(Dispersion):
    """Flat weights on |x - c| <= sqrt(3) sigma (the width is the standard deviation)."""
    law = _LAWS["rectangle"]


class UniformDispersion(Dispersion):
    """Flat weights on [c - sigma, c + sigma]."""
    law = _LAWS["uniform"]


class LogNormalDispersion(Dispersion):
    """w = exp(-(ln x - ln c)^2 / (2 t^2)) / (x t), t = sigma/c."""
    law = _LAWS["lognormal"]


class GaussianDispersion(Dispersion):
    """w = exp(-(x - c)^2 / (2 sigma^2))."""
    law = _LAWS["gaussian"]


class SchulzDispersion(Dispersion):
    """w = z^z R^(z-1) e^(-R z) / (c Gamma(z)), R = x/c, z = (c/sigma)^2."""
    law = _LAWS["schulz"]


class BoltzmannDispersion(Dispersion):
    """w = exp(-|x - c| / sigma)."""
    law = _LAWS["boltzmann"]


class ArrayDispersion(Dispersion):
    """Points and weights given by the caller (:meth:`set_weights`); those
    inside the parameter limits are used as they are."""
    law = _LAWS["array"]

    def __init__(self, npts=None, width=None, nsigmas=None):
        super().__init__(npts, width, nsigmas)
        self.values = np.array([0.0])
        self.weights = np.array([1.0])

    def set_weights(self, values, weights):
        self.values = np.ascontiguousarray(values, "d")
        self.weights = np.ascontiguousarray(weights, "d")
        self.npts = len(self.values)

    def _weights(self, center, sigma, lb, ub):
        inside = (self.values >= lb) & (self.values <= ub)
        return self.values[inside], self.weights[inside]


#: distribution name -> class, in SasView's menu order
DISTRIBUTIONS = OrderedDict((cls.type, cls) for cls in sorted(
    Dispersion.__subclasses__(), key=lambda cls: list(_LAWS).index(cls.type)))
#: the same registry under the name SasView's fitting panel imports
MODELS = DISTRIBUTIONS


def get_weights(disperser, n, width, nsigmas, value, limits, relative):
    """
    (points, weights summing to one) of distribution *disperser* with *n*
    points over *nsigmas* widths of *width* for a parameter of *value* within
    *limits* = (lb, ub); *relative* as in :meth:`Dispersion.get_weights`.
    """
    if disperser == "array":
        raise NotImplementedError("an array distribution has no width: pass its points and "
                                  "weights to the mesh directly")
    x, w = DISTRIBUTIONS[disperser](n, width, nsigmas).get_weights(value, limits[0], limits[1], relative)
    return x, w / np.sum(w)
