"""
Polydispersity distributions: (values, weights) for one parameter.

Same distributions, defaults and point placement as the reference
(sasmodels/weights.py:21-291): points are ``center + linspace(-n*sigma,
+n*sigma, npts)`` clipped to the parameter limits, the width is relative to
the value for volume parameters and absolute (and centred on zero) for
orientation jitter, and :func:`get_weights` normalises the weights to sum 1.
"""

from collections import OrderedDict

import numpy as np
from scipy.special import gammaln


class Dispersion:
    """
    Base distribution with *npts* points spanning *nsigmas* widths.
    Subclasses implement ``_weights(center, sigma, lb, ub)``.
    """
    type = "base disperser"
    default = dict(npts=35, width=0, nsigmas=3)

    def __init__(self, npts=None, width=None, nsigmas=None):
        self.npts = self.default["npts"] if npts is None else int(npts)
        self.width = self.default["width"] if width is None else width
        self.nsigmas = self.default["nsigmas"] if nsigmas is None else nsigmas

    def get_pars(self):
        """Parameters of the distribution as a dictionary."""
        pars = {"type": self.type}
        pars.update(self.__dict__)
        return pars

    def set_weights(self, values, weights):
        """Only :class:`ArrayDispersion` accepts explicit weights."""
        raise RuntimeError("set_weights is only available for ArrayDispersion")

    def get_weights(self, center, lb, ub, relative):
        """
        Points and unnormalised weights for a parameter of value *center*
        with limits [*lb*, *ub*]; *relative* widths scale with the value.
        """
        sigma = self.width * center if relative else self.width
        if not relative:
            center = 0     # orientation jitter is centred on zero
        if sigma == 0 or self.npts < 2:
            if lb <= center <= ub:
                return np.array([center], "d"), np.array([1.], "d")
            return np.array([], "d"), np.array([], "d")
        return self._weights(center, sigma, lb, ub)

    def _weights(self, center, sigma, lb, ub):
        raise NotImplementedError

    def _linspace(self, center, sigma, lb, ub):
        x = center + np.linspace(-self.nsigmas*sigma, +self.nsigmas*sigma, self.npts)
        return x[(x >= lb) & (x <= ub)]


class GaussianDispersion(Dispersion):
    """w = exp(-(x-c)^2 / (2 sigma^2))"""
    type = "gaussian"
    default = dict(npts=35, width=0, nsigmas=3)

    def _weights(self, center, sigma, lb, ub):
        x = self._linspace(center, sigma, lb, ub)
        return x, np.exp((x - center)**2 / (-2.0 * sigma * sigma))


class UniformDispersion(Dispersion):
    """w = 1 on [c - sigma, c + sigma]"""
    type = "uniform"
    default = dict(npts=35, width=0, nsigmas=None)

    def _weights(self, center, sigma, lb, ub):
        x = np.linspace(center - sigma, center + sigma, self.npts)
        x = x[(x >= lb) & (x <= ub)]
        return x, np.ones_like(x)


class RectangleDispersion(Dispersion):
    """w = 1 on |x - c| <= sqrt(3) sigma"""
    type = "rectangle"
    default = dict(npts=35, width=0, nsigmas=1.73205)

    def _weights(self, center, sigma, lb, ub):
        x = self._linspace(center, sigma, lb, ub)
        x = x[np.fabs(x - center) <= np.fabs(sigma)*np.sqrt(3.0)]
        return x, np.ones_like(x)


class LogNormalDispersion(Dispersion):
    """w = exp(-(ln x - ln c)^2 / (2 s^2)) / (x s), s = sigma/c"""
    type = "lognormal"
    default = dict(npts=80, width=0, nsigmas=8)

    def _weights(self, center, sigma, lb, ub):
        x = self._linspace(center, sigma, max(lb, 1e-8), max(ub, 1e-8))
        sig = np.fabs(sigma/center)
        return x, np.exp(-0.5*((np.log(x) - np.log(center))/sig)**2)/(x*sig)


class SchulzDispersion(Dispersion):
    """w = z^z R^(z-1) e^(-Rz) / (c Gamma(z)), R = x/c, z = (c/sigma)^2"""
    type = "schulz"
    default = dict(npts=80, width=0, nsigmas=8)

    def _weights(self, center, sigma, lb, ub):
        x = self._linspace(center, sigma, max(lb, 1e-8), max(ub, 1e-8))
        R = x/center
        z = (center/sigma)**2
        arg = z*np.log(z) + (z - 1)*np.log(R) - R*z - np.log(center) - gammaln(z)
        return x, np.exp(arg)


class ArrayDispersion(Dispersion):
    """User-supplied points and weights (set with :meth:`set_weights`)."""
    type = "array"
    default = dict(npts=35, width=0, nsigmas=1)

    def __init__(self, npts=None, width=None, nsigmas=None):
        Dispersion.__init__(self, npts, width, nsigmas)
        self.values = np.array([0.], "d")
        self.weights = np.array([1.], "d")

    def set_weights(self, values, weights):
        self.values = np.ascontiguousarray(values, "d")
        self.weights = np.ascontiguousarray(weights, "d")
        self.npts = len(values)

    def _weights(self, center, sigma, lb, ub):
        idx = (self.values >= lb) & (self.values <= ub)
        return self.values[idx], self.weights[idx]


class BoltzmannDispersion(Dispersion):
    """w = exp(-|x - c| / sigma)"""
    type = "boltzmann"
    default = dict(npts=35, width=0, nsigmas=3)

    def _weights(self, center, sigma, lb, ub):
        x = self._linspace(center, sigma, lb, ub)
        return x, np.exp(-np.abs(x - center)/np.abs(sigma))


#: name -> distribution class, in the reference's GUI order (weights.py:221-229)
DISTRIBUTIONS = OrderedDict((d.type, d) for d in (
    RectangleDispersion,
    UniformDispersion,
    ArrayDispersion,
    LogNormalDispersion,
    GaussianDispersion,
    SchulzDispersion,
    BoltzmannDispersion,
))
MODELS = DISTRIBUTIONS


def get_weights(disperser, n, width, nsigmas, value, limits, relative):
    """
    Return *(values, weights)* for a polydisperse parameter, weights
    normalised to sum to one (weights.py:264-291).
    """
    if disperser == "array":
        raise NotImplementedError("Don't handle arrays through get_weights;"
                                  " use values and weights directly")
    cls = DISTRIBUTIONS[disperser]
    obj = cls(n, width, nsigmas)
    v, w = obj.get_weights(value, limits[0], limits[1], relative)
    return v, w/np.sum(w)
