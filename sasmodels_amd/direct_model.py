"""
Script-level evaluation: the mesh of a parameter dictionary, and one kernel
call on it.

The mesh has one ``(value, points, weights)`` entry per call parameter,
assembled by the rules of the reference's call path
(sasmodels/direct_model.py:41-154), which decide the values vector a fit
sends to the kernel:

* a parameter reads ``name`` (default: its table default); a polydisperse
  one also reads ``name_pd`` (width, 0), ``name_pd_n`` (points, 0),
  ``name_pd_nsigma`` (3) and ``name_pd_type`` ('gaussian');
* it is dispersed only when it is active for the kernel's dimension (1D:
  volume parameters; 2D: volume and orientation), ``name_pd_n`` is not 0 and
  ``name_pd`` is not 0; otherwise it is the single point of weight 1 at its
  value, or at 0 for an orientation parameter, whose points are jitter
  about the view angle;
* a key that names no call parameter raises TypeError.
"""

import numpy as np

from . import weights
from .details import make_kernel_args

#: parameter holding the effective-radius mode in Fq calls (product.py)
RADIUS_MODE_ID = "radius_effective_mode"

#: suffix -> default of the keys describing a parameter's distribution
PD_KEYS = (("_pd", 0.0), ("_pd_n", 0), ("_pd_nsigma", 3.0), ("_pd_type", "gaussian"))


def call_kernel(calculator, pars, cutoff=0., mono=False):
    """I(q) of *calculator* (a :class:`~kernel.Kernel`) at *pars*."""
    details, values, magnetic = make_kernel_args(calculator, get_mesh(calculator.info, pars, calculator.dim, mono))
    return calculator(details, values, cutoff, magnetic)


def call_Fq(calculator, pars, cutoff=0., mono=False):
    """(F, F^2, R_eff, V_shell, V_form/V_shell) of *calculator* at *pars*;
    ``radius_effective_mode`` in *pars* selects R_eff (default 1)."""
    pars = dict(pars)
    mode = int(pars.pop(RADIUS_MODE_ID, 1.0))
    details, values, magnetic = make_kernel_args(calculator, get_mesh(calculator.info, pars, calculator.dim, mono))
    return calculator.Fq(details, values, cutoff, magnetic, mode)


def get_mesh(model_info, values, dim="1d", mono=False):
    """
    One ``(value, points, weights)`` per call parameter of *model_info* for
    the parameter dictionary *values*.  *dim* ('1d', '2d', or anything else
    for "every polydisperse parameter") selects the active distributions;
    *mono* switches all of them off.
    """
    table = model_info.parameters
    if mono:
        active = frozenset()
    elif dim == "1d":
        active = table.pd_1d
    elif dim == "2d":
        active = table.pd_2d
    else:
        active = frozenset(p.name for p in table.call_parameters if p.polydisperse)
    unread = dict(values)
    mesh = [_distribution(p, unread, p.name in active) for p in table.call_parameters]
    if unread:
        raise TypeError("parameters not used by %s: %s" % (model_info.name, ", ".join(sorted(unread))))
    return mesh


def _distribution(par, unread, active):
    """Pop *par*'s keys from *unread*; return its mesh entry."""
    value = float(unread.pop(par.name, par.default))
    if not par.polydisperse:
        return value, np.array([value]), np.array([1.0])
    width, npts, nsigma, kind = [unread.pop(par.name + suffix, default) for suffix, default in PD_KEYS]
    if active and npts != 0 and width != 0.0:
        x, w = weights.get_weights(kind, npts, width, nsigma, value, par.limits, par.relative_pd)
        return value, np.asarray(x, "d"), np.asarray(w, "d")
    return value, np.array([value if par.relative_pd else 0.0]), np.array([1.0])
