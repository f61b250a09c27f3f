"""
Script-level evaluation: build the mesh for a parameter set and call a kernel.

Restates the reference's call assembly (sasmodels/direct_model.py:41-154):
every call parameter becomes ``(value, dispersity, weights)``; a parameter
is polydisperse only if it is active for the kernel's dimension
(``pd_1d``/``pd_2d``) and has ``<name>_pd_n`` > 0 and ``<name>_pd`` != 0;
defaults are ``_pd_n`` = 0, ``_pd_nsigma`` = 3 and ``_pd_type`` =
'gaussian'; inactive orientation parameters carry jitter 0.
"""

import numpy as np

from . import weights
from .details import make_kernel_args

#: parameter holding the effective-radius mode in Fq calls (product.py)
RADIUS_MODE_ID = "radius_effective_mode"


def call_kernel(calculator, pars, cutoff=0., mono=False):
    """Evaluate *calculator* (a :class:`~kernel.Kernel`) at *pars*; returns I(q)."""
    mesh = get_mesh(calculator.info, pars, dim=calculator.dim, mono=mono)
    call_details, values, is_magnetic = make_kernel_args(calculator, mesh)
    return calculator(call_details, values, cutoff, is_magnetic)


def call_Fq(calculator, pars, cutoff=0., mono=False):
    """Like :func:`call_kernel` but returns (F, F^2, R_eff, V_shell, V_form/V_shell)."""
    pars = dict(pars)
    R_eff_type = int(pars.pop(RADIUS_MODE_ID, 1.0))
    mesh = get_mesh(calculator.info, pars, dim=calculator.dim, mono=mono)
    call_details, values, is_magnetic = make_kernel_args(calculator, mesh)
    return calculator.Fq(call_details, values, cutoff, is_magnetic, R_eff_type)


def get_mesh(model_info, values, dim="1d", mono=False):
    """
    List of *(value, dispersity, weights)* per call parameter; unknown keys
    raise TypeError.
    """
    parameters = model_info.parameters
    if mono:
        active = lambda name: False
    elif dim == "1d":
        active = lambda name: name in parameters.pd_1d
    elif dim == "2d":
        active = lambda name: name in parameters.pd_2d
    else:
        active = lambda name: True
    values = dict(values)
    mesh = [_pop_par_weights(p, values, active(p.name))
            for p in parameters.call_parameters]
    if values:
        raise TypeError("Unused parameters in call: %s" % ", ".join(values.keys()))
    return mesh


def _pop_par_weights(parameter, values, active=True):
    value = float(values.pop(parameter.name, parameter.default))
    if parameter.polydisperse:
        npts = values.pop(parameter.name + "_pd_n", 0)
        width = values.pop(parameter.name + "_pd", 0.0)
        nsigma = values.pop(parameter.name + "_pd_nsigma", 3.0)
        distribution = values.pop(parameter.name + "_pd_type", "gaussian")
        relative = parameter.relative_pd
        if npts == 0 or width == 0.0 or not active:
            pd = [value if relative else 0.0], [1.0]
        else:
            pd = weights.get_weights(distribution, npts, width, nsigma,
                                     value, parameter.limits, relative)
    else:
        pd = [value], [1.0]
    return value, np.asarray(pd[0], "d"), np.asarray(pd[1], "d")
