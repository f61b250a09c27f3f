"""
The two buffers of a kernel call: the int32 mesh descriptor and the values
vector.

Their layouts are the drop-in contract with the reference's kernels
(sasmodels/details.py:73-88, :225-254; kernel_iq.c:47-58), so buffers built
here or by the reference work with either backend:

    descriptor  pd_par[max_pd] | pd_length[max_pd] | pd_offset[max_pd] |
                pd_stride[max_pd] | num_eval | num_weights | num_active |
                theta_par
    values      one scalar per call parameter (scale, background, the kernel
                parameters, then the magnetic block), the distribution points
                of every kernel parameter back to back, their weights in the
                same order, zeros up to a multiple of 32, in the kernel dtype

Mesh level 0 is the innermost loop (stride 1); levels are the parameters in
decreasing distribution length, ties in the order that reversing numpy's
argsort of the lengths gives (details.py:207) -- the order decides the
kernels' summation order and what a pd_start/pd_stop chunk of the mesh is.
In the magnetic block each (M0, mtheta, mphi) triple becomes the vector
(mx, my, mz) = M0 (sin t cos p, sin t sin p, cos t) in place, the kernels'
convention (details.py:288-306).

:func:`make_kernel_args` fills both buffers in one pass over the mesh.
"""

import numpy as np

from .modelinfo import NUM_COMMON_PARS, NUM_MAGNETIC_PARS


class _Trailer:
    """One of the four scalars after the per-level arrays (-4 .. -1)."""
    def __init__(self, index, doc):
        self.index = index
        self.__doc__ = doc

    def __get__(self, obj, owner=None):
        return self if obj is None else obj.buffer[self.index]

    def __set__(self, obj, value):
        obj.buffer[self.index] = value


class CallDetails:
    """
    Mesh descriptor of one call: one int32 ``buffer`` and views of its
    fields.  ``length``/``offset`` keep the per-parameter distribution sizes
    and starts the descriptor was made from (products and mixtures re-slice
    them per part).
    """
    num_eval = _Trailer(-4, "points in the whole mesh")
    num_weights = _Trailer(-3, "length of the points (and of the weights) block of the values")
    num_active = _Trailer(-2, "parameters with more than one point")
    theta_par = _Trailer(-1, "kernel index of theta, -1 without orientation")

    def __init__(self, model_info):
        max_pd = model_info.parameters.max_pd
        self.info = model_info
        self.buffer = np.zeros(4 * max_pd + 4, np.int32)
        fields = self.buffer[:4 * max_pd].reshape(4, max_pd)
        self.pd_par, self.pd_length, self.pd_offset, self.pd_stride = fields
        self.theta_par = model_info.parameters.theta_offset
        self.length = self.offset = None


def make_details(model_info, length, offset, num_weights):
    """
    Descriptor for kernel parameters whose distributions have *length*
    points starting at *offset* in the points block (*num_weights* long).
    """
    length = np.asarray(length)
    offset = np.asarray(offset)
    max_pd = model_info.parameters.max_pd
    active = int(np.count_nonzero(length > 1))
    if active > max_pd:
        raise ValueError("%s takes at most %d polydisperse parameters, got %d"
                         % (model_info.name, max_pd, active))
    cd = CallDetails(model_info)
    levels = np.argsort(length)[::-1][:max_pd]
    cd.pd_par[:] = levels
    cd.pd_length[:] = length[levels]
    cd.pd_offset[:] = offset[levels]
    stride = 1
    for k, n in enumerate(length[levels]):
        cd.pd_stride[k] = stride
        stride *= int(n)
    if stride > np.iinfo(np.int32).max:
        raise ValueError("mesh of %d points exceeds the int32 descriptor" % stride)
    cd.num_eval = stride
    cd.num_weights = num_weights
    cd.num_active = active
    cd.length, cd.offset = length, offset
    return cd


def make_kernel_args(kernel, mesh):
    """
    ``(call_details, values, is_magnetic)`` for *mesh*, one
    ``(value, points, weights)`` per call parameter (direct_model.get_mesh);
    *kernel* supplies ``info`` and ``dtype``.
    """
    table = kernel.info.parameters
    nvalues, npars = table.nvalues, table.npars
    dists = mesh[NUM_COMMON_PARS:NUM_COMMON_PARS + npars]
    length = np.array([len(w) for _, _, w in dists], dtype=int)
    offset = np.zeros(npars, dtype=int)
    if npars:
        np.cumsum(length[:-1], out=offset[1:])
    num_weights = int(length.sum())

    used = nvalues + 2 * num_weights
    values = np.zeros(used + (-used) % 32, dtype=kernel.dtype)
    values[:len(mesh)] = [v for v, _, _ in mesh]
    points = values[nvalues:nvalues + num_weights]
    weights = values[nvalues + num_weights:used]
    for start, n, (_, x, w) in zip(offset, length, dists):
        points[start:start + n] = x
        weights[start:start + n] = w

    call_details = make_details(kernel.info, length, offset, num_weights)
    return call_details, values, _spin_vectors(table, values)


def _spin_vectors(table, values):
    """(M0, mtheta, mphi) -> (mx, my, mz) in place; True when any M0 != 0."""
    if not table.nmagnetic:
        return False
    end = table.nvalues
    block = values[end - NUM_MAGNETIC_PARS * table.nmagnetic:end].reshape(-1, NUM_MAGNETIC_PARS)
    M0 = block[:, 0].copy()
    if not M0.any():
        return False
    theta, phi = np.radians(block[:, 1]), np.radians(block[:, 2])
    block[:, 0] = M0 * np.sin(theta) * np.cos(phi)
    block[:, 1] = M0 * np.sin(theta) * np.sin(phi)
    block[:, 2] = M0 * np.cos(theta)
    return True
