"""
Kernel call details: the int32 mesh descriptor and the packed values vector.

Produces buffers with the same layout as the reference
(sasmodels/details.py:43-254), so buffers built by either side can be
passed to either backend:

* ``CallDetails.buffer`` = ``pd_par[max_pd] | pd_length[max_pd] |
  pd_offset[max_pd] | pd_stride[max_pd] | num_eval | num_weights |
  num_active | theta_par``;
* mesh levels are the parameters sorted by decreasing distribution length
  (``argsort(length)[::-1]``), level 0 innermost with stride 1;
* ``values`` = scalars, then every parameter's distribution values, then its
  weights, zero padded to a multiple of 32, in the kernel dtype;
* magnetic (M0, theta, phi) triples are converted in place to (mx, my, mz).
"""

import numpy as np
from numpy import cos, radians, sin

from .modelinfo import NUM_COMMON_PARS, NUM_MAGNETIC_PARS


class CallDetails:
    """Mesh descriptor for one kernel call (details.py:43-180)."""
    parts = None

    def __init__(self, model_info):
        parameters = model_info.parameters
        max_pd = parameters.max_pd
        self.buffer = np.empty(4*max_pd + 4, "i4")
        self._pd_par = self.buffer[0*max_pd:1*max_pd]
        self._pd_length = self.buffer[1*max_pd:2*max_pd]
        self._pd_offset = self.buffer[2*max_pd:3*max_pd]
        self._pd_stride = self.buffer[3*max_pd:4*max_pd]
        self.theta_par = parameters.theta_offset
        self.offset = None
        self.length = None
        self.info = model_info

    @property
    def pd_par(self):
        """Parameter index of each mesh level"""
        return self._pd_par

    @property
    def pd_length(self):
        """Distribution length of each mesh level"""
        return self._pd_length

    @property
    def pd_offset(self):
        """Offset of each level's distribution in the value/weight vectors"""
        return self._pd_offset

    @property
    def pd_stride(self):
        """Mesh stride of each level"""
        return self._pd_stride

    @property
    def num_eval(self):
        """Total mesh size"""
        return self.buffer[-4]

    @num_eval.setter
    def num_eval(self, v):
        self.buffer[-4] = v

    @property
    def num_weights(self):
        """Total length of the distribution vectors"""
        return self.buffer[-3]

    @num_weights.setter
    def num_weights(self, v):
        self.buffer[-3] = v

    @property
    def num_active(self):
        """Number of distributions longer than one point"""
        return self.buffer[-2]

    @num_active.setter
    def num_active(self, v):
        self.buffer[-2] = v

    @property
    def theta_par(self):
        """Kernel-parameter index of theta, or -1"""
        return self.buffer[-1]

    @theta_par.setter
    def theta_par(self, v):
        self.buffer[-1] = v

    def show(self, values=None):
        """Print the mesh descriptor."""
        print("===== %s details ====" % self.info.name)
        print("num_active:%d  num_eval:%d  num_weights:%d  theta=%d"
              % (self.num_active, self.num_eval, self.num_weights, self.theta_par))
        if self.pd_par.size:
            print("pd_par", self.pd_par)
            print("pd_length", self.pd_length)
            print("pd_offset", self.pd_offset)
            print("pd_stride", self.pd_stride)
        if values is not None:
            nvalues = self.info.parameters.nvalues
            print("scale, background", values[:2])
            print("val", values[2:nvalues])
            print("pd", values[nvalues:nvalues + self.num_weights])
            print("wt", values[nvalues + self.num_weights:nvalues + 2*self.num_weights])


def make_details(model_info, length, offset, num_weights):
    """
    Mesh descriptor for distribution *length* and *offset* per kernel
    parameter (details.py:183-221).
    """
    num_active = np.sum(length > 1)
    max_pd = model_info.parameters.max_pd
    if num_active > max_pd:
        raise ValueError("Too many polydisperse parameters")
    idx = np.argsort(length)[::-1][:max_pd]
    pd_stride = np.cumprod(np.hstack((1, length[idx])))
    call_details = CallDetails(model_info)
    call_details.pd_par[:max_pd] = idx
    call_details.pd_length[:max_pd] = length[idx]
    call_details.pd_offset[:max_pd] = offset[idx]
    call_details.pd_stride[:max_pd] = pd_stride[:-1]
    call_details.num_eval = pd_stride[-1]
    call_details.num_weights = num_weights
    call_details.num_active = num_active
    call_details.length = length
    call_details.offset = offset
    return call_details


_ZEROS = tuple([0.]*31)


def make_kernel_args(kernel, mesh):
    """
    Convert a mesh of (value, dispersity, weight) per call parameter into
    ``(call_details, values, is_magnetic)`` (details.py:225-254).
    """
    npars = kernel.info.parameters.npars
    nvalues = kernel.info.parameters.nvalues
    scalars = [value for value, _dispersity, _weight in mesh]
    _, dispersity, weight = (zip(*mesh[NUM_COMMON_PARS:npars + NUM_COMMON_PARS])
                             if npars else ((), (), ()))
    length = np.array([len(w) for w in weight])
    offset = np.cumsum(np.hstack((0, length)))
    call_details = make_details(kernel.info, length, offset[:-1], offset[-1])
    data_len = nvalues + 2*sum(len(v) for v in dispersity)
    extra = (32 - data_len % 32) % 32
    data = np.hstack((scalars,) + tuple(dispersity) + tuple(weight) + _ZEROS[:extra])
    data = data.astype(kernel.dtype)
    is_magnetic = convert_magnetism(kernel.info.parameters, data)
    return call_details, data, is_magnetic


def convert_magnetism(parameters, values):
    """
    Convert (M0, theta, phi) per sld to (mx, my, mz) in place; True if any
    M0 is non-zero (details.py:288-306).
    """
    nmagpars = NUM_MAGNETIC_PARS*parameters.nmagnetic
    mag = values[parameters.nvalues - nmagpars:parameters.nvalues]
    mag = mag.reshape(-1, NUM_MAGNETIC_PARS)
    if np.any(mag[:, 0] != 0.0):
        M0 = mag[:, 0].copy()
        theta, phi = radians(mag[:, 1]), radians(mag[:, 2])
        mag[:, 0] = M0*sin(theta)*cos(phi)
        mag[:, 1] = M0*sin(theta)*sin(phi)
        mag[:, 2] = M0*cos(theta)
        return True
    return False
