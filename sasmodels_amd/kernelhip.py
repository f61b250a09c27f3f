"""
HIP backend: :class:`HipModel` / :class:`HipKernel` over libsaship.so.

The sibling of the reference's DLL, OpenCL and CUDA backends
(sasmodels/kerneldll.py:332-475, kernelcl.py:404-654,
kernelcuda.py:471-518), with the same interface:

* ``HipModel(model_info).make_kernel([q])`` or ``make_kernel([qx, qy])``;
* ``HipKernel(call_details, values, cutoff, magnetic)`` -> I(q), and
  ``HipKernel.Fq(...)``, through :class:`~sasmodels_amd.kernel.Kernel`;
* the 1D kernel is used for 1D data even when magnetic; 2D data uses the
  magnetic kernel when *magnetic* is true (kerneldll.py:391-392, :454).

q is uploaded once per kernel; each call uploads the mesh descriptor and
values, runs the mesh on the GPU and returns the reference result layout.
The computation is FP32 on the device (FP64 for the weight sums), so
``dtype`` is float64 for the host buffers that cross the boundary.
"""

import ctypes as ct

import numpy as np

from . import _lib
from .kernel import Kernel, KernelModel

KIND_IQ, KIND_IQXY, KIND_IMAGNETIC = 0, 1, 2

_device_ready = False


def use_hip():
    """True if libsaship loads and sees a GPU."""
    try:
        return _lib.device_count() > 0
    except OSError:
        return False


def _ensure_device():
    global _device_ready
    if not _device_ready:
        if _lib.device_count() == 0:
            raise _lib.HipError("no HIP device visible")
        _lib.check(_lib.lib().saship_init(0))
        _device_ready = True


class HipModel(KernelModel):
    """A model compiled for the GPU (one per model and quadrature)."""

    def __init__(self, model_info, dtype=None):
        if model_info.kernel_id is None:
            raise ValueError("model %s has no hip kernel" % model_info.id)
        self.info = model_info
        self.dtype = np.dtype("d")
        self._handle = None
        self._quad = None

    def _model_handle(self):
        quad = int(self.info.integration_size or 0)
        if self._handle is None or quad != self._quad:
            _ensure_device()
            self.release()
            h = ct.c_void_p()
            _lib.check(_lib.lib().saship_model_create(self.info.kernel_id, quad, ct.byref(h)))
            self._handle, self._quad = h, quad
        return self._handle

    def make_kernel(self, q_vectors):
        """Bind the model to *q_vectors* ([q] or [qx, qy])."""
        return HipKernel(self, q_vectors)

    def release(self):
        if self._handle is not None:
            _lib.lib().saship_model_release(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class HipInput:
    """q points in the layout the kernels read: q or interleaved (qx, qy)."""

    def __init__(self, q_vectors):
        self.nq = np.asarray(q_vectors[0]).size
        self.is_2d = len(q_vectors) == 2
        if self.is_2d:
            self.q = np.empty((self.nq, 2), "d")
            self.q[:, 0] = np.asarray(q_vectors[0], "d").ravel()
            self.q[:, 1] = np.asarray(q_vectors[1], "d").ravel()
        else:
            self.q = np.ascontiguousarray(np.asarray(q_vectors[0], "d").ravel())
        self.dtype = np.dtype("d")

    def release(self):
        self.q = None


class HipKernel(Kernel):
    """A :class:`HipModel` bound to q points on the GPU."""

    def __init__(self, model, q_vectors):
        self._model = model
        self.info = model.info
        self.dtype = np.dtype("d")
        self.q_input = HipInput(q_vectors)
        self.dim = "2d" if self.q_input.is_2d else "1d"
        nout = 2 if self.dim == "1d" else 1
        self.result = np.empty(self.q_input.nq*nout + 4, "d")
        mh = model._model_handle()
        h = ct.c_void_p()
        _lib.check(_lib.lib().saship_kernel_create(
            mh, self.q_input.nq, self.q_input.q.ctypes.data, 1 if self.q_input.is_2d else 0,
            ct.byref(h)))
        self._handle = h

    def _call_kernel(self, call_details, values, cutoff, magnetic, radius_effective_mode,
                     pd_start=0, pd_stop=None):
        if self._handle is None:
            raise RuntimeError("kernel has been released")
        kind = KIND_IQ if self.dim == "1d" else (KIND_IMAGNETIC if magnetic else KIND_IQXY)
        details = np.ascontiguousarray(call_details.buffer, "i4")
        vals = np.ascontiguousarray(values, "d")
        stop = int(call_details.num_eval) if pd_stop is None else int(pd_stop)
        _lib.check(_lib.lib().saship_kernel_call(
            self._handle, kind, details.ctypes.data, details.size,
            vals.ctypes.data, vals.size, float(cutoff), int(radius_effective_mode),
            int(pd_start), stop, self.result.ctypes.data))

    def set_timing(self, enable=True):
        """Sum the device time of the evaluation kernel over later calls."""
        _lib.check(_lib.lib().saship_kernel_set_timing(self._handle, 1 if enable else 0))

    def get_timing(self):
        """(total evaluation-kernel milliseconds, number of calls) since set_timing."""
        ms, n = ct.c_double(), ct.c_int32()
        _lib.check(_lib.lib().saship_kernel_get_timing(self._handle, ct.byref(ms), ct.byref(n)))
        return ms.value, n.value

    def release(self):
        if getattr(self, "_handle", None) is not None:
            _lib.lib().saship_kernel_release(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass
