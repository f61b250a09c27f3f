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
import itertools
import threading
import weakref

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


def probe_fma(iters=4096):
    """Measured non-packed FP32 FMA rate of the device in TFLOP/s (saship_probe_fma)."""
    _ensure_device()
    out = ct.c_double()
    _lib.check(_lib.lib().saship_probe_fma(int(iters), ct.byref(out)))
    return out.value


class PinnedPool:
    """
    Output arrays of batched calls in pinned host memory, which the batch's
    download writes by DMA directly (no staging buffer, no host copy).
    A buffer goes back to the pool when the last array or view over it is
    gone, so a caller that keeps every result only grows the pool; past
    *cap_bytes* of buffers in use, arrays are ordinary host memory again.
    The first buffer of a size comes with a spare: a caller that holds the
    previous result while the next call runs (a fit loop) needs two, and
    pinned allocation costs far more than a call.
    """

    def __init__(self, cap_bytes=256 << 20):
        self.cap = int(cap_bytes)
        self._free = {}           # bytes -> [address]
        self._held = 0            # bytes of pinned buffers, in use or free
        self._mu = threading.Lock()

    def _alloc(self, nbytes):
        """A new pinned buffer's address, or None (over the cap, or refused); under _mu."""
        if self._held + nbytes > self.cap:
            return None
        ptr = ct.c_void_p()
        if _lib.lib().saship_host_alloc(nbytes, ct.byref(ptr)) != 0 or not ptr.value:
            return None
        self._held += nbytes
        return ptr.value

    def empty(self, n):
        """A float64 array of n values (pinned when the pool has room)."""
        nbytes = 8*max(int(n), 1)
        with self._mu:
            first = nbytes not in self._free
            stack = self._free.setdefault(nbytes, [])
            addr = stack.pop() if stack else self._alloc(nbytes)
            if first and addr is not None:
                spare = self._alloc(nbytes)
                if spare is not None:
                    stack.append(spare)
        if addr is None:
            return np.empty(n, "d")
        buf = (ct.c_double*(nbytes//8)).from_address(addr)
        fin = weakref.finalize(buf, self._give_back, nbytes, addr)
        fin.atexit = False
        return np.frombuffer(buf, dtype=np.float64, count=n)

    def _give_back(self, nbytes, addr):
        with self._mu:
            self._free[nbytes].append(addr)


#: the process's pool of pinned output arrays
pinned = PinnedPool()


class HipModel(KernelModel):
    """A model compiled for the GPU (one per model and quadrature)."""

    def __init__(self, model_info, dtype=None):
        if model_info.kernel_id is None:
            raise ValueError("model %s has no hip kernel" % model_info.id)
        self.info = model_info
        self.dtype = np.dtype("d")
        # one device handle per quadrature size; kernels keep using the
        # handle they were made with, so none is freed before release()
        self._handles = {}

    def _model_handle(self):
        quad = int(self.info.integration_size or 0)
        if quad not in self._handles:
            _ensure_device()
            h = ct.c_void_p()
            _lib.check(_lib.lib().saship_model_create(self.info.kernel_id, quad, ct.byref(h)))
            self._handles[quad] = h
        return self._handles[quad]

    def make_kernel(self, q_vectors):
        """Bind the model to *q_vectors* ([q] or [qx, qy])."""
        return HipKernel(self, q_vectors)

    def release(self):
        """Free the device tables (kernels made from this model must be released first)."""
        for h in self._handles.values():
            _lib.lib().saship_model_release(h)
        self._handles = {}

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
        nout = 2 if self.dim == "1d" and self.info.have_Fq else 1
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

    def set_separable(self, enable=True):
        """Allow (default) or forbid the separable cylinder evaluation."""
        _lib.check(_lib.lib().saship_kernel_set_path(self._handle, 1 if enable else 0))

    def set_graph(self, enable=True):
        """Replay repeated calls of one shape from a captured hipGraph, or launch eagerly (default)."""
        _lib.check(_lib.lib().saship_kernel_set_graph(self._handle, int(bool(enable))))

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


class HipBatch:
    """
    Many parameter sets, each with its own q points, evaluated in one
    persistent launch (see saship_batch_call in include/saship.h).

    *q_list* holds one ``[q]`` or ``[qx, qy]`` per set.  Calling the batch
    with a list of ``(call_details, values)`` pairs, one per set, returns
    the list of I(q); :meth:`Fq` returns the per-set Fq tuples.  Each set's
    result equals what ``HipKernel`` gives for that set alone.  With
    *use_graph* the device work of a call is captured once as a hipGraph and
    replayed while the shapes stay the same (a fit loop).
    """

    def __init__(self, model, q_list, use_graph=False):
        if not q_list:
            raise ValueError("empty batch")
        self._model = model
        self.info = model.info
        self.dim = "2d" if len(q_list[0]) == 2 else "1d"
        if any((len(qv) == 2) != (self.dim == "2d") for qv in q_list):
            raise ValueError("all sets of a batch must be 1D or all 2D")
        inputs = [HipInput(qv) for qv in q_list]
        self.nq = np.array([inp.nq for inp in inputs], "i4")
        self.qoff = np.concatenate([[0], np.cumsum(self.nq)]).astype(np.int64)
        q = np.concatenate([inp.q.ravel() for inp in inputs]) if self.qoff[-1] else np.zeros(1)
        self._q = np.ascontiguousarray(q, "d")
        mh = model._model_handle()
        h = ct.c_void_p()
        _lib.check(_lib.lib().saship_batch_create(
            mh, len(q_list), self.nq.ctypes.data, self._q.ctypes.data,
            1 if self.dim == "2d" else 0, ct.byref(h)))
        self._handle = h
        self._n_out = None          # smeared points per set (set_resolution)
        self._uniform = bool(self.nq.size) and int(self.nq.min()) == int(self.nq.max()) > 0
        if use_graph:
            self.set_options(use_graph=True)

    @property
    def nsets(self):
        return len(self.nq)

    def set_options(self, use_graph=None, pts_per_item=None, waves_per_cu=None):
        """Change the launch shape; unspecified options keep their values."""
        g, p, w = ct.c_int32(), ct.c_int32(), ct.c_int32()
        _lib.check(_lib.lib().saship_batch_get_options(self._handle, ct.byref(g), ct.byref(p), ct.byref(w)))
        _lib.check(_lib.lib().saship_batch_set_options(
            self._handle, int(g.value if use_graph is None else bool(use_graph)),
            int(p.value if pts_per_item is None else pts_per_item),
            int(w.value if waves_per_cu is None else waves_per_cu)))

    def set_separable(self, enable=True):
        """Allow (default) or forbid the separable cylinder evaluation."""
        _lib.check(_lib.lib().saship_batch_set_path(self._handle, int(bool(enable))))

    def set_resolution(self, resolutions):
        """
        Smear every set's I(q) on the device: *resolutions* holds one
        resolution object per set whose q_calc is that set's q (Pinhole1D,
        Slit1D, Pinhole2D, Slit2D; None or Perfect1D: no smearing for that set), or
        is None to switch smearing off.  Calls then return I at each
        resolution's measured q (pixels for Pinhole2D).
        """
        if resolutions is None:
            _lib.check(_lib.lib().saship_batch_set_resolution(self._handle, None, None, None, None, None, None, 0))
            self._n_out = None
            return
        from .resolution import smear_bands
        if len(resolutions) != self.nsets:
            raise ValueError("expected %d resolutions, got %d" % (self.nsets, len(resolutions)))
        parts = [smear_bands(res, int(self.nq[s])) for s, res in enumerate(resolutions)]
        offsets, total = [], 0
        for p in parts:
            offsets.append(p[4] + total)
            total += p[5].size
        keep = [np.ascontiguousarray(np.array([p[0] for p in parts], np.int32)),
                np.ascontiguousarray(np.concatenate([p[1] for p in parts]).astype(np.int32)),
                np.ascontiguousarray(np.concatenate([p[2] for p in parts]).astype(np.int32)),
                np.ascontiguousarray(np.concatenate([p[3] for p in parts]).astype(np.int64)),
                np.ascontiguousarray(np.concatenate(offsets).astype(np.int64)),
                np.ascontiguousarray(np.concatenate([p[5] for p in parts]) if total else np.zeros(1))]
        if not any(len(p) > 6 for p in parts):
            _lib.check(_lib.lib().saship_batch_set_resolution(
                self._handle, *[k.ctypes.data for k in keep], int(total)))
        else:
            # repeated bands (Slit2D); a set without them repeats its band once
            reps, steps, roffs, rws, rtotal = [], [], [], [], 0
            for p in parts:
                n = int(p[0])
                rep, step, roff, rw = p[6] if len(p) > 6 else (np.ones(n, np.int32), np.zeros(n, np.int64),
                                                              np.zeros(n, np.int64), np.ones(1))
                reps.append(rep)
                steps.append(step)
                roffs.append(np.asarray(roff, np.int64) + rtotal)
                rws.append(np.asarray(rw, float))
                rtotal += len(rw)
            keep += [np.ascontiguousarray(np.concatenate(reps).astype(np.int32)),
                     np.ascontiguousarray(np.concatenate(steps).astype(np.int64)),
                     np.ascontiguousarray(np.concatenate(roffs).astype(np.int64)),
                     np.ascontiguousarray(np.concatenate(rws))]
            _lib.check(_lib.lib().saship_batch_set_resolution_repeat(
                self._handle, *[k.ctypes.data for k in keep[:6]], int(total),
                *[k.ctypes.data for k in keep[6:]], int(rtotal)))
        self._res_keep = keep
        self._n_out = keep[0].astype(np.int64)

    def set_structure(self, pars, s_model=5):
        """
        Multiply every set's form factor by a structure factor on the device
        (saship_batch_set_structure): *pars* is an (nsets, 9) array of
        (radius_effective, volfraction, structure_factor_mode,
        radius_effective_mode, volfraction_in_form, s2, s3, s4, s5) per set,
        s2..s5 the structure factor's further parameters (0 past its last);
        *s_model* its kernel id (5 hard sphere, 14 sticky hard sphere,
        15 square well, 16 Hayter MSA); None switches the product off.
        Calls then return the P@S intensity of each set.
        """
        if pars is None:
            _lib.check(_lib.lib().saship_batch_set_structure(self._handle, 0, None, 0))
            return
        pars = np.ascontiguousarray(pars, "d").reshape(-1, 9)
        _lib.check(_lib.lib().saship_batch_set_structure(
            self._handle, int(s_model), pars.ctypes.data, pars.shape[0]))

    def _pack(self, calls):
        if len(calls) != self.nsets:
            raise ValueError("expected %d (call_details, values) pairs, got %d" % (self.nsets, len(calls)))
        dets, vals = zip(*calls)
        det = np.concatenate([getattr(cd, "buffer", cd) for cd in dets]).astype(np.int32, copy=False)
        voff = np.fromiter(itertools.accumulate(map(len, vals), initial=0), np.int64, len(vals) + 1)
        vals = np.concatenate(vals).astype(np.float64, copy=False)
        return np.ascontiguousarray(det), np.ascontiguousarray(vals), voff, det.size//self.nsets

    def _nout(self):
        """Outputs per q in the raw result: <F^2> and <F> for 1D models with <F>."""
        return 2 if self.dim == "1d" and self.info.have_Fq else 1

    def _kind(self, magnetic):
        return KIND_IQ if self.dim == "1d" else (KIND_IMAGNETIC if magnetic else KIND_IQXY)

    def raw(self, calls, cutoff=0.0, magnetic=False, radius_effective_mode=0, want_result=True, want_iq=True,
            out=None):
        """(raw accumulators of all sets back to back, I(q) of all sets back to back).
        *out*: a float64 array of the I(q) size to write I(q) into (a fit
        loop that reuses it avoids faulting in new memory on every call);
        without it I(q) goes into an array from the pinned pool, which the
        download fills directly."""
        return self.raw_packed(self._pack(calls), cutoff, magnetic, radius_effective_mode, want_result, want_iq,
                               out)

    def raw_packed(self, packed, cutoff=0.0, magnetic=False, radius_effective_mode=0, want_result=True,
                   want_iq=True, out=None):
        """:meth:`raw` on inputs already packed as (details of every set back
        to back, values of every set back to back, values offsets [nsets+1],
        details per set)."""
        det, vals, voff, ndet = packed
        if det.dtype != np.int32 or vals.dtype != np.float64 or voff.dtype != np.int64 \
                or det.size != ndet*self.nsets or voff.size != self.nsets + 1 or voff[-1] > vals.size:
            raise ValueError("packed inputs do not match the batch")
        det, vals, voff = (np.ascontiguousarray(a) for a in (det, vals, voff))
        nout = self._nout()
        result = np.empty(int(np.sum(nout*self.nq + 4)), "d") if want_result else None
        n_iq = int(self.qoff[-1]) if self._n_out is None else int(self._n_out.sum())
        if out is not None and want_iq:
            if not (isinstance(out, np.ndarray) and out.dtype == np.float64 and out.size == n_iq
                    and out.flags.c_contiguous and out.flags.writeable):
                raise ValueError("out must be a writeable contiguous float64 array of %d values" % n_iq)
            iq = out.reshape(-1)
        else:
            iq = pinned.empty(n_iq) if want_iq else None
        _lib.check(_lib.lib().saship_batch_call(
            self._handle, self._kind(magnetic), det.ctypes.data, ndet, vals.ctypes.data, voff.ctypes.data,
            float(cutoff), int(radius_effective_mode),
            result.ctypes.data if want_result else None, iq.ctypes.data if want_iq else None))
        return result, iq

    def __call__(self, calls, cutoff=0.0, magnetic=False, radius_effective_mode=0, out=None):
        """List of I(q), one per set (views of one array: *out* when given, see :meth:`raw`)."""
        _, iq = self.raw(calls, cutoff, magnetic, radius_effective_mode, want_result=False, want_iq=True, out=out)
        return self._sets(iq)

    def call_packed(self, packed, cutoff=0.0, magnetic=False, radius_effective_mode=0, out=None):
        """:meth:`__call__` on packed inputs (see :meth:`raw_packed`)."""
        _, iq = self.raw_packed(packed, cutoff, magnetic, radius_effective_mode, want_result=False, want_iq=True,
                                out=out)
        return self._sets(iq)

    def _sets(self, iq):
        """I(q) of all sets back to back -> one view per set."""
        if self._n_out is not None:
            return np.split(iq, np.cumsum(self._n_out)[:-1])
        if self._uniform:
            return list(iq.reshape(self.nsets, -1))      # rows of one array: views, as np.split gives
        return np.split(iq, self.qoff[1:-1])

    def Fq(self, calls, cutoff=0.0, magnetic=False, radius_effective_mode=0):
        """List of (F1, F2, R_eff, V_shell, V_form/V_shell), one per set."""
        result, _ = self.raw(calls, cutoff, magnetic, radius_effective_mode, want_result=True, want_iq=False)
        nout = self._nout()
        out = []
        off = 0
        for nq in self.nq:
            r = result[off:off + nout*nq + 4]
            off += nout*nq + 4
            total = r[nout*nq] or 1.0
            shell = r[nout*nq + 2]/total or 1.0
            F1 = r[1:nout*nq:nout]/total if nout == 2 else None
            out.append((F1, r[0:nout*nq:nout]/total, r[nout*nq + 3]/total, shell,
                        (r[nout*nq + 1]/total)/shell))
        return out

    def autotune(self, calls, cutoff=0.0, magnetic=False):
        """Pick the fastest (pts_per_item, waves_per_cu) on these inputs."""
        det, vals, voff, ndet = self._pack(calls)
        p, w = ct.c_int32(), ct.c_int32()
        _lib.check(_lib.lib().saship_batch_autotune(
            self._handle, self._kind(magnetic), det.ctypes.data, ndet, vals.ctypes.data, voff.ctypes.data,
            float(cutoff), ct.byref(p), ct.byref(w)))
        return p.value, w.value

    def replay(self, nsteps):
        """Device time in ms of nsteps back-to-back runs of the last call's
        device work, inputs and outputs resident in device memory."""
        ms = ct.c_double()
        _lib.check(_lib.lib().saship_batch_replay(self._handle, int(nsteps), ct.byref(ms)))
        return ms.value

    def set_timing(self, enable=True):
        _lib.check(_lib.lib().saship_batch_set_timing(self._handle, 1 if enable else 0))

    def get_timing(self):
        ms, n = ct.c_double(), ct.c_int32()
        _lib.check(_lib.lib().saship_batch_get_timing(self._handle, ct.byref(ms), ct.byref(n)))
        return ms.value, n.value

    def release(self):
        if getattr(self, "_handle", None) is not None:
            _lib.lib().saship_batch_release(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass
