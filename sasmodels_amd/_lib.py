"""
ctypes binding of libsaship.so (declarations in include/saship.h).

The library is built in-tree next to this file (``__graft_entry__.build()``
or ``python -m sasmodels_amd.build``).  There is no fallback: if the library
is missing or fails to load, :func:`lib` raises.
"""

import ctypes as ct
import os
import threading

LIBNAME = "libsaship.so"
LIBPATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIBNAME)

_lock = threading.Lock()
_lib = None

c_int32_p = ct.POINTER(ct.c_int32)
c_double_p = ct.POINTER(ct.c_double)

#: exported symbol -> (restype, argtypes)
SIGNATURES = {
    "saship_last_error": (ct.c_char_p, []),
    "saship_device_count": (ct.c_int, [c_int32_p]),
    "saship_init": (ct.c_int, [ct.c_int32]),
    "saship_model_create": (ct.c_int, [ct.c_int32, ct.c_int32, ct.POINTER(ct.c_void_p)]),
    "saship_model_release": (ct.c_int, [ct.c_void_p]),
    "saship_kernel_create": (ct.c_int, [ct.c_void_p, ct.c_int32, ct.c_void_p, ct.c_int32,
                                        ct.POINTER(ct.c_void_p)]),
    "saship_kernel_call": (ct.c_int, [ct.c_void_p, ct.c_int32, ct.c_void_p, ct.c_int32,
                                      ct.c_void_p, ct.c_int32, ct.c_double, ct.c_int32,
                                      ct.c_int32, ct.c_int32, ct.c_void_p]),
    "saship_kernel_call_noalloc": (ct.c_int, [ct.c_void_p, ct.c_int32, ct.c_int32, ct.c_void_p,
                                              ct.c_void_p, ct.c_int32, ct.c_void_p, ct.c_int32,
                                              ct.c_double, ct.c_int32, ct.c_int32, ct.c_int32,
                                              ct.c_void_p, ct.c_void_p]),
    "saship_structure_eval": (ct.c_int, [ct.c_int32, ct.c_void_p, ct.c_int32, ct.c_void_p, ct.c_int32,
                                         ct.c_void_p, ct.c_void_p]),
    "saship_probe_fma": (ct.c_int, [ct.c_int32, ct.POINTER(ct.c_double)]),
    "saship_host_alloc": (ct.c_int, [ct.c_int64, ct.POINTER(ct.c_void_p)]),
    "saship_host_free": (ct.c_int, [ct.c_void_p]),
    "saship_batch_replay": (ct.c_int, [ct.c_void_p, ct.c_int32, ct.POINTER(ct.c_double)]),
    "saship_batch_set_resolution": (ct.c_int, [ct.c_void_p]*7 + [ct.c_int64]),
    "saship_batch_set_resolution_repeat": (ct.c_int, [ct.c_void_p]*7 + [ct.c_int64] + [ct.c_void_p]*4 + [ct.c_int64]),
    "saship_batch_set_structure": (ct.c_int, [ct.c_void_p, ct.c_int32, ct.c_void_p, ct.c_int32]),
    "saship_kernel_set_path": (ct.c_int, [ct.c_void_p, ct.c_int32]),
    "saship_batch_set_path": (ct.c_int, [ct.c_void_p, ct.c_int32]),
    "saship_kernel_set_graph": (ct.c_int, [ct.c_void_p, ct.c_int32]),
    "saship_kernel_set_timing": (ct.c_int, [ct.c_void_p, ct.c_int32]),
    "saship_kernel_get_timing": (ct.c_int, [ct.c_void_p, c_double_p, c_int32_p]),
    "saship_kernel_release": (ct.c_int, [ct.c_void_p]),
    "saship_batch_create": (ct.c_int, [ct.c_void_p, ct.c_int32, ct.c_void_p, ct.c_void_p, ct.c_int32,
                                       ct.POINTER(ct.c_void_p)]),
    "saship_batch_call": (ct.c_int, [ct.c_void_p, ct.c_int32, ct.c_void_p, ct.c_int32, ct.c_void_p,
                                     ct.c_void_p, ct.c_double, ct.c_int32, ct.c_void_p, ct.c_void_p]),
    "saship_batch_set_options": (ct.c_int, [ct.c_void_p, ct.c_int32, ct.c_int32, ct.c_int32]),
    "saship_batch_get_options": (ct.c_int, [ct.c_void_p, c_int32_p, c_int32_p, c_int32_p]),
    "saship_batch_autotune": (ct.c_int, [ct.c_void_p, ct.c_int32, ct.c_void_p, ct.c_int32, ct.c_void_p,
                                         ct.c_void_p, ct.c_double, c_int32_p, c_int32_p]),
    "saship_batch_set_timing": (ct.c_int, [ct.c_void_p, ct.c_int32]),
    "saship_batch_get_timing": (ct.c_int, [ct.c_void_p, c_double_p, c_int32_p]),
    "saship_batch_release": (ct.c_int, [ct.c_void_p]),
    "saship_mlp_sizes": (ct.c_int, [c_int32_p, c_int32_p, ct.POINTER(ct.c_int64), ct.c_int32]),
    "saship_mlp_pack": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_void_p]),
    "saship_mlp_forward": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_int32, ct.c_void_p,
                                      ct.c_void_p, ct.c_void_p]),
    "saship_mlp_train_step": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_int32,
                                         ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p]),
    "saship_mlp_infer": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_int32, ct.c_void_p,
                                    ct.c_int32, ct.c_void_p, ct.c_int32, ct.c_void_p, ct.c_int32,
                                    ct.c_void_p]),
}


class HipError(RuntimeError):
    """An error reported by libsaship."""


def _share_hip_runtime():
    """
    Keep one HIP runtime per process.  PyTorch-ROCm wheels carry their own
    libamdhip64.so (same soname, libamdhip64.so.7, as the system one this
    library is linked against), and torch loads it by file name, so
    loading libsaship.so first and torch second would put two HIP (and
    HSA) runtimes in the process; torch then sees no device and stream
    handles cannot cross between the two.  When torch is installed, load
    its runtime first (by path, RTLD_GLOBAL): libsaship.so then binds to
    it by soname, and torch's own later load finds the same file.
    """
    import importlib.util
    try:
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        spec = None
    if spec is None or not spec.origin:
        return
    path = os.path.join(os.path.dirname(spec.origin), "lib", "libamdhip64.so")
    if os.path.exists(path):
        ct.CDLL(path, mode=ct.RTLD_GLOBAL)


def lib():
    """Load libsaship.so once (raises OSError if it is missing)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIBPATH):
                raise OSError("%s not built; run __graft_entry__.build() or "
                              "python -m sasmodels_amd.build" % LIBPATH)
            _share_hip_runtime()
            handle = ct.CDLL(LIBPATH)
            for name, (restype, argtypes) in SIGNATURES.items():
                fn = getattr(handle, name)
                fn.restype = restype
                fn.argtypes = argtypes
            _lib = handle
        return _lib


def check(status):
    """Raise :class:`HipError` with the library's message on a non-zero status."""
    if status != 0:
        msg = lib().saship_last_error().decode("utf-8", "replace")
        raise HipError("libsaship error %d: %s" % (status, msg))
    return status


def device_count():
    """Number of HIP devices the library sees (0 without a GPU)."""
    n = ct.c_int32(0)
    check(lib().saship_device_count(ct.byref(n)))
    return n.value
