"""
ctypes binding of the FP64 CPU oracle (liboracle.so, built from sasoracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, always as the checker.  The product package
(sasmodels_amd) never imports this module.

``OracleKernel`` has the reference Kernel surface (call with
``(call_details, values, cutoff, magnetic)``, ``Fq(...)``) and normalises
the raw accumulators the way sasmodels/kernel.py:80-188 does, written out
here so that the oracle does not depend on the code it checks.
"""

import ctypes as ct
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

KIND_IQ, KIND_IQXY, KIND_IMAGNETIC = 0, 1, 2
#: models whose 1D kernel returns <F^2> only (one output per q)
NO_FQ = (8, 9, 10) + tuple(range(22, 41))
MODEL_IDS = {"sphere": 0, "cylinder": 1, "ellipsoid": 2,
             "core_shell_cylinder": 3, "core_shell_sphere": 4,
             "parallelepiped": 6, "triaxial_ellipsoid": 7,
             "bcc_paracrystal": 8, "fcc_paracrystal": 9, "sc_paracrystal": 10,
             "hollow_cylinder": 11, "vesicle": 12, "fuzzy_sphere": 13,
             "core_multi_shell": 17, "onion": 18, "core_shell_ellipsoid": 19,
             "elliptical_cylinder": 20, "core_shell_bicelle": 21, "guinier": 22, "gaussian_peak": 23,
             "lorentz": 24, "dab": 25, "gel_fit": 26, "mass_fractal": 27, "surface_fractal": 28,
             "mass_surface_fractal": 29, "star_polymer": 30, "mono_gauss_coil": 31, "fractal": 32,
             "fractal_core_shell": 33, "lamellar_hg": 34, "lamellar_stack_caille": 35,
             "lamellar_hg_stack_caille": 36, "lamellar_stack_paracrystal": 37, "linear_pearls": 38,
             "polymer_micelle": 39, "raspberry": 40}

_lib = None
_D = ct.c_double
_P = ct.POINTER(ct.c_double)


def build():
    """Compile liboracle.so (and oracle/_ref where the reference tree exists)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def lib():
    """Load liboracle.so, building it first if needed."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            build()
        L = ct.CDLL(LIBPATH)
        L.sasoracle_kernel.restype = ct.c_int
        L.sasoracle_kernel.argtypes = [ct.c_int, ct.c_int, ct.c_int, ct.c_int32, ct.c_int32,
                                       ct.c_int32, ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                       ct.c_void_p, _D, ct.c_int32]
        L.sasoracle_set_threads.argtypes = [ct.c_int]
        for name in ("sasoracle_j1", "sasoracle_2j1x_x", "sasoracle_sph_j1c", "sasoracle_sinx_x"):
            getattr(L, name).restype = _D
            getattr(L, name).argtypes = [_D]
        L.sasoracle_gauss.restype = ct.c_int
        L.sasoracle_gauss.argtypes = [ct.c_int, ct.c_void_p, ct.c_void_p]
        L.sasoracle_fq.restype = ct.c_int
        L.sasoracle_fq.argtypes = [ct.c_int, ct.c_int, _D, ct.c_void_p, _P, _P]
        L.sasoracle_iqac.restype = _D
        L.sasoracle_iqac.argtypes = [ct.c_int, _D, _D, ct.c_void_p]
        L.sasoracle_form_volume.restype = _D
        L.sasoracle_form_volume.argtypes = [ct.c_int, ct.c_void_p]
        L.sasoracle_iqabc.restype = _D
        L.sasoracle_iqabc.argtypes = [ct.c_int, _D, _D, _D, ct.c_void_p]
        L.sasoracle_view_jitter_abc.argtypes = [_D]*6 + [ct.c_void_p]
        L.sasoracle_hardsphere.restype = _D
        L.sasoracle_hardsphere.argtypes = [_D, _D, _D]
        for name in ("sasoracle_stickyhardsphere", "sasoracle_squarewell"):
            getattr(L, name).restype = _D
            getattr(L, name).argtypes = [_D]*5
        L.sasoracle_hayter_msa.restype = _D
        L.sasoracle_hayter_msa.argtypes = [_D]*7
        L.sasoracle_shell_volume.restype = _D
        L.sasoracle_shell_volume.argtypes = [ct.c_int, ct.c_void_p]
        L.sasoracle_radius_effective.restype = _D
        L.sasoracle_radius_effective.argtypes = [ct.c_int, ct.c_int, ct.c_void_p]
        L.sasoracle_view_jitter_row.argtypes = [_D, _D, _D, _D, _P, _P]
        L.sasoracle_magnetic_sld.restype = _D
        L.sasoracle_magnetic_sld.argtypes = [ct.c_int, _D, _D, _D, _D, _D, _D, _D, _D]
        L.sasoracle_spin_weights.argtypes = [_D, _D, ct.c_void_p]
        _lib = L
    return _lib


def set_threads(n):
    """OpenMP threads used over q (results do not depend on it)."""
    lib().sasoracle_set_threads(int(n))


def raw_call(model, kind, quad_n, q, details, values, cutoff=0.0, reff_mode=0,
             pd_start=0, pd_stop=None, result=None):
    """
    One reference-shaped kernel call (kernel_iq.c:317-328).  *q* is [nq] for
    the 1D kernel or [nq, 2] for 2D; returns the raw result vector.
    """
    mid = MODEL_IDS[model] if isinstance(model, str) else int(model)
    q = np.ascontiguousarray(q, "d")
    nq = q.shape[0]
    details = np.ascontiguousarray(details, "i4")
    values = np.ascontiguousarray(values, "d")
    nout = 2 if kind == KIND_IQ and mid not in NO_FQ else 1
    if result is None:
        result = np.zeros(nout*nq + 4, "d")
    stop = int(details[-4]) if pd_stop is None else int(pd_stop)
    rc = lib().sasoracle_kernel(mid, kind, int(quad_n), nq, int(pd_start), stop,
                                details.ctypes.data, values.ctypes.data, q.ctypes.data,
                                result.ctypes.data, float(cutoff), int(reff_mode))
    if rc != 0:
        raise ValueError("sasoracle_kernel failed with %d" % rc)
    return result


FLOPCOUNT_PATH = os.path.join(HERE, "libflopcount.so")
_flib = None
#: categories tallied by libflopcount.so, in its order
FLOP_CATEGORIES = ("add", "mul", "div", "sqrt", "trans")


def flopcount_lib():
    """Load libflopcount.so (the oracle on a counting scalar), building it first if needed."""
    global _flib
    if _flib is None:
        if not os.path.exists(FLOPCOUNT_PATH):
            build()
        L = ct.CDLL(FLOPCOUNT_PATH)
        L.flopcount_kernel.restype = ct.c_int
        L.flopcount_kernel.argtypes = [ct.c_int, ct.c_int, ct.c_int, ct.c_int32, ct.c_int32, ct.c_int32,
                                       ct.c_void_p, ct.c_void_p, ct.c_int32, ct.c_void_p, ct.c_void_p,
                                       ct.c_int32, _D, ct.c_int32, ct.c_void_p]
        L.flopcount_set_threads.argtypes = [ct.c_int]
        _flib = L
    return _flib


def count_flops(model, kind, quad_n, q, details, values, cutoff=0.0, reff_mode=0, threads=1):
    """
    Operations the reference formulas take for one whole-mesh kernel call
    (sasoracle.c on a counting scalar): dict of the FLOP_CATEGORIES tallies,
    "flops" (their sum: add, sub, mul, div, sqrt and each transcendental
    count 1) and "result" (the raw result, equal to raw_call's).
    """
    mid = MODEL_IDS[model] if isinstance(model, str) else int(model)
    q = np.ascontiguousarray(q, "d")
    nq = q.shape[0]
    details = np.ascontiguousarray(getattr(details, "buffer", details), "i4")
    values = np.ascontiguousarray(values, "d")
    nout = 2 if kind == KIND_IQ and mid not in NO_FQ else 1
    result = np.zeros(nout*nq + 4, "d")
    counts = np.zeros(len(FLOP_CATEGORIES), np.uint64)
    L = flopcount_lib()
    L.flopcount_set_threads(int(threads))
    rc = L.flopcount_kernel(mid, kind, int(quad_n), nq, 0, int(details[-4]), details.ctypes.data,
                            values.ctypes.data, values.size, q.ctypes.data, result.ctypes.data, result.size,
                            float(cutoff), int(reff_mode), counts.ctypes.data)
    if rc != 0:
        raise ValueError("flopcount_kernel failed with %d" % rc)
    out = {k: int(v) for k, v in zip(FLOP_CATEGORIES, counts)}
    out["flops"] = sum(out[k] for k in FLOP_CATEGORIES)
    out["result"] = result
    return out


def normalise(result, nq, nout):
    """(F1, F2, R_eff, V_shell, V_form/V_shell) from raw accumulators (kernel.py:170-188)."""
    total = result[nout*nq]
    if total == 0.:
        total = 1.
    form = result[nout*nq + 1]/total
    shell = result[nout*nq + 2]/total
    reff = result[nout*nq + 3]/total
    if shell == 0.:
        shell = 1.
    F1 = result[1:nout*nq:nout]/total if nout == 2 else None
    F2 = result[0:nout*nq:nout]/total
    return F1, F2, reff, shell, form/shell


class OracleKernel:
    """
    CPU FP64 kernel with the reference Kernel call surface.  *info* needs
    ``id`` and ``integration_size``; *q_vectors* is [q] or [qx, qy].
    """
    def __init__(self, info, q_vectors, chunk=None):
        self.info = info
        self.dim = "2d" if len(q_vectors) == 2 else "1d"
        self.dtype = np.dtype("d")
        if self.dim == "2d":
            self.q = np.column_stack([np.asarray(v, "d").ravel() for v in q_vectors])
        else:
            self.q = np.asarray(q_vectors[0], "d").ravel()
        self.nq = self.q.shape[0]
        self.chunk = chunk

    def raw(self, call_details, values, cutoff, magnetic, reff_mode=0):
        kind = KIND_IQ if self.dim == "1d" else (KIND_IMAGNETIC if magnetic else KIND_IQXY)
        quad = int(getattr(self.info, "integration_size", 0) or 0)
        buf = np.ascontiguousarray(getattr(call_details, "buffer", call_details), "i4")
        num_eval = int(buf[-4])
        if not self.chunk:
            return raw_call(self.info.id, kind, quad, self.q, buf, values, cutoff, reff_mode)
        # mesh split in pieces, accumulated as the reference's drivers do
        result = None
        for start in range(0, max(num_eval, 1), self.chunk):
            stop = min(start + self.chunk, num_eval)
            result = raw_call(self.info.id, kind, quad, self.q, buf, values, cutoff,
                              reff_mode, start, stop, result)
        return result

    def Fq(self, call_details, values, cutoff, magnetic, radius_effective_mode=0):
        nout = 2 if self.dim == "1d" and getattr(self.info, "have_Fq", True) else 1
        res = self.raw(call_details, values, cutoff, magnetic, radius_effective_mode)
        return normalise(res, self.nq, nout)

    def Iq(self, call_details, values, cutoff, magnetic):
        _, F2, _, shell, _ = self.Fq(call_details, values, cutoff, magnetic, 0)
        return values[0]/shell*F2 + values[1]
    __call__ = Iq


def hardsphere(q, radius_effective, volfraction):
    """Hard-sphere S(q) (sasmodels/models/hardsphere.c), q scalar or array."""
    f = lib().sasoracle_hardsphere
    q = np.asarray(q, "d")
    return np.array([f(float(v), float(radius_effective), float(volfraction)) for v in q.ravel()]
                    ).reshape(q.shape)


def structure(name, q, radius_effective, volfraction, *extra):
    """S(q) of structure factor *name* ("hardsphere", "stickyhardsphere",
    "squarewell", "hayter_msa") with its further parameters *extra*; q
    scalar or array."""
    q = np.asarray(q, "d")
    if name == "hardsphere":
        return hardsphere(q, radius_effective, volfraction)
    f = getattr(lib(), "sasoracle_" + name)
    return np.array([f(float(v), float(radius_effective), float(volfraction), *map(float, extra))
                     for v in q.ravel()]).reshape(q.shape)


def product(q, fq, scale, background, radius_effective, volfraction, beta_mode, er_mode,
            volfrac_in_p=False, s_name="hardsphere", s_extra=()):
    """
    P@S intensity from P's normalised (F, F^2, R_eff, V_shell, V_form/V_shell)
    (sasmodels/product.py:425-528): S at P's R_eff (er_mode > 0) or
    *radius_effective*, volume fraction volfraction * V_form/V_shell;
    I = scale*volfraction/V_shell * PS + background with PS = F^2 S, or
    F^2 + F1^2 (S - 1) with beta; when P has its own volfraction (already in
    its F) the number density is scale/V_shell.  S is structure factor
    *s_name* with further parameters *s_extra*.  *q* is the 1D q or the |q|
    of 2D points.  Returns (I, S).
    """
    F1, F2, reff, shell, ratio = fq
    radius = reff if er_mode > 0 else radius_effective
    S = structure(s_name, q, radius, volfraction*ratio, *s_extra)
    PS = F2 + F1**2*(S - 1.0) if beta_mode else F2*S
    density = scale/shell if volfrac_in_p else scale*volfraction/shell
    return density*PS + background, S


def ref_lib(name):
    """ctypes handle of oracle/_ref/<name>.so (reference sources), or None."""
    path = os.path.join(REF_DIR, name + ".so")
    if not os.path.exists(path):
        return None
    L = ct.CDLL(path)
    if hasattr(L, "ref_sq"):
        L.ref_sq.restype = _D
        L.ref_sq.argtypes = [_D, _D, _D]
        return L
    if hasattr(L, "ref_sq4"):
        L.ref_sq4.restype = _D
        L.ref_sq4.argtypes = [_D]*5
        return L
    if hasattr(L, "ref_sq6"):
        L.ref_sq6.restype = _D
        L.ref_sq6.argtypes = [_D]*7
        return L
    if hasattr(L, "ref_iq"):
        L.ref_iq.restype = _D
        L.ref_iq.argtypes = [_D, ct.c_void_p]
        if hasattr(L, "ref_iqabc"):
            L.ref_iqabc.restype = _D
            L.ref_iqabc.argtypes = [_D, _D, _D, ct.c_void_p]
        if hasattr(L, "ref_reff"):
            L.ref_reff.restype = _D
            L.ref_reff.argtypes = [ct.c_int, ct.c_void_p]
        L.ref_form_volume.restype = _D
        L.ref_form_volume.argtypes = [ct.c_void_p]
        return L
    L.ref_fq.argtypes = [_D, ct.c_void_p, _P, _P]
    L.ref_form_volume.restype = _D
    L.ref_form_volume.argtypes = [ct.c_void_p]
    L.ref_reff.restype = _D
    L.ref_reff.argtypes = [ct.c_int, ct.c_void_p]
    if hasattr(L, "ref_shell_volume"):
        L.ref_shell_volume.restype = _D
        L.ref_shell_volume.argtypes = [ct.c_void_p]
    if hasattr(L, "ref_iqac"):
        L.ref_iqac.restype = _D
        L.ref_iqac.argtypes = [_D, _D, ct.c_void_p]
    if hasattr(L, "ref_iqabc"):
        L.ref_iqabc.restype = _D
        L.ref_iqabc.argtypes = [_D, _D, _D, ct.c_void_p]
    for name in ("ref_j1", "ref_2j1x_x", "ref_sph_j1c", "ref_sinx_x"):
        if hasattr(L, name):
            getattr(L, name).restype = _D
            getattr(L, name).argtypes = [_D]
    if hasattr(L, "ref_gauss"):
        L.ref_gauss.restype = ct.c_int
        L.ref_gauss.argtypes = [ct.c_int, ct.c_void_p, ct.c_void_p]
    return L
