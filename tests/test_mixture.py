"""
Mixture models (A+B, A*B) on the host, checked without a GPU.

* Parameter tables follow the reference's naming and order
  (core.py:452-521 test_composite_order / test_composite).
* MixtureKernel's split of a mixture call into part calls
  (mixture.py:250-339) is checked by running it over oracle kernels and
  comparing with each part evaluated on its own from a parameter dict:
  I = scale (sum of A_scale I_A ... or product of I_A ...) + background.
"""

import numpy as np
import pytest

from oracle import oracle
from sasmodels_amd import core, direct_model, mixture, modelinfo, product
from sasmodels_amd.details import make_kernel_args
from test_product import _OracleP, _OracleS, oracle_product


def _names(model_string):
    info = core.load_model_info(model_string)
    return [p.name for p in info.parameters.kernel_parameters]


def test_composite_names():
    a = ("sld", "sld_solvent", "radius", "length", "theta", "phi",
         "radius_effective", "volfraction", "structure_factor_mode", "radius_effective_mode")
    b = ("sld", "sld_solvent", "radius")
    assert _names("cylinder@hardsphere*sphere") == ["A_" + p for p in a] + ["B_" + p for p in b]
    assert _names("sphere+cylinder")[:5] == ["A_scale", "A_sld", "A_sld_solvent", "A_radius", "B_scale"]
    assert _names("sphere*cylinder+ellipsoid")[0] == "AB_scale"
    assert "C_scale" in _names("sphere*cylinder+ellipsoid")
    # a second product part moves to the next free letters
    assert _names("sphere*cylinder+ellipsoid*core_shell_sphere")[10:12] == ["CD_scale", "C_sld"]


@pytest.mark.parametrize("fst,snd", [
    ("cylinder+sphere", "sphere+cylinder"),
    ("cylinder*sphere", "sphere*cylinder"),
    ("cylinder@hardsphere*sphere", "sphere*cylinder@hardsphere"),
    ("ellipsoid+sphere*cylinder@hardsphere", "sphere*cylinder@hardsphere+ellipsoid"),
    ("ellipsoid+cylinder@hardsphere*sphere", "cylinder@hardsphere*sphere+ellipsoid"),
    ("ellipsoid+sphere*cylinder@hardsphere", "ellipsoid+cylinder@hardsphere*sphere"),
])
def test_composite_order(fst, snd):
    """The same parameters whatever the order (core.py:452-505)."""
    strip = lambda names: sorted([c for c in n if c == c.lower()] for n in names)
    assert strip(_names(fst)) == strip(_names(snd))


def oracle_kernel(info, q_vectors):
    """A kernel for *info* whose plain models run on the oracle."""
    if info.composition is None:
        k = oracle.OracleKernel(info, q_vectors)
        k.dim = "2d" if len(q_vectors) == 2 else "1d"
        return k
    kind, parts = info.composition
    if kind == "product":
        pk = _OracleP(parts[0].id, q_vectors[0])
        pk.q_input = None
        return product.ProductKernel(info, pk, _OracleS(q_vectors[0]), q_vectors)
    return mixture.MixtureKernel(info, [oracle_kernel(p, q_vectors) for p in parts], q_vectors)


def _direct(name, q_vectors, pars, dim="1d"):
    info = modelinfo.make_model_info(name)
    k = oracle.OracleKernel(info, q_vectors)
    kl = type("K", (), {"info": info, "dtype": np.dtype("d")})()
    det, val, mag = make_kernel_args(kl, direct_model.get_mesh(info, pars, dim=dim))
    return k(det, val, 0.0, mag)


def _mix_call(info, q_vectors, pars, dim="1d"):
    kern = oracle_kernel(info, q_vectors)
    kl = type("K", (), {"info": info, "dtype": np.dtype("d")})()
    det, val, mag = make_kernel_args(kl, direct_model.get_mesh(info, pars, dim=dim))
    return kern(det, val, 0.0, mag), kern


Q = np.logspace(-3, -0.5, 60)


def test_sum_of_parts():
    info = core.load_model_info("sphere+cylinder")
    pars = dict(scale=1.5, background=0.1, A_scale=0.3, A_radius=30.0, A_radius_pd=0.1, A_radius_pd_n=11,
                B_scale=2.0, B_radius=20.0, B_length=300.0, B_length_pd=0.1, B_length_pd_n=7)
    got, kern = _mix_call(info, [Q], pars)
    a = _direct("sphere", [Q], dict(scale=0.3, background=0.0, radius=30.0, radius_pd=0.1, radius_pd_n=11))
    b = _direct("cylinder", [Q], dict(scale=2.0, background=0.0, radius=20.0, length=300.0,
                                      length_pd=0.1, length_pd_n=7))
    np.testing.assert_allclose(got, 1.5*(a + b) + 0.1, rtol=1e-13)
    parts = kern.results()
    assert list(parts) == ["A_Sphere(Q)", "B_Cylinder(Q)"]
    np.testing.assert_allclose(parts["A_Sphere(Q)"][1], a, rtol=1e-13)


def test_product_of_parts_with_structure_factor():
    info = core.load_model_info("cylinder@hardsphere*sphere")
    pars = dict(scale=2.0, background=0.01, A_radius=25.0, A_volfraction=0.3, A_structure_factor_mode=1,
                A_radius_effective_mode=1, A_radius_pd=0.1, A_radius_pd_n=9, B_radius=15.0)
    got, _ = _mix_call(info, [Q], pars)
    # the P@S part by the oracle product from its parameter dict
    a = oracle_product("cylinder", Q, dict(scale=1.0, background=0.0, radius=25.0, volfraction=0.3,
                                           structure_factor_mode=1, radius_effective_mode=1,
                                           radius_pd=0.1, radius_pd_n=9))[0]
    b = _direct("sphere", [Q], dict(scale=1.0, background=0.0, radius=15.0))
    np.testing.assert_allclose(got, 2.0*a*b + 0.01, rtol=1e-12)


def test_sum_with_product_part_and_magnetism_2d():
    info = core.load_model_info("sphere*core_shell_sphere+sphere")
    x = np.linspace(-0.1, 0.1, 9)
    qx, qy = [v.ravel() for v in np.meshgrid(x, x + 0.003)]
    pars = dict(scale=0.7, background=0.02, AB_scale=1.3, C_scale=0.4,
                A_radius=40.0, A_sld_M0=1.5, A_sld_mtheta=30.0, B_radius=30.0, B_thickness=8.0,
                B_sld_shell_M0=2.0, B_sld_shell_mphi=45.0, C_radius=20.0, C_sld_solvent_M0=0.5,
                up_frac_i=0.2, up_frac_f=0.7, up_theta=30.0)
    got, _ = _mix_call(info, [qx, qy], pars, dim="2d")
    spin = dict(up_frac_i=0.2, up_frac_f=0.7, up_theta=30.0)
    a = _direct("sphere", [qx, qy], dict(spin, scale=1.3, background=0.0, radius=40.0, sld_M0=1.5,
                                         sld_mtheta=30.0), dim="2d")
    b = _direct("core_shell_sphere", [qx, qy], dict(spin, scale=1.0, background=0.0, radius=30.0,
                                                    thickness=8.0, sld_shell_M0=2.0, sld_shell_mphi=45.0),
                dim="2d")
    c = _direct("sphere", [qx, qy], dict(spin, scale=0.4, background=0.0, radius=20.0,
                                         sld_solvent_M0=0.5), dim="2d")
    np.testing.assert_allclose(got, 0.7*(a*b + c) + 0.02, rtol=1e-12)


def test_product_restarts_on_exact_zero():
    """mixture.py:226-230: a running product holding an exact zero is replaced
    by the next part's I (a reference behaviour kept for parity)."""
    info = core.load_model_info("sphere*sphere")
    pars = dict(scale=1.0, background=0.0, A_sld=1.0, A_sld_solvent=1.0, B_radius=30.0)
    got, _ = _mix_call(info, [Q], pars)
    b = _direct("sphere", [Q], dict(scale=1.0, background=0.0, radius=30.0))
    np.testing.assert_allclose(got, b, rtol=1e-13)
