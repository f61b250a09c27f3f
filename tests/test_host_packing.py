"""
Host side of the boundary: distributions, mesh assembly, the CallDetails
buffer, the values vector and the Fq normalisation (sasmodels/weights.py,
direct_model.py:97-154, details.py:183-306, kernel.py:101-188).
"""

import numpy as np
import pytest
from scipy import stats
from scipy.spatial.transform import Rotation

from sasmodels_amd import core, details, direct_model, generate, modelinfo, weights
from sasmodels_amd.kernel import Kernel


class _Stub:
    """Object with the attributes make_kernel_args reads."""
    def __init__(self, name, dim="1d"):
        self.info = modelinfo.make_model_info(name)
        self.dtype = np.dtype("d")
        self.dim = dim


# ---------------------------------------------------------------- distributions
def test_gaussian_points_and_shape():
    x, w = weights.get_weights("gaussian", 35, 0.1, 3, 50.0, (0, np.inf), True)
    np.testing.assert_allclose(x, 50 + np.linspace(-15, 15, 35), rtol=1e-15)
    pdf = stats.norm.pdf(x, loc=50, scale=5)
    np.testing.assert_allclose(w, pdf/pdf.sum(), rtol=1e-13)


def test_lognormal_shape():
    x, w = weights.get_weights("lognormal", 35, 0.1, 3, 20.0, (0, np.inf), True)
    np.testing.assert_allclose(x, 20*(1 + np.linspace(-0.3, 0.3, 35)), rtol=1e-15)
    pdf = stats.lognorm.pdf(x, s=0.1, scale=20.0)
    np.testing.assert_allclose(w, pdf/pdf.sum(), rtol=1e-12)


def test_schulz_shape_is_gamma():
    c, sig = 40.0, 0.2
    x, w = weights.get_weights("schulz", 25, sig, 3, c, (0, np.inf), True)
    z = 1/sig**2
    pdf = stats.gamma.pdf(x, a=z, scale=c/z)
    np.testing.assert_allclose(w, pdf/pdf.sum(), rtol=1e-11)


def test_boltzmann_uniform_rectangle():
    x, w = weights.get_weights("boltzmann", 21, 0.2, 3, 10.0, (0, np.inf), True)
    pdf = stats.laplace.pdf(x, loc=10, scale=2.0)
    np.testing.assert_allclose(w, pdf/pdf.sum(), rtol=1e-13)
    x, w = weights.get_weights("uniform", 11, 0.1, 3, 10.0, (0, np.inf), True)
    np.testing.assert_allclose(x, np.linspace(9, 11, 11))
    np.testing.assert_allclose(w, np.full(11, 1/11))
    x, w = weights.get_weights("rectangle", 35, 0.1, 1.73205, 10.0, (0, np.inf), True)
    assert np.all(np.abs(x - 10) <= 0.1*10*np.sqrt(3)) and np.allclose(w, 1/len(w))


def test_limits_truncate_and_orientation_is_absolute():
    x, w = weights.get_weights("gaussian", 11, 0.5, 3, 10.0, (0, np.inf), True)
    assert x.min() >= 0 and len(x) < 11
    x, w = weights.get_weights("gaussian", 5, 10.0, 3, 60.0, (-90, 90), False)
    np.testing.assert_allclose(x, np.linspace(-30, 30, 5))      # centred on 0, absolute width
    x, w = weights.get_weights("gaussian", 35, 0.0, 3, 7.0, (0, np.inf), True)
    assert list(x) == [7.0] and list(w) == [1.0]
    with pytest.raises(NotImplementedError):
        weights.get_weights("array", 3, 0.1, 3, 1.0, (0, 1), True)


# ---------------------------------------------------------------- mesh and details
def test_get_mesh_activity_rules():
    info = modelinfo.make_model_info("cylinder")
    pars = dict(radius_pd=0.1, radius_pd_n=5, theta_pd=10, theta_pd_n=3, theta=30)
    mesh1 = direct_model.get_mesh(info, pars, dim="1d")
    names = [p.name for p in info.parameters.call_parameters]
    assert len(mesh1[names.index("radius")][1]) == 5
    th = mesh1[names.index("theta")]
    assert th[0] == 30 and list(th[1]) == [0.0] and list(th[2]) == [1.0]   # inactive in 1D
    mesh2 = direct_model.get_mesh(info, pars, dim="2d")
    assert len(mesh2[names.index("theta")][1]) == 3
    with pytest.raises(TypeError):
        direct_model.get_mesh(info, dict(radius_pdd=3))
    meshm = direct_model.get_mesh(info, pars, mono=True)
    assert all(len(m[1]) == 1 for m in meshm)


def test_call_details_layout():
    k = _Stub("cylinder")
    pars = dict(radius_pd=0.1, radius_pd_n=5, length_pd=0.1, length_pd_n=3)
    mesh = direct_model.get_mesh(k.info, pars)
    cd, values, mag = details.make_kernel_args(k, mesh)
    mp = k.info.parameters.max_pd
    assert mp == 4 and cd.buffer.dtype == np.int32 and cd.buffer.size == 4*mp + 4
    assert list(cd.pd_par[:2]) == [2, 3]                    # radius (5), length (3) first
    assert list(cd.pd_length[:2]) == [5, 3]
    assert list(cd.pd_stride[:2]) == [1, 5]
    assert cd.num_eval == 15 and cd.num_active == 2 and cd.theta_par == 4
    lengths = [len(m[2]) for m in mesh[2:8]]
    assert cd.num_weights == sum(lengths)
    nv = k.info.parameters.nvalues
    assert nv == 18 and values.size % 32 == 0
    np.testing.assert_array_equal(values[:nv], [m[0] for m in mesh])
    off = nv
    for m in mesh[2:8]:
        np.testing.assert_array_equal(values[off:off + len(m[1])], m[1])
        off += len(m[1])
    for m in mesh[2:8]:
        np.testing.assert_array_equal(values[off:off + len(m[2])], m[2])
        off += len(m[2])
    assert not values[off:].any() and not mag
    # offsets point at each parameter's own distribution
    for lev in range(2):
        par, o, n = cd.pd_par[lev], cd.pd_offset[lev], cd.pd_length[lev]
        np.testing.assert_array_equal(values[nv + o:nv + o + n], mesh[2 + par][1])


@pytest.mark.parametrize("name,nvalues,max_pd", [
    ("sphere", 15, 1), ("cylinder", 18, 4), ("ellipsoid", 18, 4),
    ("core_shell_cylinder", 23, 5), ("core_shell_sphere", 20, 2)])
def test_parameter_table_sizes(name, nvalues, max_pd):
    info = modelinfo.make_model_info(name)
    assert info.parameters.nvalues == nvalues and info.parameters.max_pd == max_pd


def test_too_many_polydisperse_parameters():
    k = _Stub("sphere")
    mesh = direct_model.get_mesh(k.info, dict(radius_pd=0.1, radius_pd_n=3))
    mesh[2] = (mesh[2][0], np.array([1., 2.]), np.array([.5, .5]))   # sld made polydisperse
    with pytest.raises(ValueError):
        details.make_kernel_args(k, mesh)


def test_convert_magnetism_matches_rotation():
    k = _Stub("core_shell_sphere", "2d")
    pars = dict(sld_core_M0=2.0, sld_core_mtheta=35.0, sld_core_mphi=120.0,
                sld_shell_M0=0.5, sld_shell_mtheta=-70.0, sld_shell_mphi=15.0)
    mesh = direct_model.get_mesh(k.info, pars, dim="2d")
    _, values, mag = details.make_kernel_args(k, mesh)
    assert mag
    base = 2 + k.info.parameters.npars + 4
    for j, (m0, th, ph) in enumerate([(2.0, 35.0, 120.0), (0.5, -70.0, 15.0), (0, 0, 0)]):
        want = m0*Rotation.from_euler("yz", [th, ph], degrees=True).apply([0, 0, 1.0])
        np.testing.assert_allclose(values[base + 3*j:base + 3*j + 3], want, atol=1e-15)
    _, values0, mag0 = details.make_kernel_args(k, direct_model.get_mesh(k.info, {}, dim="2d"))
    assert not mag0


# ---------------------------------------------------------------- normalisation
class _FakeKernel(Kernel):
    def __init__(self, result, nq, dim="1d"):
        self.info = modelinfo.make_model_info("sphere")
        self.dim = dim
        self.q_input = type("Q", (), {"nq": nq})()
        self.result = np.asarray(result, "d")

    def _call_kernel(self, *args):
        pass


def test_fq_normalisation_and_guards():
    k = _FakeKernel([4.0, 2.0, 9.0, 3.0, 2.0, 10.0, 10.0, 6.0], 2)
    F1, F2, R, V, Vr = k.Fq(None, np.array([1.0, 0.0]), 0, False, 1)
    np.testing.assert_allclose(F2, [2.0, 4.5])
    np.testing.assert_allclose(F1, [1.0, 1.5])
    assert (R, V, Vr) == (3.0, 5.0, 1.0)
    I = k.Iq(None, np.array([2.0, 0.5]), 0, False)
    np.testing.assert_allclose(I, 2.0/5.0*F2 + 0.5)
    z = _FakeKernel([0.0]*8, 2)
    F1, F2, R, V, Vr = z.Fq(None, np.array([1.0, 0.0]), 0, False)
    assert V == 1.0 and Vr == 0.0 and not F2.any()


def test_backend_selection_errors():
    with pytest.raises(ValueError):
        core.build_model(core.load_model_info("sphere"), platform="dll")
    with pytest.raises(ValueError):
        core.load_model_info("sphere@cylinder")
    with pytest.raises(ValueError):
        core.load_model_info("barbell+sphere")
    with pytest.raises(ValueError):
        modelinfo.make_model_info("barbell")
    info = core.load_model_info("cylinder")
    generate.set_integration_size(info, 150)
    assert info.integration_size == 150 and core.list_models()[1] == "cylinder"
