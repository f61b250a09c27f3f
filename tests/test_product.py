"""
P@S products on the host, checked without a GPU.

* The oracle's hard-sphere S(q) is pinned bit for bit against the
  reference's hardsphere.c compiled into oracle/_ref, against the
  reference's known answers (hardsphere.py:107-111), and against what the
  mathematics fixes: the q -> 0 compressibility limit (1 - phi)^4/(1 + 2 phi)^2
  and a direct quadrature of the Fourier transform of the Percus-Yevick
  direct correlation function.
* The oracle's product (oracle.product) is pinned against every "@S" known
  answer of the sphere, cylinder and ellipsoid model files, intermediates
  included.
* The product parameter table follows the reference's layout
  (product.py:171-245), and ProductKernel's split of a P@S call into P and
  S calls (product.py:425-528) reproduces the known answers when its P and S
  kernels are oracle kernels.
"""

import numpy as np
import pytest
from scipy import integrate

from kat_tables import HARDSPHERE, HAYTER_MSA, PRODUCT_KATS, SQUAREWELL, STICKYHARDSPHERE, is_near
from oracle import oracle
from sasmodels_amd import core, direct_model, modelinfo, product
from sasmodels_amd.details import make_kernel_args


# ------------------------------------------------------------ hard-sphere S(q)
def test_hardsphere_matches_reference_source():
    ref = oracle.ref_lib("ref_hardsphere")
    if ref is None:
        pytest.skip("oracle/_ref/ref_hardsphere.so not built (reference tree absent)")
    rng = np.random.default_rng(3)
    f = oracle.lib().sasoracle_hardsphere
    cases = [(10**rng.uniform(-6, 0), 10**rng.uniform(-1, 3), rng.uniform(0, 0.7))
             for _ in range(5000)]
    cases += [(0.0, 50.0, 0.2), (0.1, 0.0, 0.3), (0.1, -40.0, 0.3), (2.5e-7, 10.0, 0.1),
              (0.0025, 10.0, 0.4), (0.1, 50.0, 0.0)]
    bad = [c for c in cases if f(*c) != ref.ref_sq(*c)]
    assert not bad, bad[:5]


def test_hardsphere_kat():
    pars, q, want = HARDSPHERE[0]
    got = oracle.hardsphere(q, pars["radius_effective"], pars["volfraction"])
    assert all(is_near(w, g) for w, g in zip(want, got)), (want, got)


def _py_sq(q, R, phi):
    """1/(1 - n c(q)) with c(r) the Percus-Yevick direct correlation function
    (Wertheim), its Fourier transform by adaptive quadrature."""
    sigma = 2.0*R
    l1 = (1 + 2*phi)**2/(1 - phi)**4
    l2 = -(1 + phi/2)**2/(1 - phi)**4
    c = lambda r: -(l1 + 6*phi*l2*(r/sigma) + 0.5*phi*l1*(r/sigma)**3)
    n = 6*phi/(np.pi*sigma**3)
    integrand = lambda r: c(r)*r*r*np.sinc(q*r/np.pi)
    nc = 4*np.pi*n*integrate.quad(integrand, 0, sigma, epsabs=0, epsrel=1e-11, limit=200)[0]
    return 1.0/(1.0 - nc)


@pytest.mark.parametrize("phi", [0.05, 0.2, 0.45])
def test_hardsphere_against_correlation_integral(phi):
    R = 40.0
    for q in [1e-4, 5e-4, 2e-3, 0.01, 0.03, 0.08, 0.2, 0.6]:
        got = oracle.hardsphere(q, R, phi)
        assert got == pytest.approx(_py_sq(q, R, phi), rel=1e-7), q


def test_hardsphere_forward_limit_and_zero_radius():
    for phi in (0.01, 0.3, 0.6):
        assert oracle.hardsphere(0.0, 30.0, phi) == pytest.approx((1 - phi)**4/(1 + 2*phi)**2, rel=1e-14)
    assert oracle.hardsphere(0.3, 0.0, 0.4) == 1.0
    # no interaction without particles
    np.testing.assert_allclose(oracle.hardsphere(np.logspace(-3, 0, 20), 50.0, 0.0), 1.0, rtol=1e-14)


# ------------------------------------------------------------ product info
def test_product_parameter_layout():
    info = core.load_model_info("cylinder@hardsphere")
    names = [p.id for p in info.parameters.call_parameters]
    assert names[:12] == ["scale", "background", "sld", "sld_solvent", "radius", "length",
                          "theta", "phi", "radius_effective", "volfraction",
                          "structure_factor_mode", "radius_effective_mode"]
    assert names[12:16] == ["up_frac_i", "up_frac_f", "up_theta", "up_phi"]
    assert info.parameters.max_pd == 4 and info.parameters.nvalues == 22
    mode = info.parameters["radius_effective_mode"]
    assert mode.default == 1 and mode.limits == (0, 7) and mode.choices[0] == "unconstrained"
    assert info.parameters["structure_factor_mode"].choices == ["P*S", "P*(1+beta*(S-1))"]
    assert info.composition[0] == "product"
    assert core.load_model_info("sphere@hardsphere").parameters.nvalues == 19


def test_product_info_errors():
    with pytest.raises(ValueError):
        core.load_model_info("sphere@cylinder")
    with pytest.raises(ValueError):
        core.load_model_info("hardsphere@hardsphere")
    with pytest.raises(ValueError):
        core.load_model_info("sphere@hardsphere@hardsphere")
    s_bad = modelinfo.make_model_info("sphere")
    with pytest.raises(TypeError):
        product.make_product_info(modelinfo.make_model_info("cylinder"), s_bad)


# ------------------------------------------------------------ the oracle product
def _p_fq(p_name, q, pars):
    """P's (F, F^2, R_eff, V_shell, V_ratio) by the oracle from a parameter dict."""
    info = modelinfo.make_model_info(p_name)
    kern = oracle.OracleKernel(info, [np.asarray(q, "d")])
    kern.info = info
    kl = type("K", (), {"info": info, "dtype": np.dtype("d")})()
    det, val, mag = make_kernel_args(kl, direct_model.get_mesh(info, pars, dim="1d"))
    return kern, det, val, mag


def oracle_product(p_name, q, pars, s_name="hardsphere"):
    pars = dict(pars)
    s_info = modelinfo.make_model_info(s_name)
    s_defaults = s_info.parameters.defaults
    extra = tuple(pars.pop(p.id, p.default) for p in s_info.parameters.kernel_parameters[2:])
    scale = pars.pop("scale", 1.0)
    background = pars.pop("background", modelinfo.DEFAULT_BACKGROUND)
    volfrac = pars.pop("volfraction", s_defaults["volfraction"])
    reff = pars.pop("radius_effective", s_defaults["radius_effective"])
    beta = pars.pop("structure_factor_mode", 0) > 0
    er_mode = int(pars.pop("radius_effective_mode", 1))
    in_p = "volfraction" in modelinfo.make_model_info(p_name).parameters
    if in_p:
        # P's own volfraction (vesicle): P and S both see the product's value
        pars["volfraction"] = volfrac
    kern, det, val, mag = _p_fq(p_name, q, pars)
    fq = kern.Fq(det, val, 0.0, mag, er_mode)
    I, S = oracle.product(np.asarray(q, "d"), fq, scale, background, reff, volfrac, beta, er_mode, in_p,
                          s_name, extra)
    return I, S, fq, (scale if in_p else scale*volfrac)/fq[3]


@pytest.mark.parametrize("case", PRODUCT_KATS, ids=lambda c: "%s-%s" % (c[0], sorted(c[1].items())))
def test_oracle_product_kats(case):
    p_name, pars, q, want = case[:4]
    I, S, fq, combined = oracle_product(p_name, q, pars)
    assert all(is_near(w, g) for w, g in zip(want, I)), (want, I)
    if len(case) > 4:
        parts = case[4]
        F1, F2, reff, shell, ratio = fq
        assert all(is_near(w, g) for w, g in zip(parts["P(Q)"], combined*F2))
        assert all(is_near(w, g) for w, g in zip(parts["S(Q)"], S))
        beta = F1**2/F2
        assert all(is_near(w, g) for w, g in zip(parts["S_eff(Q)"], 1 + beta*(S - 1)))
        assert is_near(parts["volume"], shell) and is_near(parts["volume_ratio"], ratio)
        assert is_near(parts["radius_effective"], reff)


# ------------------------------------------------------------ ProductKernel's call split
class _OracleS:
    """Hard-sphere S with the Kernel surface, checking the S call it receives."""
    dim = "1d"
    dtype = np.dtype("d")

    def __init__(self, q):
        self.q = np.asarray(q, "d")

    def Iq(self, details, values, cutoff, magnetic):
        buf = details.buffer
        assert buf.size == 4 and buf[0] == 1 and buf[-1] == -1 and not magnetic
        assert values.size % 32 == 0 and values[0] == 1.0 and values[1] == 0.0
        return oracle.hardsphere(self.q, values[2], values[3])


class _OracleP:
    def __init__(self, p_name, q):
        self.info = modelinfo.make_model_info(p_name)
        self.k = oracle.OracleKernel(self.info, [np.asarray(q, "d")])
        self.dim, self.dtype = "1d", np.dtype("d")

    def Fq(self, details, values, cutoff, magnetic, radius_effective_mode=0):
        assert values[0] == 1.0 and values[1] == 0.0 and values.size % 32 == 0
        return self.k.Fq(details, values, cutoff, magnetic, radius_effective_mode)


@pytest.mark.parametrize("case", PRODUCT_KATS, ids=lambda c: "%s-%s" % (c[0], sorted(c[1].items())))
def test_product_kernel_split_reproduces_kats(case):
    p_name, pars, q, want = case[:4]
    info = core.load_model_info(p_name + "@hardsphere")
    pk = _OracleP(p_name, q)
    pk.q_input = None
    kernel = product.ProductKernel(info, pk, _OracleS(q), [np.asarray(q, "d")])
    kl = type("K", (), {"info": info, "dtype": np.dtype("d")})()
    det, val, mag = make_kernel_args(kl, direct_model.get_mesh(info, pars, dim="1d"))
    got = kernel(det, val, 0.0, mag)
    ref, S, _, _ = oracle_product(p_name, q, pars)
    np.testing.assert_allclose(got, ref, rtol=1e-13)
    assert all(is_near(w, g) for w, g in zip(want, got))
    if len(case) > 4:
        parts = kernel.results()
        for key in ("P(Q)", "S(Q)", "S_eff(Q)"):
            assert all(is_near(w, g) for w, g in zip(case[4][key], parts[key][1])), key
        for key in ("volume", "volume_ratio", "radius_effective"):
            assert is_near(case[4][key], parts[key]), key


def test_product_kernel_split_volfraction_in_p():
    """A form factor with its own volfraction (vesicle): P keeps it, S's
    volfraction slot borrows P's distribution slot (product.py:478-486), and
    the number density is scale/V_shell (product.py:513-517); hollow, so the
    S volume fraction carries the form:shell ratio."""
    q = np.logspace(-3, np.log10(0.5), 40)
    info = core.load_model_info("vesicle@hardsphere")
    names = [p.id for p in info.parameters.call_parameters]
    assert names[2:10] == ["sld", "sld_solvent", "volfraction", "radius", "thickness", "radius_effective",
                           "structure_factor_mode", "radius_effective_mode"]
    kl = type("K", (), {"info": info, "dtype": np.dtype("d")})()
    for beta, er in ((1, 1), (0, 0)):
        pars = dict(radius=60.0, thickness=15.0, radius_pd=0.1, radius_pd_n=9, thickness_pd=0.1,
                    thickness_pd_n=3, volfraction=0.2, structure_factor_mode=beta, radius_effective_mode=er,
                    scale=1.3, background=0.01)
        pk = _OracleP("vesicle", q)
        pk.q_input = None
        kernel = product.ProductKernel(info, pk, _OracleS(q), [q])
        det, val, mag = make_kernel_args(kl, direct_model.get_mesh(info, pars, dim="1d"))
        got = kernel(det, val, 0.0, mag)
        want, _, fq, _ = oracle_product("vesicle", q, pars)
        assert fq[4] > 1.5                      # form:shell ratio of the hollow shape
        np.testing.assert_allclose(got, want, rtol=1e-13)


# ------------------------------------------------------------ sticky hard spheres, square well
@pytest.mark.parametrize("name", ["stickyhardsphere", "squarewell"])
def test_structure_factors_match_reference_source(name):
    ref = oracle.ref_lib("ref_" + name)
    if ref is None:
        pytest.skip("oracle/_ref/ref_%s.so not built (reference tree absent)" % name)
    rng = np.random.default_rng(len(name))
    f = getattr(oracle.lib(), "sasoracle_" + name)
    for _ in range(3000):
        q, r = 10**rng.uniform(-4, 0), 10**rng.uniform(0, 3)
        if name == "stickyhardsphere":
            args = (q, r, rng.uniform(0, 0.6), rng.uniform(0.01, 0.1), rng.uniform(-1, 2))
        else:
            args = (q, r, rng.uniform(0, 0.08), rng.uniform(0, 1.5), rng.uniform(1, 2))
        a, b = f(*args), ref.ref_sq4(*args)
        assert a == b or (np.isnan(a) and np.isnan(b)), args


@pytest.mark.parametrize("name,kats", [("stickyhardsphere", STICKYHARDSPHERE), ("squarewell", SQUAREWELL)])
def test_structure_factor_kats(name, kats):
    for pars, q, want in kats:
        extra = (pars["perturb"], pars["stickiness"]) if name == "stickyhardsphere" else \
            (pars["welldepth"], pars["wellwidth"])
        got = oracle.structure(name, q, pars["radius_effective"], pars["volfraction"], *extra)
        assert all(is_near(w, g) for w, g in zip(want, got)), (want, got)


def test_squarewell_without_well_is_the_hard_sphere():
    """welldepth 0 leaves the hard-core part, whose coefficients differ from
    the Percus-Yevick ones only at order eta^3 ((1 + 2 eta)^2 + eta^3 (eta - 4)
    in place of (1 + 2 eta)^2): the two agree to ~10 eta^3."""
    q = np.logspace(-2.5, 0, 60)
    for phi in (0.001, 0.01, 0.03):
        sw = oracle.structure("squarewell", q, 40.0, phi, 0.0, 1.3)
        hs = oracle.hardsphere(q, 40.0, phi)
        rel = np.max(np.abs(sw/hs - 1))
        assert rel < 10*phi**3, (phi, rel)


def test_sticky_forward_limit():
    """q -> 0: S = ((1 - eta)^2/(1 + 2 eta - mu))^2 with mu = lambda eta (1 - eta)
    and lambda the smaller root of the contact quadratic (Baxter)."""
    for phi, eps, tau in ((0.1, 0.05, 0.2), (0.3, 0.02, 1.0), (0.05, 0.1, 0.5)):
        eta = phi/(1 - eps)**3
        lam = min(np.roots([eta/12, -(tau + eta/(1 - eta)), (1 + eta/2)/(1 - eta)**2]).real)
        mu = lam*eta*(1 - eta)
        want = ((1 - eta)**2/(1 + 2*eta - mu))**2
        got = oracle.structure("stickyhardsphere", [1e-5], 50.0, phi, eps, tau)[0]
        assert got == pytest.approx(want, rel=1e-6)


# ------------------------------------------------------------ Hayter-Penfold rescaled MSA
def _hayter_args(rng):
    """Parameters over the reference's random() ranges (hayter_msa.py:127-145)
    and q over [1e-5, 1]: about 40 % of them fail to converge (NaN), as the
    reference notes."""
    return (10**rng.uniform(-5, 0), 10**rng.uniform(0, 4.7), 10**rng.uniform(-3, -0.3),
            10**rng.uniform(-2, 2.3), rng.uniform(270, 450), 10**rng.uniform(-3, 2), 10**rng.uniform(0, 2.3))


def test_hayter_msa_matches_reference_source():
    """The restatement against the reference's hayter_msa.c compiled as is:
    the same value to the last bit, or NaN on both sides."""
    ref = oracle.ref_lib("ref_hayter_msa")
    if ref is None:
        pytest.skip("oracle/_ref/ref_hayter_msa.so not built (reference tree absent)")
    rng = np.random.default_rng(16)
    f = oracle.lib().sasoracle_hayter_msa
    nan = 0
    for _ in range(4000):
        args = _hayter_args(rng)
        a, b = f(*args), ref.ref_sq6(*args)
        assert a == b or (np.isnan(a) and np.isnan(b)), args
        nan += np.isnan(a)
    assert 0 < nan < 3000          # both branches exercised


def test_hayter_msa_kats():
    for pars, q, want in HAYTER_MSA:
        got = oracle.structure("hayter_msa", q, pars["radius_effective"], pars["volfraction"], pars["charge"],
                               pars["temperature"], pars["concentration_salt"], pars["dielectconst"])
        assert all(is_near(w, g) for w, g in zip(want, got)), (want, got)


def test_hayter_msa_uncharged_limit_is_percus_yevick():
    """Without Coulomb coupling the MSA closure for hard spheres is the
    Percus-Yevick one: as the charge goes to zero S(q) tends to the hard-sphere
    S(q) (an independent restatement), the difference shrinking with it."""
    q = np.logspace(-3, -0.3, 40)
    for phi in (0.01, 0.3):
        hs = oracle.hardsphere(q, 40.0, phi)
        d = [np.max(np.abs(oracle.structure("hayter_msa", q, 40.0, phi, z, 300.0, 0.0, 78.0)/hs - 1))
             for z in (1e-3, 1e-2)]
        assert d[0] < 1e-4 and d[0] < d[1], (phi, d)



def test_structure_eval_on_host_matches_oracle():
    """saship_structure_eval on the host (the header the device kernels use,
    compiled for the CPU): S(q) of every structure factor against the
    oracle's restatement of the reference, and the Hayter MSA's solved
    coefficients reproduce S through structure_sq_prepared."""
    import ctypes as ct
    from sasmodels_amd import _lib
    L = _lib.lib()
    q = np.logspace(-4, 0, 120)
    rng = np.random.default_rng(21)
    cases = [(5, "hardsphere", [50.0, 0.2]), (14, "stickyhardsphere", [30.0, 0.15, 0.05, 0.3]),
             (15, "squarewell", [40.0, 0.05, 1.0, 1.3])]
    for _ in range(8):
        cases.append((16, "hayter_msa", [10**rng.uniform(0.5, 2.5), 10**rng.uniform(-2.5, -0.5),
                                         10**rng.uniform(0, 2), rng.uniform(280, 350), 10**rng.uniform(-3, -1),
                                         rng.uniform(60, 80)]))
    for sid, name, pars in cases:
        p = np.array(pars + [0.0]*(6 - len(pars)))
        prep, sq = np.zeros(8), np.zeros(q.size)
        _lib.check(L.saship_structure_eval(sid, p.ctypes.data, q.size, q.ctypes.data, 0,
                                           prep.ctypes.data, sq.ctypes.data))
        want = oracle.structure(name, q, *pars)
        if name == "hayter_msa" and np.isnan(prep[0]):
            assert np.isnan(sq).all() and np.isnan(want).all()
            continue
        # the header's own hard sphere has no cancellation window: 2e-6 below 2qR = 0.3
        tol = 2e-6 if name == "hardsphere" else 1e-9
        np.testing.assert_allclose(sq, want, rtol=tol, err_msg=name)
