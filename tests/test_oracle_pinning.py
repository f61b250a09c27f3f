"""
Pins the FP64 oracle (oracle/sasoracle.c) before it is trusted as the checker.

1. Against the reference's own model and special-function sources compiled
   unmodified into oracle/_ref (skipped where the reference tree was absent
   at build time): special functions bit for bit, form factors to the
   precision of the reference's printed Gauss tables.
2. Against the reference's known answers (model files and
   direct_model.test_simple_interface), fed through the host packing.
3. Against mathematics that does not share the oracle's formulas: scipy
   Bessel functions, numpy's Gauss-Legendre rule, closed forms.
"""

import ctypes as ct
import math

import numpy as np
import pytest
from scipy import special

from kat_tables import DIRECT_MODEL, MODEL_KATS, is_near, q_vectors
from oracle import oracle
from sasmodels_amd import direct_model, generate, modelinfo

P = ct.POINTER(ct.c_double)
NPARS = {"sphere": 3, "cylinder": 6, "ellipsoid": 6, "core_shell_cylinder": 8, "core_shell_sphere": 5,
         "parallelepiped": 8, "triaxial_ellipsoid": 8, "hollow_cylinder": 7, "vesicle": 5,
         "fuzzy_sphere": 4, "core_shell_ellipsoid": 9, "elliptical_cylinder": 8, "core_shell_bicelle": 10}


def _ptr(a):
    return a.ctypes.data


def _need_ref(name):
    lib = oracle.ref_lib(name)
    if lib is None:
        pytest.skip("oracle/_ref/%s.so not built (reference tree absent)" % name)
    return lib


# ---------------------------------------------------------------- 1. reference sources
@pytest.mark.parametrize("fn,refname,lib", [
    ("sasoracle_j1", "ref_j1", "ref_cylinder"),
    ("sasoracle_2j1x_x", "ref_2j1x_x", "ref_cylinder"),
    ("sasoracle_sph_j1c", "ref_sph_j1c", "ref_sphere"),
    ("sasoracle_sinx_x", "ref_sinx_x", "ref_sphere"),
])
def test_special_functions_bit_exact(fn, refname, lib):
    ref = _need_ref(lib)
    rng = np.random.default_rng(7)
    x = np.concatenate([rng.uniform(-12, 12, 4000), 10**rng.uniform(-4, 4, 4000),
                        [0.0, 0.1, -0.1, 5.0, -5.0, 1e-300]])
    f_o, f_r = getattr(oracle.lib(), fn), getattr(ref, refname)
    mismatches = [v for v in x if f_o(v) != f_r(v)]
    assert not mismatches, mismatches[:5]


def _rand_pars(model, rng):
    p = rng.uniform(1, 400, NPARS[model])
    if model in ("sphere", "cylinder", "ellipsoid", "parallelepiped", "triaxial_ellipsoid", "fuzzy_sphere"):
        p[0:2] = rng.uniform(-2, 8, 2)
    elif model == "vesicle":
        p[0:2] = rng.uniform(-2, 8, 2)
        p[2] = rng.uniform(0, 1)
    elif model == "hollow_cylinder":
        p[3:5] = rng.uniform(-2, 8, 2)
    elif model == "core_shell_cylinder":
        p[0:3] = rng.uniform(-2, 8, 3)
    elif model == "core_shell_ellipsoid":
        p[1], p[3] = 10**rng.uniform(-1.5, 1.5, 2)       # axial ratios, oblate to prolate
        p[4:7] = rng.uniform(-2, 8, 3)
        if rng.uniform() < 0.2:
            p[1] = p[3] = 1.0                             # the spherical shortcut of the curvature radius
    elif model == "elliptical_cylinder":
        p[1] = 1.0 + 10**rng.uniform(-3, 1)               # axis ratio >= 1
        p[3:5] = rng.uniform(-2, 8, 2)
    elif model == "core_shell_bicelle":
        p[4:8] = rng.uniform(-2, 8, 4)
    else:
        p[2:5] = rng.uniform(-2, 8, 3)
    return np.ascontiguousarray(p)


@pytest.mark.parametrize("model,refname,quad", [
    ("sphere", "ref_sphere", 0),
    ("cylinder", "ref_cylinder", 0), ("cylinder", "ref_cylinder_g76", 76),
    ("ellipsoid", "ref_ellipsoid", 0), ("ellipsoid", "ref_ellipsoid_g76", 76),
    ("core_shell_cylinder", "ref_core_shell_cylinder", 0),
    ("core_shell_cylinder", "ref_core_shell_cylinder_g76", 76),
    ("core_shell_sphere", "ref_core_shell_sphere", 0),
    ("parallelepiped", "ref_parallelepiped", 0), ("parallelepiped", "ref_parallelepiped_g76", 76),
    ("triaxial_ellipsoid", "ref_triaxial_ellipsoid", 0),
    ("triaxial_ellipsoid", "ref_triaxial_ellipsoid_g76", 76),
    ("hollow_cylinder", "ref_hollow_cylinder", 0), ("hollow_cylinder", "ref_hollow_cylinder_g76", 76),
    ("vesicle", "ref_vesicle", 0), ("fuzzy_sphere", "ref_fuzzy_sphere", 0),
    ("core_shell_ellipsoid", "ref_core_shell_ellipsoid", 0),
    ("core_shell_ellipsoid", "ref_core_shell_ellipsoid_g76", 76),
    ("elliptical_cylinder", "ref_elliptical_cylinder", 0),
    ("elliptical_cylinder", "ref_elliptical_cylinder_g76", 76),
    ("core_shell_bicelle", "ref_core_shell_bicelle", 0),
    ("core_shell_bicelle", "ref_core_shell_bicelle_g76", 76),
])
def test_form_factors_match_reference_sources(model, refname, quad):
    ref = _need_ref(refname)
    L = oracle.lib()
    mid = oracle.MODEL_IDS[model]
    rng = np.random.default_rng(11 + quad)
    worst = 0.0
    for _ in range(400):
        p = _rand_pars(model, rng)
        q = 10**rng.uniform(-3, 0)
        a1, a2, b1, b2 = (ct.c_double() for _ in range(4))
        L.sasoracle_fq(mid, quad, q, _ptr(p), ct.byref(a1), ct.byref(a2))
        ref.ref_fq(q, _ptr(p), ct.byref(b1), ct.byref(b2))
        # F^2 relative; F relative to the RMS amplitude sqrt(F^2)
        if b2.value == 0.0:       # the amplitude underflows (e.g. fuzzy sphere at large q*fuzziness)
            assert a2.value == 0.0 and a1.value == b1.value
        else:
            worst = max(worst, abs(a2.value - b2.value)/abs(b2.value),
                        abs(a1.value - b1.value)/math.sqrt(abs(b2.value))*1e-2)
        assert L.sasoracle_form_volume(mid, _ptr(p)) == ref.ref_form_volume(_ptr(p))
        if hasattr(ref, "ref_shell_volume"):
            assert L.sasoracle_shell_volume(mid, _ptr(p)) == ref.ref_shell_volume(_ptr(p))
        for mode in range(1, 11):
            assert L.sasoracle_radius_effective(mid, mode, _ptr(p)) == ref.ref_reff(mode, _ptr(p))
        if hasattr(ref, "ref_iqac"):
            qab, qc = rng.uniform(0, 0.5, 2)
            u, v = L.sasoracle_iqac(mid, qab, qc, _ptr(p)), ref.ref_iqac(qab, qc, _ptr(p))
            worst = max(worst, abs(u - v)/abs(v) if v else abs(u))
        if hasattr(ref, "ref_iqabc"):
            qa, qb, qc = rng.uniform(-0.5, 0.5, 3)
            u, v = L.sasoracle_iqabc(mid, qa, qb, qc, _ptr(p)), ref.ref_iqabc(qa, qb, qc, _ptr(p))
            assert u == v, (u, v)
    # bit-exact without quadrature; the reference's printed Gauss tables
    # (15-16 digits, 500/5000-point weights to ~4e-10) bound the rest
    limit = 0.0 if model in ("sphere", "core_shell_sphere", "vesicle", "fuzzy_sphere") else 1e-9
    assert worst <= limit, worst


@pytest.mark.parametrize("n", [20, 76])
def test_gauss_tables_match_reference(n):
    ref = _need_ref("ref_cylinder")
    z, w, zr, wr = (np.zeros(n) for _ in range(4))
    oracle.lib().sasoracle_gauss(n, _ptr(z), _ptr(w))
    assert ref.ref_gauss(n, _ptr(zr), _ptr(wr)) == n
    np.testing.assert_allclose(z, zr, atol=1e-15)
    np.testing.assert_allclose(w, wr, rtol=5e-15)


# ---------------------------------------------------------------- 2. reference known answers
def _kat_cases():
    for model, kats in MODEL_KATS.items():
        for k, t in enumerate(kats):
            yield pytest.param(model, t, id="%s-%d" % (model, k))


@pytest.mark.parametrize("model,test", list(_kat_cases()))
def test_reference_kats(model, test):
    info = modelinfo.make_model_info(model)
    qv, xs = q_vectors(test[1])
    kernel = oracle.OracleKernel(info, qv)
    pars = dict(test[0])
    if len(test) == 3:
        actual = direct_model.call_kernel(kernel, pars)
        ys = test[2] if isinstance(test[2], list) else [test[2]]
        for y, a in zip(ys, actual):
            assert is_near(y, a, 5), (y, a)
    else:
        F1, F2, R, V, Vr = direct_model.call_Fq(kernel, pars)
        if test[2] is not None:
            assert is_near(test[2], F1[0], 5)
        if test[3] is not None:
            assert is_near(test[3], F2[0], 5)
        assert is_near(test[4], R, 5) and is_near(test[5], V, 5) and is_near(test[6], Vr, 5)


@pytest.mark.parametrize("case", range(len(DIRECT_MODEL)))
def test_direct_model_kats_full_precision(case):
    """The reference prints these to full double precision; sphere has no
    quadrature, so the oracle reproduces them to the last bit or two."""
    model, pars, x, y = DIRECT_MODEL[case]
    info = modelinfo.make_model_info(model)
    qv, _ = q_vectors(x)
    actual = direct_model.call_kernel(oracle.OracleKernel(info, qv), dict(pars))
    np.testing.assert_allclose(actual, y, rtol=4e-16, atol=0)


def test_magnetic_kat_full_precision():
    """All six spin cross sections (sphere.py:115-121) to the last bit."""
    pars, x, y = MODEL_KATS["sphere"][7]
    info = modelinfo.make_model_info("sphere")
    qv, _ = q_vectors(x)
    actual = direct_model.call_kernel(oracle.OracleKernel(info, qv), dict(pars))
    np.testing.assert_allclose(actual, y, rtol=4e-16, atol=0)


# ---------------------------------------------------------------- 3. mathematics
def test_j1_against_scipy():
    x = np.concatenate([np.linspace(-30, 30, 3001), 10**np.linspace(-6, 3, 500)])
    got = np.array([oracle.lib().sasoracle_j1(v) for v in x])
    np.testing.assert_allclose(got, special.j1(x), rtol=0, atol=2e-15)
    got2 = np.array([oracle.lib().sasoracle_2j1x_x(v) for v in x[x != 0]])
    np.testing.assert_allclose(got2, 2*special.j1(x[x != 0])/x[x != 0], rtol=1e-12, atol=1e-15)


def test_sph_j1c_against_scipy():
    x = np.concatenate([np.linspace(1e-4, 50, 5000), [0.05, 0.0999, 0.1001]])
    got = np.array([oracle.lib().sasoracle_sph_j1c(v) for v in x])
    want = 3*special.spherical_jn(1, x)/x
    np.testing.assert_allclose(got, want, rtol=0, atol=2e-14)


@pytest.mark.parametrize("n", [20, 76, 150, 500])
def test_gauss_against_leggauss(n):
    z, w = np.zeros(n), np.zeros(n)
    assert oracle.lib().sasoracle_gauss(n, _ptr(z), _ptr(w)) == n
    zl, wl = np.polynomial.legendre.leggauss(n)
    np.testing.assert_allclose(z, zl, atol=2e-16)
    np.testing.assert_allclose(w, wl, rtol=0, atol=3e-14)   # leggauss end weights carry ~1e-14 abs error
    # exact on polynomials of degree <= 2n-1
    for deg in (0, 2, 10, 2*n - 2):
        np.testing.assert_allclose(np.sum(w*z**deg), 2.0/(deg + 1), rtol=1e-12)


def test_sphere_closed_form_and_forward_limit():
    """Monodisperse sphere: I(q) = scale*V*drho^2*1e-4 * [3 j1(qR)/(qR)]^2 + bkg."""
    info = modelinfo.make_model_info("sphere")
    q = np.logspace(-5, 0, 50)
    pars = dict(radius=75, sld=3.5, sld_solvent=1.25, scale=0.3, background=0.01)
    I = direct_model.call_kernel(oracle.OracleKernel(info, [q]), pars)
    R = 75.0
    V = 4*np.pi/3*R**3
    x = q*R
    want = 0.3*V*(3.5 - 1.25)**2*1e-4*(3*special.spherical_jn(1, x)/x)**2 + 0.01
    np.testing.assert_allclose(I, want, rtol=1e-12)


def test_cylinder_orientation_average_against_scipy_quad():
    """<F^2> of a cylinder by adaptive integration over the polar angle."""
    from scipy.integrate import quad
    info = modelinfo.make_model_info("cylinder")
    generate.set_integration_size(info, 150)
    q = np.array([0.003, 0.02, 0.07])
    R, L, drho = 25.0, 180.0, 2.5
    F1, F2, _, V, _ = direct_model.call_Fq(oracle.OracleKernel(info, [q]),
                                           dict(radius=R, length=L, sld=3.5, sld_solvent=1.0))
    for qi, f2 in zip(q, F2):
        def amp(a):
            x, y = qi*R*np.sin(a), qi*L/2*np.cos(a)
            jc = 2*special.j1(x)/x if x else 1.0
            return jc*np.sinc(y/np.pi)
        num, _ = quad(lambda a: amp(a)**2*np.sin(a), 0, np.pi/2, epsabs=0, epsrel=1e-13, limit=200)
        want = 1e-4*(drho*V)**2*num
        assert abs(f2/want - 1) < 1e-10


def test_ellipsoid_orientation_average_against_scipy_quad():
    from scipy.integrate import quad
    info = modelinfo.make_model_info("ellipsoid")
    generate.set_integration_size(info, 150)
    q = np.array([0.004, 0.03])
    rp, re = 30.0, 90.0
    _, F2, _, V, _ = direct_model.call_Fq(oracle.OracleKernel(info, [q]),
                                          dict(radius_polar=rp, radius_equatorial=re, sld=2, sld_solvent=1))
    for qi, f2 in zip(q, F2):
        def f(u):
            x = qi*np.sqrt((re*np.sqrt(1 - u*u))**2 + (rp*u)**2)
            return (3*special.spherical_jn(1, x)/x)**2
        num, _ = quad(f, 0, 1, epsabs=0, epsrel=1e-13)
        assert abs(f2/(1e-4*V**2*num) - 1) < 1e-10


@pytest.mark.parametrize("model", ["parallelepiped", "triaxial_ellipsoid"])
def test_nested_orientation_average_against_dblquad(model):
    """The 1D <F^2> of the qabc shapes is the orientation average of the 2D
    amplitude: (2/pi) int_0^pi/2 int_0^pi/2 f^2(q n) sin(theta) dtheta dphi,
    n = (sin t cos p, sin t sin p, cos t), by scipy's adaptive dblquad."""
    from scipy import integrate
    L = oracle.lib()
    mid = oracle.MODEL_IDS[model]
    p = np.ascontiguousarray([3.0, 1.0, 35.0, 75.0, 180.0, 0, 0, 0])
    s = (p[0] - p[1])*L.sasoracle_form_volume(mid, _ptr(p))

    def amp(q, t, f):
        na, nb, nc = math.sin(t)*math.cos(f), math.sin(t)*math.sin(f), math.cos(t)
        if model == "parallelepiped":
            sinc = lambda x: math.sin(x)/x if x else 1.0
            return sinc(0.5*q*p[2]*na)*sinc(0.5*q*p[3]*nb)*sinc(0.5*q*p[4]*nc)
        x = q*math.sqrt((p[2]*na)**2 + (p[3]*nb)**2 + (p[4]*nc)**2)
        return 3*(math.sin(x) - x*math.cos(x))/x**3

    for q in (0.005, 0.03, 0.08):
        f1, f2 = ct.c_double(), ct.c_double()
        L.sasoracle_fq(mid, 150, q, _ptr(p), ct.byref(f1), ct.byref(f2))
        avg2 = integrate.dblquad(lambda t, f: amp(q, t, f)**2*math.sin(t), 0, math.pi/2, 0, math.pi/2,
                                 epsabs=0, epsrel=1e-11)[0]*2/math.pi
        avg1 = integrate.dblquad(lambda t, f: amp(q, t, f)*math.sin(t), 0, math.pi/2, 0, math.pi/2,
                                 epsabs=0, epsrel=1e-11)[0]*2/math.pi
        assert f2.value == pytest.approx(1e-4*s*s*avg2, rel=1e-8), q
        assert f1.value == pytest.approx(1e-2*s*avg1, rel=1e-8), q


def test_qabc_rotation_is_orthonormal_and_reduces_to_qac():
    """The (theta, phi, psi) view and jitter rotation: its first two columns
    are orthonormal, psi = dpsi = 0 gives the qac rotation's third row, and
    the view alone maps the beam-plane direction of the particle's c axis."""
    L = oracle.lib()
    rng = np.random.default_rng(5)
    r = np.zeros(6)
    for _ in range(200):
        th, ph, ps, dth, dph, dps = rng.uniform(-180, 180, 6)
        L.sasoracle_view_jitter_abc(th, ph, ps, dth, dph, dps, _ptr(r))
        c1, c2 = r[0::2], r[1::2]
        assert abs(c1 @ c1 - 1) < 1e-13 and abs(c2 @ c2 - 1) < 1e-13 and abs(c1 @ c2) < 1e-13
        L.sasoracle_view_jitter_abc(th, ph, 0.0, dth, dph, 0.0, _ptr(r))
        r31, r32 = ct.c_double(), ct.c_double()
        L.sasoracle_view_jitter_row(th, ph, dth, dph, ct.byref(r31), ct.byref(r32))
        assert r[4] == pytest.approx(r31.value, abs=1e-15) and r[5] == pytest.approx(r32.value, abs=1e-15)


@pytest.mark.parametrize("model", ["bcc_paracrystal", "fcc_paracrystal", "sc_paracrystal"])
def test_paracrystals_match_reference_sources(model):
    """The reference's paracrystal sources with their printed Gauss-150 table:
    2D bit for bit, the 1D product rule to 1e-9 (the printed table has 15-16
    digits; sharp peaks amplify the difference)."""
    ref = _need_ref("ref_" + model)
    L = oracle.lib()
    mid = oracle.MODEL_IDS[model]
    rng = np.random.default_rng(mid)
    worst = 0.0
    for _ in range(40):
        p = np.ascontiguousarray([rng.uniform(80, 400), rng.uniform(0.01, 0.3), rng.uniform(10, 60),
                                  rng.uniform(-2, 8), rng.uniform(-2, 8), 0, 0, 0])
        q = 10**rng.uniform(-3, -0.3)
        f1, f2 = ct.c_double(), ct.c_double()
        L.sasoracle_fq(mid, 0, q, _ptr(p), ct.byref(f1), ct.byref(f2))
        want = ref.ref_iq(q, _ptr(p))
        worst = max(worst, abs(f2.value - want)/abs(want))
        qa, qb, qc = rng.uniform(-0.3, 0.3, 3)
        assert L.sasoracle_iqabc(mid, qa, qb, qc, _ptr(p)) == ref.ref_iqabc(qa, qb, qc, _ptr(p))
        assert L.sasoracle_form_volume(mid, _ptr(p)) == ref.ref_form_volume(_ptr(p))
    assert worst < 1e-9, worst


@pytest.mark.parametrize("model", ["bcc_paracrystal", "sc_paracrystal"])
def test_paracrystal_average_against_dblquad(model):
    """The 1D lattice factor is the orientation average of the 2D one:
    scipy's adaptive dblquad over the sphere against the product rule at a
    size where it has converged (2000 points; the default Gauss-150 can miss
    the lattice planes' sharp ridges by ~1 %, which the known answers
    include)."""
    from scipy import integrate
    L = oracle.lib()
    mid = oracle.MODEL_IDS[model]
    p = np.ascontiguousarray([150.0, 0.25, 30.0, 4.0, 1.0, 0, 0, 0])
    for q in (0.01, 0.04):
        f1, f2 = ct.c_double(), ct.c_double()
        L.sasoracle_fq(mid, 2000, q, _ptr(p), ct.byref(f1), ct.byref(f2))

        def integrand(t, f):
            n = (math.sin(t)*math.cos(f), math.sin(t)*math.sin(f), math.cos(t))
            return L.sasoracle_iqabc(mid, q*n[0], q*n[1], q*n[2], _ptr(p))*math.sin(t)
        avg = integrate.dblquad(integrand, 0, 2*math.pi, 0, math.pi, epsabs=0, epsrel=1e-10)[0]/(4*math.pi)
        assert f2.value == pytest.approx(avg, rel=1e-9), q


def _fq_o(model, q, p, quad=0):
    a1, a2 = ct.c_double(), ct.c_double()
    p = np.ascontiguousarray(p, "d")
    oracle.lib().sasoracle_fq(oracle.MODEL_IDS[model], quad, float(q), _ptr(p), ct.byref(a1), ct.byref(a2))
    return a1.value, a2.value


def test_new_shapes_reduce_to_known_ones():
    """core_shell_ellipsoid with unit axial ratios is core_shell_sphere; a
    bicelle without rim or faces is a cylinder; an elliptical cylinder of
    axis ratio 1 is a cylinder (its nested rule against the cylinder's own
    rule, at sizes where both have converged)."""
    rng = np.random.default_rng(41)
    for _ in range(100):
        q = 10**rng.uniform(-3, -0.5)
        r, t = rng.uniform(5, 100, 2)
        slds = rng.uniform(-2, 8, 3)
        e1, e2 = _fq_o("core_shell_ellipsoid", q, [r, 1.0, t, 1.0, *slds, 0, 0])
        c1, c2 = _fq_o("core_shell_sphere", q, [r, t, *slds])
        # the rule integrates a constant: the weights sum to 2 up to rounding
        assert math.isclose(e2, c2, rel_tol=1e-11) and math.isclose(abs(e1), abs(c1), rel_tol=1e-11)
        r, length = rng.uniform(5, 100), rng.uniform(10, 400)
        core, face, solvent = rng.uniform(-2, 8, 3)
        b1, b2 = _fq_o("core_shell_bicelle", q, [r, 0.0, 0.0, length, core, face, face, solvent, 0, 0], 76)
        y1, y2 = _fq_o("cylinder", q, [core, solvent, r, length, 0, 0], 76)
        assert math.isclose(b2, y2, rel_tol=1e-12) and math.isclose(b1, y1, rel_tol=1e-12, abs_tol=1e-12*abs(y1))
    for _ in range(40):
        q = 10**rng.uniform(-3, -1.5)
        r, length = rng.uniform(5, 40), rng.uniform(10, 200)      # q r_max < 5: both rules converged
        sld, solvent = rng.uniform(-2, 8, 2)
        l1, l2 = _fq_o("elliptical_cylinder", q, [r, 1.0, length, sld, solvent, 0, 0, 0], 76)
        y1, y2 = _fq_o("cylinder", q, [sld, solvent, r, length, 0, 0], 76)
        assert math.isclose(l2, y2, rel_tol=1e-10) and math.isclose(l1, y1, rel_tol=1e-10)


# ---------------------------------------------------------------- models defined by Iq alone
IQ_ONLY = {
    "guinier": lambda r: [r.uniform(-300, 300)],
    "gaussian_peak": lambda r: [r.uniform(-0.1, 0.3), 10**r.uniform(-4, -1)],
    "lorentz": lambda r: [10**r.uniform(0, 3)],
    "dab": lambda r: [10**r.uniform(0, 3)],
    "gel_fit": lambda r: [r.uniform(-5, 5), r.uniform(-5, 5), r.uniform(2, 500), r.uniform(0, 10), r.uniform(1, 100)],
    "mass_fractal": lambda r: [r.uniform(0, 50), r.choice([1.0, r.uniform(1, 6)]), r.uniform(0, 500)],
    "surface_fractal": lambda r: [r.uniform(0, 50), r.uniform(1, 3), r.uniform(1, 1000)],
    "mass_surface_fractal": lambda r: [r.choice([0.0, r.uniform(0, 6)]), r.uniform(0, 6), r.uniform(1, 5000),
                                       r.uniform(1, 200)],
    "star_polymer": lambda r: [r.uniform(0, 1e4), r.uniform(1, 6)],
    "mono_gauss_coil": lambda r: [r.uniform(0, 100), r.uniform(0, 300)],
    "fractal": lambda r: [r.uniform(0, 1), r.uniform(0.5, 50), r.choice([0.0, 1.0, r.uniform(0, 6)]),
                          r.uniform(1, 500), r.uniform(-2, 8), r.uniform(-2, 8)],
    "fractal_core_shell": lambda r: [r.uniform(1, 80), r.uniform(0, 30), *r.uniform(-2, 8, 3), r.uniform(0, 1),
                                     r.choice([0.0, 1.0, r.uniform(0, 6)]), r.uniform(1, 500)],
    "lamellar_hg": lambda r: [r.uniform(1, 40), r.uniform(0, 20), *r.uniform(-2, 8, 3)],
    "lamellar_stack_caille": lambda r: [r.uniform(1, 60), r.uniform(1, 40), r.uniform(20, 500), r.uniform(0, 0.8),
                                        *r.uniform(-2, 8, 2)],
    "lamellar_hg_stack_caille": lambda r: [r.uniform(1, 30), r.uniform(0, 10), r.uniform(1, 40), r.uniform(20, 200),
                                           r.uniform(0, 0.8), *r.uniform(-2, 8, 3)],
    "lamellar_stack_paracrystal": lambda r: [r.uniform(1, 60), r.uniform(1, 40), r.uniform(20, 500),
                                             r.uniform(0, 50), *r.uniform(-2, 8, 2)],
    "linear_pearls": lambda r: [r.uniform(1, 100), r.uniform(0, 400), r.uniform(1, 12), *r.uniform(-2, 8, 2)],
    "polymer_micelle": lambda r: [r.uniform(0, 20), r.uniform(1e3, 1e5), r.uniform(1e3, 1e5), *r.uniform(0, 7, 3),
                                  r.uniform(10, 100), r.uniform(1, 40), r.uniform(-1, 2), r.uniform(1, 50)],
    "raspberry": lambda r: [*r.uniform(-2, 8, 3), r.uniform(0.01, 0.5), r.uniform(0.001, 0.1), r.uniform(0, 1),
                            r.uniform(500, 5000), r.uniform(10, 200), r.uniform(-1, 1)],
}


@pytest.mark.parametrize("model", list(IQ_ONLY))
def test_iq_only_models_match_reference_sources(model):
    ref = _need_ref("ref_" + model)
    L = oracle.lib()
    mid = oracle.MODEL_IDS[model]
    rng = np.random.default_rng(51 + mid)
    for trial in range(500):
        p = np.ascontiguousarray(IQ_ONLY[model](rng), "d")
        q = 0.0 if trial % 50 == 0 else 10**rng.uniform(-4, 0.5)
        f1, f2 = _fq_o(model, q, p)
        want = ref.ref_iq(q, _ptr(p))
        assert f1 == 0.0 and (f2 == want or (math.isnan(f2) and math.isnan(want))), (trial, q, p, f2, want)
        assert L.sasoracle_form_volume(mid, _ptr(p)) == ref.ref_form_volume(_ptr(p))
        if hasattr(ref, "ref_reff"):
            for mode in (1, 2, 3, 4):
                assert L.sasoracle_radius_effective(mid, mode, _ptr(p)) == ref.ref_reff(mode, _ptr(p))


def test_iq_only_models_against_closed_forms():
    """Formulas that share nothing with the restatement: numpy closed forms,
    the Debye function for the coil and for a one-arm star, scipy's gamma in
    the fractal structure factors."""
    from scipy.special import gamma
    rng = np.random.default_rng(61)
    debye = lambda x: 2*(np.expm1(-x) + x)/x**2
    for _ in range(300):
        q = 10**rng.uniform(-3, 0)
        rg = rng.uniform(-200, 200)
        with np.errstate(over="ignore"):          # negative rg grows without bound, to inf as in C
            want = float(np.exp(-abs(rg)*rg*q*q/3))
        assert math.isclose(_fq_o("guinier", q, [rg])[1], want, rel_tol=1e-14)
        xi = 10**rng.uniform(0, 3)
        assert math.isclose(_fq_o("lorentz", q, [xi])[1], 1/(1 + (q*xi)**2), rel_tol=1e-14)
        assert math.isclose(_fq_o("dab", q, [xi])[1], xi**3/(1 + (q*xi)**2)**2, rel_tol=1e-14)
        r0 = rng.uniform(1, 300)
        x = (q*r0)**2
        if x > 1e-2:      # away from the cancellation of the closed form
            assert math.isclose(_fq_o("mono_gauss_coil", q, [3.0, r0])[1], 3.0*debye(x), rel_tol=1e-11)
            # one arm: the star is a Gaussian coil of the same Rg
            assert math.isclose(_fq_o("star_polymer", q, [r0*r0, 1.0])[1], debye(x), rel_tol=1e-11)
        # fractal: (1 + D Gamma(D-1) sin((D-1) atan(q xi)) / ((q r)^D (1 + 1/(q xi)^2)^((D-1)/2))) P(q)
        vf, r, D, xi = rng.uniform(0, 1), rng.uniform(1, 20), rng.uniform(1.2, 3), rng.uniform(50, 500)
        sb, ss = rng.uniform(-2, 8, 2)
        S = 1 + D*gamma(D - 1)*math.sin((D - 1)*math.atan(q*xi))/((q*r)**D*(1 + 1/(q*xi)**2)**((D - 1)/2))
        x = q*r
        P = (4/3*math.pi*r**3*(sb - ss)*3*(math.sin(x) - x*math.cos(x))/x**3)**2
        assert math.isclose(_fq_o("fractal", q, [vf, r, D, xi, sb, ss])[1], 1e-4*vf*S*P, rel_tol=1e-9)


def test_linear_pearls_printed_values_and_closed_form():
    """linear_pearls: two of the reference's three printed test values
    disagree with its own source; the oracle follows the source (compiled in
    oracle/_ref) at those inputs, and one pearl is the sphere's intensity
    per volume."""
    from kat_tables import LINEAR_PEARLS
    ref = _need_ref("ref_linear_pearls")
    agree = []
    for pars, q, printed in LINEAR_PEARLS:
        p = np.array([pars["radius"], pars["edge_sep"], pars["num_pearls"], pars["sld"], pars["sld_solvent"]])
        got = _fq_o("linear_pearls", q, p)[1]
        assert got == ref.ref_iq(q, _ptr(p))
        agree.append(bool(is_near(printed, got + 1e-3, 5)))
    assert agree == [False, True, False]      # only the second printed value is the source's
    rng = np.random.default_rng(71)
    for _ in range(100):
        q, r = 10**rng.uniform(-3, 0), rng.uniform(1, 100)
        sld, solvent = rng.uniform(-2, 8, 2)
        one = _fq_o("linear_pearls", q, [r, rng.uniform(0, 100), 1.0, sld, solvent])[1]
        sphere = _fq_o("sphere", q, [sld, solvent, r])[1]
        assert math.isclose(one, sphere/(4/3*math.pi*r**3), rel_tol=1e-12)
