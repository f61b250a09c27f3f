"""
The counting oracle (oracle/flopcount.cpp: sasoracle.c on a counting
scalar), which gives the algorithmic FLOP figures of the benchmark: it must
compute exactly what the oracle computes, and its tallies must equal hand
counts of the formulas on the branches taken.
"""

import numpy as np
import pytest

from oracle import oracle
from sasmodels_amd import direct_model, generate, modelinfo
from sasmodels_amd.details import make_kernel_args


def _args(name, pars, dim="1d", quad=0):
    info = modelinfo.make_model_info(name)
    generate.set_integration_size(info, quad)
    kl = type("K", (), {"info": info, "dtype": np.dtype("d")})()
    det, val, mag = make_kernel_args(kl, direct_model.get_mesh(info, dict(pars), dim=dim))
    return info, det, val, mag


@pytest.mark.parametrize("name,dim,quad,pars", [
    ("cylinder", "1d", 76, dict(radius_pd=0.1, radius_pd_n=5, length_pd=0.1, length_pd_n=3)),
    ("cylinder", "1d", 0, dict(radius=30, length=2000, radius_pd=0.1, radius_pd_n=3)),
    ("ellipsoid", "2d", 0, dict(theta_pd=10, theta_pd_n=3, radius_equatorial_pd=0.1, radius_equatorial_pd_n=3)),
    ("core_shell_sphere", "mag", 0, dict(sld_core_M0=1.0, up_frac_i=0.3, up_theta=30, radius_pd=0.1,
                                         radius_pd_n=3)),
])
def test_counting_oracle_computes_what_the_oracle_computes(name, dim, quad, pars):
    info, det, val, mag = _args(name, pars, "2d" if dim == "mag" else dim, quad)
    if dim == "1d":
        q, kind = np.logspace(-3, 0, 57), oracle.KIND_IQ
    else:
        x = np.linspace(-0.3, 0.3, 9)
        q = np.column_stack([a.ravel() for a in np.meshgrid(x, x)])
        kind = oracle.KIND_IMAGNETIC if dim == "mag" else oracle.KIND_IQXY
    want = oracle.raw_call(name, kind, quad, q, det.buffer, val)
    got = oracle.count_flops(name, kind, quad, q, det, val, threads=3)
    np.testing.assert_array_equal(got["result"], want)
    assert got["flops"] > 0


def _sphere_counts(q, radius):
    """Hand count of one sphere 1D call with a single mesh point (sphere.c:1-20,
    sas_3j1x_x.c:46-57, kernel_iq.c mesh pass and accumulation)."""
    # mesh pass: weight product 1 mul; volume 4pi/3 r^3: 3 mul; three weight sums:
    # sum w (1 add), sum w V_form and sum w V_shell (1 mul + 1 add each)
    add, mul, div, trans = 3, 6, 0, 0
    x = q*radius
    # per q: x = q r (1 mul), contrast (1 add), form = contrast V bes (V: 3 mul, then 2 mul),
    # F1 (1 mul), F2 (2 mul), two accumulations w*F (1 mul + 1 add each)
    mul += 1 + 5 + 1 + 2 + 2
    add += 1 + 2
    if abs(x) < 0.1:
        mul += 1 + 3        # x^2, Horner in x^2
        add += 3
    else:
        trans += 2          # sin, cos
        div += 2            # s/x, / x^2
        add += 1            # s/x - c
        mul += 2            # 3 *, x*x
    return dict(add=add, mul=mul, div=div, sqrt=0, trans=trans)


@pytest.mark.parametrize("q", [1e-4, 1.0e-3, 0.05, 0.3])
def test_tallies_match_hand_count(q):
    info, det, val, _ = _args("sphere", dict(radius=50))
    got = oracle.count_flops("sphere", oracle.KIND_IQ, 0, np.array([q]), det, val)
    want = _sphere_counts(q, 50.0)
    assert {k: got[k] for k in oracle.FLOP_CATEGORIES} == want
    assert got["flops"] == sum(want.values())


def test_tallies_add_over_q_and_count_j1_branches():
    """Counts are additive over q, and the J1 branch beyond |x| = 5 costs
    the published rational approximation plus sin, cos and sqrt."""
    info, det, val, _ = _args("cylinder", dict(radius=20, length=400), quad=76)
    qs = np.array([1e-3, 0.02, 0.3, 0.9])
    whole = oracle.count_flops("cylinder", oracle.KIND_IQ, 76, qs, det, val)
    parts = [oracle.count_flops("cylinder", oracle.KIND_IQ, 76, qs[i:i + 1], det, val) for i in range(4)]
    mesh_pass = oracle.count_flops("cylinder", oracle.KIND_IQ, 76, qs[:0], det, val)
    for k in oracle.FLOP_CATEGORIES:
        assert whole[k] - mesh_pass[k] == sum(p[k] - mesh_pass[k] for p in parts)
    # at q = 1e-3 every J1 argument is <= 5: no sqrt; at q = 0.9 (qR up to 18) some are above
    assert parts[0]["sqrt"] == mesh_pass["sqrt"]
    assert parts[3]["sqrt"] > mesh_pass["sqrt"]
    # per node: sin and cos of the angle, the sine of sinc, and sin, cos in J1's
    # large-argument branch only
    per_node = (parts[3]["trans"] - mesh_pass["trans"])/76
    assert 3 < per_node < 5
    assert (parts[0]["trans"] - mesh_pass["trans"])/76 == 3
