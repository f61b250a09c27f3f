"""
The separable cylinder evaluation (radius x length product mesh, fixed
Gauss-Legendre orientation quadrature) against the FP64 oracle and against
the generic mesh kernel on the same buffers.

The separable path evaluates J1c(q R_i sin a) and sinc(q L_j cos a / 2) once
per (q, node) and forms the weighted bilinear sums over the mesh; the
generic path evaluates every (q, node, mesh point).  Both must meet the
5e-5 tolerance on F^2 and on F (relative to the RMS amplitude).
"""

import numpy as np
import pytest

from gpu_helpers import buffers, compare_raw

pytestmark = pytest.mark.gpu

Q = np.logspace(-3, 0, 131)


def _pars(nr, nl, radius=30.0, length=300.0, pd_type="lognormal"):
    return dict(radius=radius, length=length, sld=3.5, sld_solvent=-0.5,
                radius_pd=0.12 if nr > 1 else 0.0, radius_pd_n=nr, radius_pd_type=pd_type,
                length_pd=0.15 if nl > 1 else 0.0, length_pd_n=nl, length_pd_type=pd_type)


@pytest.mark.parametrize("nr,nl", [(1, 1), (35, 35), (1, 40), (64, 1), (7, 41), (48, 5), (13, 64), (9, 65)])
def test_separable_matches_oracle_and_generic(nr, nl):
    print("separable", nr, nl, flush=True)
    hk, ok, det, val, mag = buffers("cylinder", [Q], _pars(nr, nl), 76)
    ora = ok.raw(det, val, 0.0, mag, 1)
    hk.set_separable(True)
    hk._call_kernel(det, val, 0.0, mag, 1)
    sep = hk.result.copy()
    compare_raw(sep, ora, Q.size, 2)
    hk.set_separable(False)
    hk._call_kernel(det, val, 0.0, mag, 1)
    gen = hk.result.copy()
    compare_raw(gen, ora, Q.size, 2)
    hk.set_separable(True)
    if nr*nl > 1 and nl <= 64:
        # different summation orders: the two paths cannot agree to the last bit
        assert not np.array_equal(sep[:2*Q.size], gen[:2*Q.size])


@pytest.mark.parametrize("quad", [20, 500])
def test_separable_other_rules(quad):
    print("separable quad", quad, flush=True)
    hk, ok, det, val, mag = buffers("cylinder", [Q], _pars(11, 17, radius=80, length=900), quad)
    hk._call_kernel(det, val, 0.0, mag, 3)
    compare_raw(hk.result, ok.raw(det, val, 0.0, mag, 3), Q.size, 2)


def test_separable_cutoff_chunks_and_truncation():
    """Cutoff pruning, pd_start/pd_stop pieces and a gaussian radius mesh
    clipped at zero (fewer radius points than requested)."""
    pars = _pars(21, 15, radius=10.0, length=120.0, pd_type="gaussian")
    pars.update(radius_pd=0.6, radius_pd_nsigma=3)
    hk, ok, det, val, mag = buffers("cylinder", [Q], pars, 76)
    for cutoff in (0.0, 1e-4, 5e-3):
        print("cutoff", cutoff, flush=True)
        hk._call_kernel(det, val, cutoff, mag, 2)
        compare_raw(hk.result, ok.raw(det, val, cutoff, mag, 2), Q.size, 2)
    hk._call_kernel(det, val, 0.0, mag, 2)
    whole = hk.result.copy()
    n = int(det.num_eval)
    for start in range(0, n, 37):
        hk._call_kernel(det, val, 0.0, mag, 2, pd_start=start, pd_stop=min(n, start + 37))
    compare_raw(hk.result, whole, Q.size, 2, rtol=1e-6)

