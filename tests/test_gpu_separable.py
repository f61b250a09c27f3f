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


@pytest.mark.parametrize("nr,nl,pd_type", [(35, 35, "lognormal"), (1, 40, "gaussian"), (64, 9, "schulz"),
                                          (21, 1, "uniform")])
def test_separable_whole_mesh_sums(nr, nl, pd_type):
    """A whole unpruned mesh without an effective radius: the expansion sums
    the weights as products of the two grids' sums instead of walking the
    mesh.  Raw sums and accumulators against the oracle, and the weight sums
    equal (to FP64 rounding) to the walk's, which an effective-radius mode
    forces on the same buffers."""
    print("whole mesh", nr, nl, pd_type, flush=True)
    hk, ok, det, val, mag = buffers("cylinder", [Q], _pars(nr, nl, pd_type=pd_type), 76)
    hk._call_kernel(det, val, 0.0, mag, 0)
    whole = hk.result.copy()
    compare_raw(whole, ok.raw(det, val, 0.0, mag, 0), Q.size, 2)
    hk._call_kernel(det, val, 0.0, mag, 3)
    walk = hk.result.copy()
    np.testing.assert_allclose(whole[2*Q.size:2*Q.size + 3], walk[2*Q.size:2*Q.size + 3], rtol=1e-14, atol=0)
    np.testing.assert_array_equal(whole[:2*Q.size], walk[:2*Q.size])


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



def _custom_mesh(hm, nq_dim, radius, rw, length, lw):
    """Mesh with explicit radius/length points and weights (array dispersion)."""
    from sasmodels_amd import direct_model
    mesh = direct_model.get_mesh(hm.info, dict(sld=2.0, sld_solvent=-1.0), dim=nq_dim)
    names = [p.name for p in hm.info.parameters.call_parameters]
    mesh[names.index("radius")] = (float(radius[0]), np.asarray(radius, "d"), np.asarray(rw, "d"))
    mesh[names.index("length")] = (float(length[0]), np.asarray(length, "d"), np.asarray(lw, "d"))
    return mesh


def _mixed_sign_mesh(hm, seed):
    """Radius and length weights of both signs: the accepted points of a row
    are then not a prefix of the weight order (bilinear mode)."""
    r = np.random.default_rng(seed)
    nr, nl = int(r.integers(3, 20)), int(r.integers(3, 20))
    radius = np.sort(r.uniform(5, 80, nr))
    length = np.sort(r.uniform(40, 600, nl))
    rw = r.uniform(0.2, 1.0, nr)*np.where(np.arange(nr) % 3 == 1, -1, 1)
    lw = r.uniform(0.2, 1.0, nl)*np.where(np.arange(nl) % 4 == 2, -1, 1)
    return _custom_mesh(hm, "1d", radius, rw, length, lw)


def test_separable_bilinear_mode():
    from gpu_helpers import hip_model
    from sasmodels_amd.details import make_kernel_args
    for seed in (1, 2, 3):
        hk, ok, _, _, _ = buffers("cylinder", [Q], {}, 76)
        mesh = _mixed_sign_mesh(hip_model("cylinder", 76), seed)
        det, val, mag = make_kernel_args(hk, mesh)
        print("bilinear mode seed", seed, flush=True)
        ora = ok.raw(det, val, 0.0, mag, 0)
        hk.set_separable(True)
        hk._call_kernel(det, val, 0.0, mag, 0)
        compare_raw(hk.result, ora, Q.size, 2)


def _cylinder_sets(hm, n, seed, with_mixed=True):
    from sasmodels_amd import direct_model
    from sasmodels_amd.details import make_kernel_args
    r = np.random.default_rng(seed)
    kl = type("K", (), {"info": hm.info, "dtype": np.dtype("d")})()
    q_list, calls = [], []
    for s in range(n):
        nq = int(r.integers(1, 700)) if s % 7 else 0 if s == 7 else 64
        q = np.logspace(-3, 0, nq)
        if with_mixed and s == 3:
            mesh = _mixed_sign_mesh(hm, seed + s)
        else:
            pars = dict(radius=r.uniform(5, 150), length=r.uniform(20, 1500),
                        sld=r.uniform(-1, 7), sld_solvent=r.uniform(-1, 7),
                        radius_pd=r.uniform(0, 0.3), radius_pd_n=int(r.integers(1, 65)),
                        radius_pd_type=["gaussian", "lognormal", "schulz"][s % 3],
                        length_pd=r.uniform(0, 0.3), length_pd_n=int(r.integers(1, 65)),
                        length_pd_type=["lognormal", "gaussian", "uniform"][s % 3])
            mesh = direct_model.get_mesh(hm.info, pars, dim="1d")
        q_list.append([q])
        calls.append(make_kernel_args(kl, mesh)[:2])
    return q_list, calls


@pytest.mark.parametrize("cutoff", [0.0, 1e-5, 1e-3])
def test_batched_separable(cutoff):
    """Batched separable path against the oracle and the generic batched path."""
    from gpu_helpers import hip_model
    from oracle.oracle import OracleKernel
    from sasmodels_amd.kernelhip import HipBatch
    hm = hip_model("cylinder", 76)
    q_list, calls = _cylinder_sets(hm, 24, 1234)
    batch = HipBatch(hm, q_list)
    print("batched separable cutoff", cutoff, flush=True)
    sep, _ = batch.raw(calls, cutoff)
    batch.set_separable(False)
    gen, _ = batch.raw(calls, cutoff)
    batch.release()
    off = 0
    for qv, (det, val) in zip(q_list, calls):
        nq = qv[0].size
        n = 2*nq + 4
        ref = OracleKernel(hm.info, qv).raw(det, val, cutoff, False, 0)
        compare_raw(sep[off:off + n], ref, nq, 2)
        compare_raw(gen[off:off + n], ref, nq, 2)
        off += n


def test_batched_separable_graph_replay_and_options():
    """Graph replays and launch shapes change nothing: every replay equals
    the eager call bit for bit, and other node blocks agree to rounding."""
    from gpu_helpers import hip_model
    from sasmodels_amd.kernelhip import HipBatch
    hm = hip_model("cylinder", 76)
    q_list, calls = _cylinder_sets(hm, 16, 99)
    batch = HipBatch(hm, q_list)
    eager, _ = batch.raw(calls, 1e-5)
    batch.set_options(use_graph=True)
    for rep in range(3):
        print("replay", rep, flush=True)
        again, _ = batch.raw(calls, 1e-5)
        assert np.array_equal(again, eager)
        hk, _, det, val, mag = buffers("cylinder", [Q], _pars(5, 7), 76)
        hk._call_kernel(det, val, 0.0, mag, 0)          # eager work between replays
    batch.set_options(use_graph=False)
    for nodes, waves in ((1, 8), (7, 16), (76, 32), (128, 4)):
        batch.set_options(pts_per_item=nodes, waves_per_cu=waves)
        other, _ = batch.raw(calls, 1e-5)
        off = 0
        for qv in q_list:
            nq = qv[0].size
            compare_raw(other[off:off + 2*nq + 4], eager[off:off + 2*nq + 4], nq, 2, rtol=1e-6)
            off += 2*nq + 4
    batch.release()


@pytest.mark.parametrize("separable", [True, False])
def test_batched_few_items(separable):
    """Fewer work items than queue counters (down to one item): every item
    is still handed out once, by the separable and the generic kernel."""
    from gpu_helpers import hip_model
    from oracle.oracle import OracleKernel
    from sasmodels_amd.kernelhip import HipBatch
    hm = hip_model("cylinder", 76)
    q = np.logspace(-3, -0.5, 40)
    hk, _, det, val, mag = buffers("cylinder", [q], _pars(35, 35), 76)
    ref = OracleKernel(hm.info, [q]).raw(det, val, 0.0, False, 0)
    for nsets in (1, 3):
        batch = HipBatch(hm, [[q]]*nsets)
        batch.set_separable(separable)
        for nodes in (128, 16, 1):
            print("few items", separable, nsets, nodes, flush=True)
            batch.set_options(pts_per_item=nodes, waves_per_cu=8)
            raw, _ = batch.raw([(det, val)]*nsets, 0.0)
            n = 2*q.size + 4
            for s in range(nsets):
                compare_raw(raw[s*n:(s + 1)*n], ref, q.size, 2)
        batch.release()


def _adaptive_sets(n, seed):
    """Random cylinder sets for the adaptive rule (q up to 1, q*max(R, L/2)
    through every rule threshold; the last set reaches the 5000-point rule)."""
    rng = np.random.default_rng(seed)
    sets = []
    for s in range(n):
        q = np.logspace(-3, 0, int(rng.integers(60, 250)))
        nr, nl = [(1, 35), (7, 1), (35, 20), (9, 9)][s % 4]
        pd_type = ["lognormal", "gaussian"][s % 2]
        pars = dict(radius=rng.uniform(5, 150), length=rng.uniform(10, 1500), sld=rng.uniform(1, 6),
                    sld_solvent=rng.uniform(-1, 1),
                    radius_pd=rng.uniform(0.02, 0.25) if nr > 1 else 0.0, radius_pd_n=nr, radius_pd_type=pd_type,
                    length_pd=rng.uniform(0.02, 0.25) if nl > 1 else 0.0, length_pd_n=nl, length_pd_type=pd_type)
        sets.append((q, pars))
    sets.append((np.logspace(-2, 0, 64), dict(radius=40.0, length=2400.0, length_pd=0.1, length_pd_n=5)))
    # R and L/2 alike with wide distributions: both grids cross a threshold
    # at the same q (neither grid whole under the rule)
    sets.append((np.logspace(-3, 0, 128), dict(radius=50.0, length=100.0, radius_pd=0.25, radius_pd_n=15,
                                               length_pd=0.25, length_pd_n=15)))
    # a disc (R well above L/2): the radius grid alone crosses the thresholds
    sets.append((np.logspace(-3, 0, 100), dict(radius=300.0, length=20.0, radius_pd=0.2, radius_pd_n=20,
                                               length_pd=0.2, length_pd_n=9)))
    return sets


@pytest.mark.parametrize("graph", [False, True])
def test_batched_separable_adaptive(graph):
    """The adaptive rule (quad 0) on the separable path: every mesh point
    keeps the reference's rule (20/76/500/5000 by q*max(R, L/2), FP64), the
    mesh split into rule rectangles.  Against the FP64 oracle, the generic
    kernel, and the raw sums; replays equal the eager call."""
    from sasmodels_amd.kernelhip import HipBatch
    from gpu_helpers import hip_model, masked_relerr, RTOL
    from oracle.oracle import OracleKernel
    from sasmodels_amd import direct_model
    from sasmodels_amd.details import make_kernel_args
    model = hip_model("cylinder", 0)
    sets = _adaptive_sets(12, 51)
    kl = type("K", (), {"info": model.info, "dtype": np.dtype("d")})()
    calls = [make_kernel_args(kl, direct_model.get_mesh(model.info, p, dim="1d"))[:2] for _, p in sets]
    batch = HipBatch(model, [[q] for q, _ in sets], use_graph=graph)
    batch.set_options(pts_per_item=16)
    for rep in range(2 if graph else 1):
        print("adaptive separable graph", graph, "rep", rep, flush=True)
        raw, iq = batch.raw(calls, 0.0, False, 0)
        off = 0
        for (q, pars), (det, val), i0 in zip(sets, calls, np.cumsum([0] + [q.size for q, _ in sets])[:-1]):
            nq = q.size
            ora = OracleKernel(model.info, [q]).raw(det, val, 0.0, False, 0)
            compare_raw(raw[off:off + 2*nq + 4], ora, nq, 2)
            ref = OracleKernel(model.info, [q])(det, val, 0.0, False)
            assert masked_relerr(iq[i0:i0 + nq], ref) <= RTOL, pars
            off += 2*nq + 4
    batch.set_separable(False)
    generic = np.concatenate(batch(calls))
    batch.set_separable(True)
    sep = np.concatenate(batch(calls))
    sel = generic > 0
    assert np.max(np.abs(sep[sel]/generic[sel] - 1)) <= RTOL
    # the two paths sum in different orders: equal to the last bit would mean
    # the separable path was not taken
    assert not np.array_equal(sep, generic)
    batch.release()


def test_batched_separable_adaptive_fallbacks():
    """Meshes the rule rectangles cannot split (pruned corners at a cutoff)
    take the generic kernel for the whole batch; results still match the oracle."""
    from sasmodels_amd.kernelhip import HipBatch
    from gpu_helpers import hip_model, masked_relerr, RTOL
    from oracle.oracle import OracleKernel
    from sasmodels_amd import direct_model
    from sasmodels_amd.details import make_kernel_args
    model = hip_model("cylinder", 0)
    sets = _adaptive_sets(4, 53)
    kl = type("K", (), {"info": model.info, "dtype": np.dtype("d")})()
    calls = [make_kernel_args(kl, direct_model.get_mesh(model.info, p, dim="1d"))[:2] for _, p in sets]
    batch = HipBatch(model, [[q] for q, _ in sets])
    for cutoff in (1e-3, 1e-12):
        print("adaptive fallback cutoff", cutoff, flush=True)
        out = batch(calls, cutoff)
        for (q, pars), (det, val), o in zip(sets, calls, out):
            ref = OracleKernel(model.info, [q])(det, val, cutoff, False)
            assert masked_relerr(o, ref) <= RTOL, (cutoff, pars)
    batch.release()


def test_single_call_separable_adaptive():
    """Single calls of the cylinder under the adaptive rule take the
    rule-rectangle separable path for a whole mesh (a one-set batch), the
    generic kernel for a pd_start/pd_stop piece; both against the oracle."""
    for q, pars in _adaptive_sets(5, 57):
        print("single adaptive", pars["radius"], pars["length"], flush=True)
        hk, ok, det, val, mag = buffers("cylinder", [q], pars, 0)
        hk._call_kernel(det, val, 0.0, mag, 1)
        sep = hk.result.copy()
        compare_raw(sep, ok.raw(det, val, 0.0, mag, 1), q.size, 2)
        hk.set_separable(False)
        hk._call_kernel(det, val, 0.0, mag, 1)
        gen = hk.result.copy()
        hk.set_separable(True)
        compare_raw(gen, ok.raw(det, val, 0.0, mag, 1), q.size, 2)
        if int(det.num_eval) > 1:
            assert not np.array_equal(sep[:2*q.size], gen[:2*q.size])
            # a piece of the mesh, then the rest accumulated: the generic path
            n = int(det.num_eval)
            hk._call_kernel(det, val, 0.0, mag, 1, 0, n//2)
            hk._call_kernel(det, val, 0.0, mag, 1, n//2, n)
            compare_raw(hk.result, ok.raw(det, val, 0.0, mag, 1), q.size, 2)
        hk.release()
