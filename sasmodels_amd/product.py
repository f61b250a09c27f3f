"""
Form factor times structure factor: the P@S models.

The same surface as the reference's product models (sasmodels/product.py):
:func:`make_product_info` merges the parameter tables, :class:`ProductModel`
binds a form-factor model P and a structure-factor model S, and
:class:`ProductKernel` evaluates

    I(q) = scale * volfraction / V_shell * <F^2> S(q)                 (P*S)
    I(q) = scale * volfraction / V_shell * (<F^2> + <F>^2 (S(q) - 1)) (beta)
           + background

with S at the form's mean effective radius (radius_effective_mode > 0) or
the radius_effective parameter (mode 0), and at the volume fraction
volfraction * V_form / V_shell.  Both P and S run on the GPU through their
own kernels; :class:`ProductBatch` evaluates many P@S parameter sets in one
device sequence with S and the combination fused into the batch's final
kernel (saship_batch_set_structure).

Combined parameter vector (product.py:27-117): scale, background, the P
parameters, radius_effective, volfraction (unless P has its own), the other
S parameters, structure_factor_mode (if P has <F>), radius_effective_mode
(if P has effective-radius modes), then P's magnetic parameters.
"""

from collections import OrderedDict
from copy import copy

import numpy as np

from .details import make_details
from .kernel import Kernel, KernelModel
from .modelinfo import (NUM_COMMON_PARS, NUM_MAGFIELD_PARS, NUM_MAGNETIC_PARS,
                        ModelInfo, ParameterTable, parse_parameter)

STRUCTURE_MODE_ID = "structure_factor_mode"
#: per-set structure parameters of the batched product (SASHIP_SPARS)
SPARS = 9
RADIUS_MODE_ID = "radius_effective_mode"
RADIUS_ID = "radius_effective"
VOLFRAC_ID = "volfraction"


def make_extra_pars(p_info):
    """
    The mode parameters a product adds after the S parameters
    (product.py:146-169): structure_factor_mode (0 "P*S", 1 beta) when P
    returns <F>, radius_effective_mode (default 1) when P has
    effective-radius modes.
    """
    pars = []
    if p_info.have_Fq:
        pars.append(parse_parameter(STRUCTURE_MODE_ID, "", 0, [["P*S", "P*(1+beta*(S-1))"]],
                                    "", "Structure factor calculation"))
    if p_info.radius_effective_modes is not None:
        pars.append(parse_parameter(RADIUS_MODE_ID, "", 1,
                                    [["unconstrained"] + list(p_info.radius_effective_modes)],
                                    "", "Effective radius calculation"))
    return pars


def _tag_parameter(par):
    """Copy of an S parameter renamed ``<id>_S`` (its id collides with P's)."""
    par = copy(par)
    suffix = par.name[len(par.id):]
    par.id = par.id + "_S"
    par.name = par.id + suffix
    return par


def make_product_info(p_info, s_info):
    """
    ModelInfo of P@S (product.py:171-245).  S must start with
    radius_effective and volfraction and have no sld parameters; P must not
    have radius_effective.  volfraction is dropped from the S block when P
    has its own.
    """
    s_kernel_pars = s_info.parameters.kernel_parameters
    if len(s_kernel_pars) < 2:
        raise TypeError("S needs %s and %s as its first parameters" % (RADIUS_ID, VOLFRAC_ID))
    if s_kernel_pars[0].id != RADIUS_ID:
        raise TypeError("S needs %s as first parameter" % RADIUS_ID)
    if s_kernel_pars[1].id != VOLFRAC_ID:
        raise TypeError("S needs %s as second parameter" % VOLFRAC_ID)
    if s_info.parameters.magnetism_index:
        raise TypeError("S should not have SLD parameters")
    if RADIUS_ID in p_info.parameters:
        raise TypeError("P should not have %s" % RADIUS_ID)
    p_has_volfrac = VOLFRAC_ID in p_info.parameters
    p_ids = set(p.id for p in p_info.parameters.kernel_parameters)
    s_list = [(_tag_parameter(p) if p.id in p_ids else p) for p in s_kernel_pars
              if not (p.id == VOLFRAC_ID and p_has_volfrac)]
    combined = list(p_info.parameters.kernel_parameters) + s_list + make_extra_pars(p_info)
    names = [p.id for p in combined]
    if len(set(names)) != len(names):
        raise TypeError("duplicate parameter names in %s@%s" % (p_info.id, s_info.id))
    parameters = ParameterTable(combined)
    # room for every polydisperse parameter of both parts
    parameters.max_pd = p_info.parameters.max_pd + s_info.parameters.max_pd

    info = ModelInfo()
    info.id = "@".join((p_info.id, s_info.id))
    info.name = "@".join((p_info.name, s_info.name))
    info.title = "Product of %s and %s" % (p_info.name, s_info.name)
    info.parameters = info.base = parameters
    info.composition = ("product", [p_info, s_info])
    info.structure_factor = False
    info.have_Fq = False
    info.radius_effective_modes = None
    info.kernel_id = None
    return info


class _Intermediates:
    """
    The parts of the last product evaluation (product.py:271-294), returned
    by calling the object: P(Q), S(Q), volume, volume_ratio,
    radius_effective and, with beta, beta(Q) and S_eff(Q).
    """
    def __init__(self, q, F, Fsq, S, scale, volume, volume_ratio, radius_effective, beta_mode):
        self.args = (q, F, Fsq, S, scale, volume, volume_ratio, radius_effective, beta_mode)

    def __call__(self):
        q, F, Fsq, S, scale, volume, volume_ratio, radius_effective, beta_mode = self.args
        parts = OrderedDict()
        parts["P(Q)"] = (q, scale*Fsq)
        parts["volume"] = volume
        parts["volume_ratio"] = volume_ratio
        parts["radius_effective"] = radius_effective
        parts["S(Q)"] = (q, S)
        if beta_mode:
            beta = F**2/Fsq
            parts["beta(Q)"] = (q, beta)
            parts["S_eff(Q)"] = (q, 1 + beta*(S - 1))
        return parts


class ProductModel(KernelModel):
    """P@S model: *P* and *S* are the component KernelModels (product.py:296-361)."""

    def __init__(self, model_info, P, S):
        self.info = model_info
        self.P = P
        self.S = S
        self.dtype = P.dtype

    def make_kernel(self, q_vectors):
        """Kernels of P and S on the same q, combined."""
        return ProductKernel(self.info, self.P.make_kernel(q_vectors),
                             self.S.make_kernel(q_vectors), q_vectors)

    def release(self):
        self.P.release()
        self.S.release()


class _Layout:
    """
    Where the P, S and mode values sit in a P@S values vector, and how the
    P@S call splits into a P call and an S call (product.py:375-528).
    """

    def __init__(self, info):
        p_info, s_info = info.composition[1]
        self.p_info, self.s_info = p_info, s_info
        self.nvalues = info.parameters.nvalues
        call = info.parameters.call_parameters
        try:
            self.volfrac_index = next(k for k, p in enumerate(call) if p.id == VOLFRAC_ID)
        except StopIteration:
            raise RuntimeError("no %s parameter in %s" % (VOLFRAC_ID, info.id))
        p_npars = p_info.parameters.npars
        s_npars = s_info.parameters.npars
        self.volfrac_in_p = self.volfrac_index < NUM_COMMON_PARS + p_npars
        have_beta = p_info.have_Fq
        have_er = p_info.radius_effective_modes is not None
        # per-parameter slices of CallDetails.length/offset
        self.p_detail = slice(0, p_npars)
        self.s_detail = slice(p_npars, p_npars + s_npars - self.volfrac_in_p)
        # value slots
        first_p, last_p = NUM_COMMON_PARS, NUM_COMMON_PARS + p_npars
        self.p_values = slice(first_p, last_p)
        self.er_index = last_p
        first_s = last_p + NUM_COMMON_PARS - self.volfrac_in_p   # after radius_effective (+ volfraction)
        last_s = first_s + s_npars - NUM_COMMON_PARS
        self.s_values = slice(first_s, last_s)
        self.beta_index = last_s if have_beta else 0
        self.er_mode_index = last_s + have_beta if have_er else 0
        first_mag = last_s + have_beta + have_er
        nmag = NUM_MAGNETIC_PARS*p_info.parameters.nmagnetic
        self.magnetic = slice(first_mag, first_mag + (nmag + NUM_MAGFIELD_PARS if nmag else 0))
        # S distributions start after the S scalar values
        self.s_dist_start = NUM_COMMON_PARS + s_npars

    def modes(self, values):
        """(scale, background, volfraction, radius_effective_mode, beta_mode)."""
        er_mode = int(values[self.er_mode_index]) if self.er_mode_index > 0 else 0
        beta_mode = bool(values[self.beta_index] > 0) if self.beta_index > 0 else False
        return values[0], values[1], values[self.volfrac_index], er_mode, beta_mode

    def p_call(self, call_details, values, scale=1.0, background=0.0):
        """(details, values) of the P call: P scalars, P magnetic values, all distributions."""
        nvalues = self.nvalues
        nweights = int(call_details.num_weights)
        dists = values[nvalues:nvalues + 2*nweights]
        details = make_details(self.p_info, call_details.length[self.p_detail],
                               call_details.offset[self.p_detail], nweights)
        parts = [[scale, background], values[self.p_values], values[self.magnetic], dists]
        n = sum(len(v) for v in parts)
        parts.append(np.zeros((32 - n % 32) % 32))
        return details, np.hstack(parts).astype("d")

    def s_call(self, call_details, values, er_mode, radius_effective, volfrac, volume_ratio):
        """
        (details, values) of the S call: S at radius_effective (from P when
        er_mode > 0, then monodisperse) and at volfrac * volume_ratio, both
        written to the scalar slot and to the first point of the slot's
        distribution with weight 1.
        """
        nvalues = self.nvalues
        nweights = int(call_details.num_weights)
        length = np.array(call_details.length[self.s_detail])
        offset = np.array(call_details.offset[self.s_detail])
        if self.volfrac_in_p:
            # S's volfraction slot borrows P.volfraction's distribution slot
            length = np.insert(length, 1, 1)
            offset = np.insert(offset, 1, call_details.offset[self.volfrac_index - NUM_COMMON_PARS])
        if er_mode > 0:
            length[0] = 1
        details = make_details(self.s_info, length, offset, nweights)
        parts = [[1.0, 0.0, values[self.er_index], 0.0], values[self.s_values],
                 values[nvalues:nvalues + 2*nweights]]
        n = sum(len(v) for v in parts)
        parts.append(np.zeros((32 - n % 32) % 32))
        s_values = np.hstack(parts).astype("d")
        dist = s_values[self.s_dist_start:]
        if er_mode > 0:
            s_values[NUM_COMMON_PARS] = dist[offset[0]] = radius_effective
            dist[offset[0] + nweights] = 1.0
        s_values[NUM_COMMON_PARS + 1] = dist[offset[1]] = volfrac*volume_ratio
        dist[offset[1] + nweights] = 1.0
        return details, s_values


class ProductKernel(Kernel):
    """P@S bound to q: runs P.Fq, then S, and combines them (product.py:362-556)."""

    def __init__(self, model_info, p_kernel, s_kernel, q):
        self.info = model_info
        self.q = q
        self.p_kernel = p_kernel
        self.s_kernel = s_kernel
        self.dtype = p_kernel.dtype
        self.dim = p_kernel.dim
        self.q_input = p_kernel.q_input
        self.results = None
        self._layout = _Layout(model_info)

    def Iq(self, call_details, values, cutoff, magnetic):
        """I(q) of P@S for the combined *call_details* and *values*."""
        lay = self._layout
        scale, background, volfrac, er_mode, beta_mode = lay.modes(values)
        if beta_mode and self.p_kernel.dim == "2d":
            raise NotImplementedError("beta not yet supported for 2D")
        p_details, p_values = lay.p_call(call_details, values)
        F, Fsq, radius_effective, shell_volume, volume_ratio = self.p_kernel.Fq(
            p_details, p_values, cutoff, magnetic, er_mode)
        s_details, s_values = lay.s_call(call_details, values, er_mode, radius_effective,
                                         volfrac, volume_ratio)
        S = self.s_kernel.Iq(s_details, s_values, cutoff, False)
        PS = Fsq + F**2*(S - 1) if beta_mode else Fsq*S
        combined_scale = scale/shell_volume
        if not lay.volfrac_in_p:
            combined_scale *= volfrac
        self.results = _Intermediates(self.q, F, Fsq, S, combined_scale, shell_volume,
                                      volume_ratio, radius_effective, beta_mode)
        return combined_scale*PS + background
    __call__ = Iq

    def Fq(self, call_details, values, cutoff, magnetic, radius_effective_mode=0):
        raise NotImplementedError("a P@S product has no <F>")

    def release(self):
        self.p_kernel.release()
        self.s_kernel.release()


class ProductBatch:
    """
    Many P@S parameter sets evaluated in one device sequence: P's batched
    mesh integration, then S(q) and the product combination in the batch's
    final kernel (saship_batch_set_structure), with the batch's hipGraph,
    autotuning and resolution smearing.  Each set's result equals
    :class:`ProductKernel` on that set.

    *model* is a :class:`ProductModel` whose P is a hip model and whose S
    is the hard-sphere, sticky hard-sphere, square-well or Hayter MSA
    structure factor;
    *q_list* holds one ``[q]`` or
    ``[qx, qy]`` per set.  Sets with radius_effective_mode > 0 in one call
    must share the mode (the batch computes one effective radius per set).
    """

    def __init__(self, model, q_list, use_graph=False):
        from .kernelhip import HipBatch
        if model.S.info.kernel_id not in (5, 14, 15, 16):
            raise ValueError("batched products support the hardsphere, stickyhardsphere, "
                             "squarewell and hayter_msa structure factors")
        self._s_model = model.S.info.kernel_id
        self.info = model.info
        self._layout = _Layout(model.info)
        if any(p.id != q.id for p, q in zip(model.S.info.parameters.kernel_parameters,
                                            self._layout.s_info.parameters.kernel_parameters)):
            raise ValueError("structure factor table mismatch")
        self.batch = HipBatch(model.P, q_list, use_graph=use_graph)
        self.dim = self.batch.dim
        self.nsets = self.batch.nsets
        self._det_cache = {}
        self._idx_cache = {}

    def _p_details(self, cd):
        """P's mesh descriptor for P@S details *cd*, cached by layout (a fit
        keeps the same polydispersity layout from call to call)."""
        lay = self._layout
        key = (cd.length[lay.p_detail].tobytes(), cd.offset[lay.p_detail].tobytes(), int(cd.num_weights))
        det = self._det_cache.get(key)
        if det is None:
            det = make_details(lay.p_info, np.asarray(cd.length[lay.p_detail]),
                               np.asarray(cd.offset[lay.p_detail]), int(cd.num_weights))
            if len(self._det_cache) > 256:
                self._det_cache.clear()
            self._det_cache[key] = det
        return det

    def _p_index(self, nweights):
        """Positions in a P@S values vector of P's values vector (scale and
        background kept: the final kernel applies them), and its padded length."""
        idx = self._idx_cache.get(nweights)
        if idx is None:
            lay = self._layout
            idx = np.concatenate([[0, 1], np.arange(lay.p_values.start, lay.p_values.stop),
                                  np.arange(lay.magnetic.start, lay.magnetic.stop),
                                  np.arange(lay.nvalues, lay.nvalues + 2*nweights)]).astype(np.intp)
            idx = (idx, idx.size + (32 - idx.size % 32) % 32)
            self._idx_cache[nweights] = idx
        return idx

    def _groups(self, calls):
        """Per run of sets sharing a mesh layout (a fit's population does):
        (set indices, P's details, P's values as rows, structure parameters as
        rows, effective-radius modes), from the stacked value vectors."""
        lay = self._layout
        cds, vals = zip(*calls)
        D = np.concatenate([cd.buffer for cd in cds]).reshape(len(cds), -1)
        lens = np.fromiter(map(len, vals), np.int64, len(vals))
        if (D == D[0]).all() and (lens == lens[0]).all():
            groups = [np.arange(len(calls))]
        else:
            key = np.concatenate([D, lens[:, None]], axis=1)
            _, inv = np.unique(key, axis=0, return_inverse=True)
            inv = inv.ravel()
            groups = [np.flatnonzero(inv == g) for g in range(inv.max() + 1)]
        ns = len(range(*lay.s_values.indices(lay.nvalues)))
        for members in groups:
            cd = cds[members[0]]
            V = np.concatenate([vals[s] for s in members]).reshape(len(members), -1)
            er = V[:, lay.er_mode_index].astype(int) if lay.er_mode_index > 0 else np.zeros(len(members), int)
            beta = V[:, lay.beta_index] > 0 if lay.beta_index > 0 else np.zeros(len(members), bool)
            if self.dim == "2d" and beta.any():
                raise NotImplementedError("beta not yet supported for 2D")
            idx, size = self._p_index(int(cd.num_weights))
            P = np.zeros((len(members), size))
            P[:, :idx.size] = V[:, idx]
            sp = np.zeros((len(members), SPARS))
            sp[:, 0] = V[:, lay.er_index]
            sp[:, 1] = V[:, lay.volfrac_index]
            sp[:, 2] = beta
            sp[:, 3] = er
            sp[:, 4] = 1.0 if lay.volfrac_in_p else 0.0
            sp[:, 5:5 + ns] = V[:, lay.s_values]
            yield members, self._p_details(cd), P, sp, er

    @staticmethod
    def _er_mode(modes):
        modes = set(int(e) for e in np.unique(np.asarray(modes, int)) if e > 0)
        if len(modes) > 1:
            raise ValueError("sets of one call need the same radius_effective_mode (got %s)" % sorted(modes))
        return modes.pop() if modes else 0

    def _split(self, calls):
        """(P calls, structure parameters per set, the call's effective-radius mode)."""
        n = len(calls)
        p_calls, spars, modes = [None]*n, np.zeros((n, SPARS)), []
        for rows, det, P, sp, er in self._groups(calls):
            spars[rows] = sp
            for k, s in enumerate(rows):
                p_calls[s] = (det, P[k])
            modes.append(er)
        return p_calls, spars, self._er_mode(np.concatenate(modes))

    def _split_packed(self, calls):
        """As :meth:`_split`, with P's calls packed for HipBatch.call_packed
        (details and values of every set back to back, in set order)."""
        n = len(calls)
        groups = list(self._groups(calls))
        spars, modes = np.zeros((n, SPARS)), []
        sizes = np.zeros(n, np.int64)
        for rows, _, P, sp, er in groups:
            spars[rows] = sp
            sizes[rows] = P.shape[1]
            modes.append(er)
        modes = np.concatenate(modes)
        ndet = groups[0][1].buffer.size
        if len(groups) == 1:
            rows, det, P, _, _ = groups[0]
            vals = P.ravel()
            dets = np.tile(det.buffer.astype(np.int32), n)
        else:
            vals = np.empty(int(sizes.sum()))
            dets = np.empty(n*ndet, np.int32)
            start = np.concatenate([[0], np.cumsum(sizes)[:-1]])
            for rows, det, P, _, _ in groups:
                vals[start[rows][:, None] + np.arange(P.shape[1])] = P
                dets.reshape(n, ndet)[rows] = det.buffer
        voff = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        return (dets, vals, voff, ndet), spars, self._er_mode(modes)

    def __call__(self, calls, cutoff=0.0, magnetic=False):
        """List of P@S I(q), one per set (smeared when set_resolution is on)."""
        packed, spars, er_mode = self._split_packed(calls)
        self.batch.set_structure(spars, self._s_model)
        return self.batch.call_packed(packed, cutoff, magnetic, radius_effective_mode=er_mode)

    def set_resolution(self, resolutions):
        """As :meth:`HipBatch.set_resolution`: smear each set's P@S I(q) on the device."""
        self.batch.set_resolution(resolutions)

    def set_options(self, **kw):
        self.batch.set_options(**kw)

    def autotune(self, calls, cutoff=0.0, magnetic=False):
        p_calls, spars, _ = self._split(calls)
        self.batch.set_structure(spars, self._s_model)
        return self.batch.autotune(p_calls, cutoff, magnetic)

    def replay(self, nsteps):
        return self.batch.replay(nsteps)

    def release(self):
        self.batch.release()
