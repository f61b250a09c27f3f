"""
Mixtures of models: ``A+B`` (sum) and ``A*B`` (product of intensities).

The same surface as the reference's mixture models (sasmodels/mixture.py):
:func:`make_mixture_info` merges the parameter tables, :class:`MixtureModel`
holds the part models and :class:`MixtureKernel` evaluates every part on
its own kernel (on the GPU) and combines

    A+B:  I = scale * (A_scale I_A + B_scale I_B + ...) + background
    A*B:  I = scale * I_A * I_B * ... + background

where each part is called with background 0 (and, for a sum, its own
scale; for a product, scale 1).

Combined parameter vector (mixture.py:32-152): scale, background, then per
part (for a sum) its ``<P>_scale`` and its parameters renamed ``<P>_<name>``
with P = A, B, C... in order of first use; parts that are themselves
mixtures keep their prefixes, moved to the next free letter when one is
already taken.  Magnetic parameters of all parts follow, in part order.
"""

from collections import OrderedDict
from copy import copy

import numpy as np

from .details import make_details
from .kernel import Kernel, KernelModel
from .modelinfo import (NUM_COMMON_PARS, NUM_MAGFIELD_PARS, NUM_MAGNETIC_PARS,
                        ModelInfo, Parameter, ParameterTable)


def _renamed(par, prefix):
    """Copy of *par* with its name and id prefixed (or re-prefixed)."""
    par = copy(par)
    par.id = prefix + par.id
    par.name = prefix + par.name
    if par.length_control is not None:
        par.length_control = prefix + par.length_control
    return par


def make_mixture_info(parts, operation="+"):
    """ModelInfo of the mixture of *parts* with *operation* '+' or '*'."""
    if operation not in "+*" or len(operation) != 1:
        raise ValueError("operation must be '+' or '*'")
    used = []                    # prefix letters in order of use
    inner = {}                   # (part index, kernel parameter index) -> renamed parameter
    # parts that are mixtures already carry A_, B_...; give each of their
    # submodels a letter not used by an earlier one
    for k, part in enumerate(parts):
        if not (part.composition and part.composition[0] == "mixture"):
            continue
        pars = part.parameters.kernel_parameters
        start = 0
        for sub in part.composition[1]:
            n = len(sub.parameters.kernel_parameters)
            letter = pars[start].name[0]
            if letter in used:
                while letter in used:
                    letter = chr(ord(letter) + 1)
                for j in range(start, start + n):
                    p = copy(pars[j])
                    p.id = letter + "_" + p.id[2:]
                    p.name = letter + "_" + p.name[2:]
                    if p.length_control is not None:
                        p.length_control = letter + "_" + p.length_control[2:]
                    inner[(k, j)] = p
            used.append(letter)
            start += n

    combined = []
    for k, part in enumerate(parts):
        prefix = ""
        if not part.composition or part.composition[0] == "product":
            letter = chr(ord("A") + len(used))
            used.append(letter)
            prefix = letter + "_"
        pars = [inner.get((k, j), p) for j, p in enumerate(part.parameters.kernel_parameters)]
        if operation == "+":
            scale_prefix = prefix
            if prefix == "" and getattr(part, "operation", "") == "*":
                # a product part scales as a whole: <letters of its parts>_scale
                letters = []
                for p in pars:
                    head = p.id.split("_")[0]
                    if head not in letters:
                        letters.append(head)
                scale_prefix = "".join(letters) + "_"
            combined.append(Parameter(scale_prefix + "scale", "", 1.0, (-np.inf, np.inf), "",
                                      "model intensity for " + part.name))
        combined.extend(_renamed(p, prefix) for p in pars)
    names = [p.id for p in combined]
    if len(set(names)) != len(names):
        raise TypeError("duplicate parameter names in mixture: %s" % names)
    parameters = ParameterTable(combined)
    parameters.max_pd = sum(part.parameters.max_pd for part in parts)

    info = ModelInfo()
    info.id = operation.join(part.id for part in parts)
    info.operation = operation
    info.name = "(" + operation.join(part.name for part in parts) + ")"
    info.title = "Mixture model with " + info.name
    info.parameters = info.base = parameters
    info.composition = ("mixture", list(parts))
    info.structure_factor = False
    info.have_Fq = False
    info.radius_effective_modes = None
    info.kernel_id = None
    return info


class MixtureModel(KernelModel):
    """Mixture of part models (mixture.py:154-182)."""

    def __init__(self, model_info, parts):
        self.info = model_info
        self.parts = parts
        self.dtype = parts[0].dtype

    def make_kernel(self, q_vectors):
        """Every part bound to the same q."""
        return MixtureKernel(self.info, [part.make_kernel(q_vectors) for part in self.parts], q_vectors)

    def release(self):
        for part in self.parts:
            part.release()


def _part_calls(info, call_details, values):
    """
    (part info, details, values) for each part of the mixture *info*
    (mixture.py:250-339): the part's slice of the length/offset vectors, its
    scalars after [part scale or 1, 0], the field and its own magnetic
    values when it has sld parameters, and the whole distribution vector
    (offsets stay valid).
    """
    op = info.operation
    parts = info.composition[1]
    npars = info.parameters.npars
    nvalues = info.parameters.nvalues
    nweights = int(call_details.num_weights)
    dists = values[nvalues:nvalues + 2*nweights]
    spin = NUM_COMMON_PARS + npars                 # up_frac_i, ... of the mixture
    mag = spin + NUM_MAGFIELD_PARS                 # first magnetic triple
    pos = NUM_COMMON_PARS                          # first value of the current part
    out = []
    for part in parts:
        n = part.parameters.npars
        if op == "+":
            scale, first = values[pos], pos + 1
        else:
            scale, first = 1.0, pos
        # index into CallDetails.length/offset (which start after scale, background)
        k0 = first - NUM_COMMON_PARS
        details = make_details(part, np.asarray(call_details.length[k0:k0 + n]),
                               np.asarray(call_details.offset[k0:k0 + n]), nweights)
        nmag = len(part.parameters.magnetism_index)
        pieces = [[scale, 0.0], values[first:first + n]]
        if nmag:
            pieces += [values[spin:spin + NUM_MAGFIELD_PARS], values[mag:mag + NUM_MAGNETIC_PARS*nmag]]
        pieces.append(dists)
        size = sum(len(v) for v in pieces)
        pieces.append(np.zeros((32 - size % 32) % 32))
        out.append((part, details, np.hstack(pieces).astype("d")))
        pos = first + n
        mag += NUM_MAGNETIC_PARS*nmag
    return out


class MixtureKernel(Kernel):
    """Mixture bound to q (mixture.py:198-247)."""

    def __init__(self, model_info, kernels, q):
        self.dim = kernels[0].dim
        self.info = model_info
        self.q = q
        self.kernels = kernels
        self.dtype = kernels[0].dtype
        self.q_input = getattr(kernels[0], "q_input", None)
        self.operation = model_info.operation
        self.results = None

    def Iq(self, call_details, values, cutoff, magnetic):
        """I(q) of the mixture for the combined *call_details* and *values*."""
        scale, background = values[0], values[1]
        total = 0.0
        results = []
        for kernel, (part, details, part_values) in zip(
                self.kernels, _part_calls(self.info, call_details, values)):
            result = np.array(kernel(details, part_values, cutoff, magnetic), "d")
            if self.operation == "+":
                total = total + result
            elif np.all(total) == 0.0:
                # the reference starts the product over whenever the running
                # product holds an exact zero (mixture.py:226-230)
                total = result
            else:
                total = total*result
            results.append((kernel, result, getattr(kernel, "results", None)))
        q = self.q

        def intermediates():
            parts = OrderedDict()
            for k, (kernel, Iq, sub) in enumerate(results):
                key = chr(ord("A") + k) + "_" + kernel.info.name + "(Q)"
                parts[key] = (q, Iq)
                if sub is not None:
                    parts[key + " parts"] = sub()
            return parts
        self.results = intermediates
        return scale*total + background
    __call__ = Iq

    def Fq(self, call_details, values, cutoff, magnetic, radius_effective_mode=0):
        raise NotImplementedError("a mixture has no <F>")

    def release(self):
        for k in self.kernels:
            k.release()
