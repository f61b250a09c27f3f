"""
Model loading: :func:`load_model_info`, :func:`build_model`, :func:`load_model`.

Mirrors the entry points of sasmodels/core.py:117-341.  The only platform
here is ``"hip"`` (libsaship.so on an AMD GPU); asking for another platform
raises, and so does a missing library or GPU -- there is no CPU fallback.
Model strings are a model name, a product ``P@S`` of a form factor and a
structure factor, or a mixture of those with ``+`` (sum) and ``*``
(product of intensities), parsed as the reference does (core.py:130-167):
split on every ``+`` first, then on ``*``, then on ``@``.
"""

from . import mixture, modelinfo, product
from .kernelhip import HipModel

PLATFORMS = ("hip",)


def list_models():
    """Names of the models the hip platform provides."""
    return list(modelinfo.MODELS)


def load_model_info(model_string):
    """ModelInfo for a model name, a ``P@S`` product or a ``+``/``*`` mixture."""
    for op in "+*":
        if op in model_string:
            parts = [load_model_info(part) for part in model_string.split(op)]
            return mixture.make_mixture_info(parts, operation=op)
    if "@" in model_string:
        parts = model_string.split("@")
        if len(parts) != 2:
            raise ValueError("a product is one form factor @ one structure factor: %r" % model_string)
        p_info, s_info = [modelinfo.make_model_info(part.strip()) for part in parts]
        if p_info.structure_factor or not s_info.structure_factor:
            raise ValueError("%s: expected form_factor@structure_factor" % model_string)
        return product.make_product_info(p_info, s_info)
    return modelinfo.make_model_info(model_string)


def build_model(model_info, dtype=None, platform="hip"):
    """
    Prepare *model_info* for evaluation.  *dtype* is accepted for interface
    compatibility: kernels compute in FP32 with FP64 normalisation sums.
    """
    if platform not in PLATFORMS:
        raise ValueError("platform %r not available; use one of %s"
                         % (platform, ", ".join(PLATFORMS)))
    if dtype is not None and str(dtype).rstrip("!") not in (
            "single", "double", "fast", "float32", "float64", "default"):
        raise ValueError("unsupported dtype %r" % (dtype,))
    if model_info.composition is not None:
        kind, parts = model_info.composition
        models = [build_model(part, dtype=dtype, platform=platform) for part in parts]
        if kind == "product":
            return product.ProductModel(model_info, *models)
        if kind == "mixture":
            return mixture.MixtureModel(model_info, models)
        raise ValueError("unknown composition %r" % kind)
    return HipModel(model_info, dtype=dtype)


def load_model(model_name, dtype=None, platform="hip"):
    """load_model_info + build_model."""
    return build_model(load_model_info(model_name), dtype=dtype, platform=platform)
