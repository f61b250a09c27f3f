"""
Model loading: :func:`load_model_info`, :func:`build_model`, :func:`load_model`.

Mirrors the entry points of sasmodels/core.py:117-341.  The only platform
here is ``"hip"`` (libsaship.so on an AMD GPU); asking for another platform
raises, and so does a missing library or GPU -- there is no CPU fallback.
Model strings are a model name or a product ``P@S`` of a form factor and a
structure factor (core.py:151-153); mixtures (``+``, ``*``) are not provided.
"""

from . import modelinfo, product
from .kernelhip import HipModel

PLATFORMS = ("hip",)


def list_models():
    """Names of the models the hip platform provides."""
    return list(modelinfo.MODELS)


def load_model_info(model_string):
    """ModelInfo for a model name or a ``P@S`` product."""
    if any(op in model_string for op in "+*"):
        raise ValueError("model mixtures (%s) are not provided by the hip backend"
                         % model_string)
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
    if model_info.composition is not None and model_info.composition[0] == "product":
        P, S = [build_model(part, dtype=dtype, platform=platform)
                for part in model_info.composition[1]]
        return product.ProductModel(model_info, P, S)
    return HipModel(model_info, dtype=dtype)


def load_model(model_name, dtype=None, platform="hip"):
    """load_model_info + build_model."""
    return build_model(load_model_info(model_name), dtype=dtype, platform=platform)
