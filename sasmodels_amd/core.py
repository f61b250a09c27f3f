"""
Model loading: :func:`load_model_info`, :func:`build_model`, :func:`load_model`.

Mirrors the entry points of sasmodels/core.py:117-341.  The only platform
here is ``"hip"`` (libsaship.so on an AMD GPU); asking for another platform
raises, and so does a missing library or GPU -- there is no CPU fallback.
"""

from . import modelinfo
from .kernelhip import HipModel

PLATFORMS = ("hip",)


def list_models():
    """Names of the models the hip platform provides."""
    return list(modelinfo.MODELS)


def load_model_info(model_string):
    """ModelInfo for a single model name (no mixtures or products)."""
    if any(op in model_string for op in "+*@"):
        raise ValueError("model composition (%s) is not provided by the hip backend"
                         % model_string)
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
    return HipModel(model_info, dtype=dtype)


def load_model(model_name, dtype=None, platform="hip"):
    """load_model_info + build_model."""
    return build_model(load_model_info(model_name), dtype=dtype, platform=platform)
