"""
sasmodels_amd -- an MI355X-native backend for the sasmodels I(q) path.

Provides the reference's evaluation surface (load_model ->
KernelModel.make_kernel(q_vectors) -> Kernel(call_details, values, cutoff,
magnetic)) for the sphere, cylinder, ellipsoid, core_shell_cylinder and
core_shell_sphere models, computed by hand-written HIP kernels in
libsaship.so.
"""

__version__ = "0.1.0"
