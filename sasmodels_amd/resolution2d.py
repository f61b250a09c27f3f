"""
2D resolution under the reference's module name (sasmodels/resolution2d.py):
Pinhole2D for 2D data and Slit2D for oriented samples measured along qx.
The implementations live in :mod:`sasmodels_amd.resolution`.
"""

from .resolution import N_SLIT_PERP, NPHI_2D as NPHI, NR_2D as NR, SIGMA_ZERO, Pinhole2D, Resolution, Slit2D

#: sigmas covered by the 2D Gaussian (resolution2d.py:18-20)
NSIGMA = 3.0

__all__ = ["Pinhole2D", "Slit2D", "Resolution", "N_SLIT_PERP", "NR", "NPHI", "NSIGMA", "SIGMA_ZERO"]
