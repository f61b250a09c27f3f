"""
Quadrature control for the orientation average.

The reference regenerates C source to swap its adaptive quadrature for a
fixed Gauss-Legendre rule (sasmodels/generate.py:291-320).  Here the kernels
take the rule size at run time, so the same call only records it on the
model info; the next model handle built from it uses that rule.
"""


def set_integration_size(info, n):
    """Use an *n*-point Gauss-Legendre rule for *info* (0: adaptive)."""
    n = int(n)
    if n < 0:
        raise ValueError("integration size must be >= 0")
    info.integration_size = n
