"""
1D resolution smearing: the q points a model must be evaluated at (q_calc)
and the weight matrix that maps I(q_calc) to the smeared I at the measured
q, with the reference's interface (sasmodels/resolution.py:27-180).

The weight matrices are built once per data set on the host; applying
them to a batch of curves runs on the GPU after the batched kernel
(:meth:`sasmodels_amd.kernelhip.HipBatch.set_resolution`), using the banded
form returned by :func:`band`.  :meth:`Resolution.apply` keeps the
reference's host semantics for single curves.

Construction follows the reference's definitions:

* pinhole: Gaussian resolution of width q_width at each q, integrated
  over the bins of q_calc by differences of erf at the bin edges (midpoint
  edges), truncated to [q - 2.5 dq, q + 3 dq] and normalised per q
  (resolution.py:182-220);
* slit: the u-substitution sum over bins for the slit length, averaged
  over 2 n_length + 1 offsets across the slit width when both are given
  (resolution.py:223-396);
* q_calc extension: linear steps below/above the data for pinhole,
  geometric steps for slit; q_calc below 0.02 min(q) is dropped
  (resolution.py:80-107, 131-166, 398-544);
* 2D pinhole: every pixel sampled at nr x nphi polar offsets (radii at bin
  centres of [0, nsigma] in units of (dq_parallel, dq_perp), angles
  2 pi k/nphi, rotated to the pixel's direction), averaged with the
  Gaussian mass of each radial bin (resolution2d.py:29-175);
* 2D slit (an oriented sample measured along qx): trapezoid integral over
  qy of a (qy, qx) grid, then the pinhole weights along qx
  (resolution2d.py:177-241).
"""

import numpy as np
from scipy.special import erf

__all__ = ["Resolution", "Perfect1D", "Pinhole1D", "Slit1D", "Pinhole2D", "Slit2D", "apply_resolution_matrix",
           "pinhole_resolution", "slit_resolution", "pinhole_extend_q", "slit_extend_q",
           "bin_edges", "interpolate", "linear_extrapolation", "geometric_extrapolation", "band",
           "smear_bands"]

MINIMUM_RESOLUTION = 1e-8
MINIMUM_ABSOLUTE_Q = 0.02          # relative to the smallest measured q
PINHOLE_N_SIGMA = (2.5, 3.0)       # (below, above) the measured q


class Resolution:
    """Base class: *q* measured points, *q_calc* evaluation points,
    *weight_matrix* [len(q_calc), len(q)] with apply(I(q_calc)) -> I(q)."""
    q = q_calc = weight_matrix = None

    def apply(self, theory):
        raise NotImplementedError("Subclass does not define the apply function")


class Perfect1D(Resolution):
    """No smearing: q_calc is q and apply is the identity."""

    def __init__(self, q):
        self.q = self.q_calc = q

    def apply(self, theory):
        return theory


class Pinhole1D(Resolution):
    """Pinhole geometry: Gaussian resolution *q_width* (1 sigma) at each *q*."""

    def __init__(self, q, q_width, q_calc=None, nsigma=PINHOLE_N_SIGMA):
        self.q, self.q_width = q, q_width
        qc = pinhole_extend_q(q, q_width, nsigma=nsigma) if q_calc is None else np.sort(q_calc)
        qc = qc[np.abs(qc) >= MINIMUM_ABSOLUTE_Q*np.min(q)]
        self.weight_matrix = pinhole_resolution(qc, q, np.maximum(q_width, MINIMUM_RESOLUTION), nsigma=nsigma)
        self.q_calc = np.abs(qc)

    def apply(self, theory):
        return apply_resolution_matrix(self.weight_matrix, theory)


class Slit1D(Resolution):
    """Slit geometry: *q_length* along the slit, *q_width* across it
    (scalars or one value per q)."""

    def __init__(self, q, q_length=None, q_width=None, q_calc=None):
        self.q_length, self.q_width = q_length, q_width
        n = len(q)
        width = np.zeros(n) if q_width is None else np.broadcast_to(np.asarray(q_width, float), (n,)).copy()
        length = np.zeros(n) if q_length is None else np.broadcast_to(np.asarray(q_length, float), (n,)).copy()
        self.q = q.flatten()
        qc = slit_extend_q(q, q_length=length, q_width=width) if q_calc is None else np.sort(q_calc)
        qc = qc[np.abs(qc) >= MINIMUM_ABSOLUTE_Q*np.min(self.q)]
        self.weight_matrix = slit_resolution(qc, self.q, q_length=length, q_width=width)
        self.q_calc = np.abs(qc)

    def apply(self, theory):
        return apply_resolution_matrix(self.weight_matrix, theory)


SIGMA_ZERO = 1.0e-10
NR_2D = {"xhigh": 10, "high": 5, "med": 5, "low": 3}
NPHI_2D = {"xhigh": 20, "high": 12, "med": 6, "low": 4}


class Pinhole2D(Resolution):
    """
    Gaussian smearing of 2D data: *data* holds qx_data, qy_data, q_data and
    optionally dqx_data (radial) and dqy_data (azimuthal); *index* selects
    the pixels.  q_calc = [qx, qy] of nbins x npix points, bin-major.
    """

    def __init__(self, data=None, index=None, nsigma=3.0, accuracy="Low", coords="polar"):
        self.nr = NR_2D[accuracy.lower()]
        self.nphi = NPHI_2D[accuracy.lower()]
        self.nsigma = nsigma
        self.coords = coords
        self.data = data
        self.index = index if index is not None else slice(None)
        self.qx_data = data.qx_data[self.index]
        self.qy_data = data.qy_data[self.index]
        self.q_data = data.q_data[self.index]
        dqx, dqy = getattr(data, "dqx_data", None), getattr(data, "dqy_data", None)
        if dqx is not None and dqy is not None:
            self.dqx_data = np.maximum(dqx[self.index], SIGMA_ZERO)
            self.dqy_data = np.maximum(dqy[self.index], SIGMA_ZERO)
            qx, qy, w = self._sample()
            self.q_calc = [qx, qy]
            self.q_calc_weights = w
        else:
            self.dqx_data = self.dqy_data = None
            self.q_calc = [self.qx_data, self.qy_data]
            self.q_calc_weights = None

    def _sample(self):
        nr, nphi, npix = self.nr, self.nphi, len(self.qx_data)
        step = self.nsigma/nr
        r = step*(np.arange(nr) + 0.5)                       # radial bin centres
        ang = 2.0*np.pi*np.arange(nphi)/nphi
        # bins ordered angle-major, radius-minor; each bin holds all pixels
        rb = np.tile(r, nphi)
        ab = np.repeat(ang, nr)
        wb = np.tile(np.exp(-0.5*(r - 0.5*step)**2) - np.exp(-0.5*(r + 0.5*step)**2), nphi)
        dx = rb[:, None]*self.dqx_data[None, :]
        dy = rb[:, None]*self.dqy_data[None, :]
        qx, qy = self.qx_data[None, :], self.qy_data[None, :]
        if self.coords == "polar":
            phi = np.arctan(qy/qx)
            qr = np.sqrt(qx*qx + qy*qy)
            rad = dx*np.cos(ab)[:, None] + qr
            tan = dy*np.sin(ab)[:, None]
            qx_s = rad*np.cos(phi) - tan*np.sin(phi)
            qy_s = rad*np.sin(phi) + tan*np.cos(phi)
        else:
            qx_s = qx + dx*np.cos(ab)[:, None]
            qy_s = qy + dy*np.sin(ab)[:, None]
        return qx_s.ravel(), qy_s.ravel(), wb

    def apply(self, theory):
        if self.q_calc_weights is None:
            return theory
        t = np.reshape(theory, (len(self.q_calc_weights), len(self.qx_data)))
        return np.average(t, axis=0, weights=self.q_calc_weights)


N_SLIT_PERP = {"xhigh": 1000, "high": 500, "med": 200, "low": 50}


class Slit2D(Resolution):
    """
    Slit smearing of an oriented sample measured along qx
    (resolution2d.py:177-241): the model is evaluated on a (qy, qx) grid,
    integrated over qy in [-q_width, q_width] by the trapezoid rule on
    2 N + 1 log-spaced qy (N = N_SLIT_PERP[accuracy], both signs and 0), then
    smeared along qx with a Gaussian of width q_length (the pinhole weights)
    when any q_length > 0.  q_calc = [qx, qy] of the grid, qy-major (the
    meshgrid layout), so ``apply`` reshapes to (ny, nx).
    """

    def __init__(self, q, q_length, q_width=0.0, q_calc=None, accuracy="low"):
        self.q, self.q_length, self.q_width = q, q_length, q_width
        length = np.full(len(q), float(q_length)) if np.isscalar(q_length) else np.asarray(q_length, float)
        width = float(q_width) if np.isscalar(q_width) else float(np.max(q_width))
        qx = np.sort(q_calc) if q_calc is not None else pinhole_extend_q(q, length, nsigma=3)
        half = np.logspace(np.log10(np.min(q)), np.log10(width), N_SLIT_PERP[accuracy])
        qy = np.concatenate([-half[::-1], [0.0], half])
        grid_x, grid_y = np.meshgrid(qx, qy)
        self.q_calc = [grid_x.ravel(), grid_y.ravel()]
        self.qx_calc, self.qy_calc = qx, qy
        self.nx, self.ny = len(qx), len(qy)
        self.dy = 2*width/self.ny
        # trapezoid weights over qy (the rule numpy's trapz applies)
        dq = np.diff(qy)
        self.qy_weights = np.concatenate([[0.0], dq/2]) + np.concatenate([dq/2, [0.0]])
        if np.any(length > 0):
            self.weights = pinhole_resolution(qx, q, np.maximum(length, MINIMUM_RESOLUTION))
        elif len(qx) == len(q) and np.all(qx == q):
            self.weights = None
        else:
            raise ValueError("Slit2D fails with q_calc != q")

    def apply(self, theory):
        iq = np.trapezoid(np.reshape(theory, (self.ny, self.nx)), x=self.qy_calc, axis=0)
        return iq if self.weights is None else apply_resolution_matrix(self.weights, iq)


def apply_resolution_matrix(weight_matrix, theory):
    """I(q) = theory . W (host)."""
    return np.dot(np.asarray(theory)[None, :], weight_matrix).ravel()


def bin_edges(x):
    """Bin edges halfway between increasing centres, the outer edges half an
    interval beyond the first and last centre."""
    x = np.asarray(x, float)
    if len(x) < 2 or (np.diff(x) < 0).any():
        raise ValueError("Expected bins to be an increasing set")
    edges = np.empty(len(x) + 1)
    edges[1:-1] = 0.5*(x[1:] + x[:-1])
    edges[0] = x[0] - 0.5*(x[1] - x[0])
    edges[-1] = x[-1] + 0.5*(x[-1] - x[-2])
    return edges


def pinhole_resolution(q_calc, q, q_width, nsigma=PINHOLE_N_SIGMA):
    """W[j, i]: weight of I(q_calc[j]) in the smeared I(q[i])."""
    lo_n, hi_n = (nsigma, nsigma) if np.isscalar(nsigma) else nsigma
    q_calc, q, q_width = (np.asarray(v, float) for v in (q_calc, q, q_width))
    cdf = erf((bin_edges(q_calc)[:, None] - q[None, :])/(np.sqrt(2.0)*q_width)[None, :])
    w = np.diff(cdf, axis=0)
    outside = (q_calc[:, None] < (q - lo_n*q_width)[None, :]) | (q_calc[:, None] > (q + hi_n*q_width)[None, :])
    w[outside] = 0.0
    w /= w.sum(axis=0)[None, :]
    return w


def _slit_length_weights(edges, qi, length):
    """Weights of the bins for a slit of length *length* at *qi* (the
    u = sqrt(q^2 - qi^2) substitution, integrated over whole bins)."""
    top = np.sqrt(qi*qi + length*length)
    u2 = edges*edges - qi*qi
    u2[edges < abs(qi)] = 0.0
    u2[edges > top] = top*top - qi*qi
    return np.diff(np.sqrt(u2))/length


def slit_resolution(q_calc, q, *, q_length=0.0, q_width=0.0, n_length=30):
    """W[j, i] for slit smearing (see the module docstring)."""
    q_calc = np.asarray(q_calc, float)
    q = np.asarray(q, float)
    width = np.broadcast_to(np.asarray(q_width, float), q.shape)
    length = np.broadcast_to(np.asarray(q_length, float), q.shape)
    edges = bin_edges(q_calc)
    W = np.zeros((len(q), len(q_calc)))
    for i, (qi, wi, li) in enumerate(zip(q, width, length)):
        if wi == 0.0 and li == 0.0:
            W[i] = q_calc == qi
        elif wi == 0.0:
            W[i] = _slit_length_weights(edges, qi, li)
        elif li == 0.0:
            raise NotImplementedError("We do not yet handle the q_length=0.0 case for all q_width.")
        else:
            for k in range(-n_length, n_length + 1):
                W[i] += _slit_length_weights(edges, qi + k*wi/n_length, li)
            W[i] /= 2*n_length + 1
    return W.T


def linear_extrapolation(q, q_min, q_max):
    """q extended to [q_min, q_max] with about the first/last step of q."""
    q = np.sort(q)
    low = high = np.empty(0)
    if q_min + 2*MINIMUM_RESOLUTION < q[0]:
        step = q[1] - q[0] if len(q) > 1 else 0.0
        n = int(np.ceil((q[0] - q_min)/step)) if step > 0 else 15
        low = np.linspace(q_min, q[0], n + 1)[:-1]
    if q_max - 2*MINIMUM_RESOLUTION > q[-1]:
        step = q[-1] - q[-2] if len(q) > 1 else 0.0
        n = int(np.ceil((q_max - q[-1])/step)) if step > 0 else 15
        high = np.linspace(q[-1], q_max, n + 1)[1:]
    return np.concatenate([low, q, high])


def geometric_extrapolation(q, q_min, q_max, points_per_decade=None):
    """q extended to [q_min, q_max] with its mean geometric step (or
    *points_per_decade*); q_min <= 0 becomes 0.02 q[0]."""
    q = np.sort(q)
    lo, hi = q[0], q[-1]
    if points_per_decade is not None:
        lstep = np.log(10.0)/points_per_decade
    elif hi > lo:
        lstep = (np.log(hi) - np.log(lo))/(len(q) - 1)
    else:
        lstep = np.log(10.0)/10
    low = high = np.empty(0)
    if q_min < lo:
        if q_min < 0:
            q_min = lo*MINIMUM_ABSOLUTE_Q
        n = int(np.ceil((np.log(q[0]) - np.log(q_min))/lstep))
        low = np.logspace(np.log10(q_min), np.log10(q[0]), n + 1)[:-1]
    if q_max > hi:
        n = int(np.ceil((np.log(q_max) - np.log(hi))/lstep))
        high = np.logspace(np.log10(hi), np.log10(q_max), n + 1)[1:]
    return np.concatenate([low, q, high])


def pinhole_extend_q(q, q_width, nsigma=PINHOLE_N_SIGMA):
    """q_calc covering [min(q - lo dq), max(q + hi dq)] (linear steps)."""
    lo_n, hi_n = (nsigma, nsigma) if np.isscalar(nsigma) else nsigma
    return linear_extrapolation(q, np.min(q - lo_n*q_width), np.max(q + hi_n*q_width))


def slit_extend_q(q, *, q_length=0, q_width=0):
    """q_calc covering the slit (geometric steps)."""
    return geometric_extrapolation(q, np.min(q - q_width), np.max(np.sqrt((q + q_width)**2 + q_length**2)))


def interpolate(q, max_step):
    """q with points inserted so that no step exceeds *max_step* (and seven
    steps beyond the end)."""
    step = np.diff(q)
    wide = step > max_step
    if not np.any(wide):
        return q
    extra = []
    for qi, si in zip(q[:-1][wide], step[wide]):
        n = np.ceil(si/max_step)
        extra.extend(qi + np.arange(1, n)*(si/n))
    extra.extend(q[-1] + np.arange(1, 8)*max_step)
    return np.sort(np.hstack((q, extra)))


def band(weight_matrix):
    """
    Banded form of a weight matrix [n_calc, n_out]: for each output i the
    first row `start[i]`, the row count `length[i]` (first to last non-zero
    entry, zeros inside kept) and the weights W[start:start+length, i],
    concatenated.  Returns (start int32, length int32, offset int64,
    weights float64).
    """
    W = np.asarray(weight_matrix, float)
    n_out = W.shape[1]
    start = np.zeros(n_out, np.int32)
    length = np.zeros(n_out, np.int32)
    parts = []
    for i in range(n_out):
        nz = np.flatnonzero(W[:, i])
        if nz.size:
            start[i], length[i] = nz[0], nz[-1] - nz[0] + 1
            parts.append(W[nz[0]:nz[-1] + 1, i])
    offset = np.zeros(n_out + 1, np.int64)
    np.cumsum(length, out=offset[1:])
    weights = np.concatenate(parts) if parts else np.zeros(0)
    return start, length, offset[:-1], weights


def smear_bands(res, n_calc):
    """
    Device form of a resolution for a set with *n_calc* q_calc points:
    (n_out, start, length, stride, offset, weights), and for Slit2D a
    seventh entry (repeat, repeat_step, repeat_offset, repeat_weights): the
    band of each output is repeated at repeat_step along the set's points
    and the repeats summed with repeat_weights.  1D weight matrices give
    contiguous bands; Pinhole2D gives one strided band per pixel over its
    polar bins, all pixels sharing one normalised weight vector; Slit2D the
    qx band of each point repeated over the qy rows with the trapezoid
    weights; None or a resolution without weights is the identity on the
    set's points.
    """
    if isinstance(res, Slit2D):
        if res.nx*res.ny != n_calc:
            raise ValueError("Slit2D has %d q_calc points, the set %d" % (res.nx*res.ny, n_calc))
        if res.weights is None:
            n_out = res.nx
            start, length = np.arange(n_out, dtype=np.int32), np.ones(n_out, np.int32)
            offset, weights = np.zeros(n_out, np.int64), np.ones(1)
        else:
            start, length, offset, weights = band(res.weights)
            n_out = res.weights.shape[1]
        rep = (np.full(n_out, res.ny, np.int32), np.full(n_out, res.nx, np.int64), np.zeros(n_out, np.int64),
               np.asarray(res.qy_weights, float))
        return n_out, start, length, np.ones(n_out, np.int64), offset, weights, rep
    W = getattr(res, "weight_matrix", None)
    wb = getattr(res, "q_calc_weights", None)
    if isinstance(res, Pinhole2D) and wb is not None:
        npix = len(res.qx_data)
        if len(wb)*npix != n_calc:
            raise ValueError("Pinhole2D has %d q_calc points, the set %d" % (len(wb)*npix, n_calc))
        return (npix, np.arange(npix, dtype=np.int32), np.full(npix, len(wb), np.int32),
                np.full(npix, npix, np.int64), np.zeros(npix, np.int64), np.asarray(wb, float)/np.sum(wb))
    if W is None:
        return (n_calc, np.arange(n_calc, dtype=np.int32), np.ones(n_calc, np.int32),
                np.ones(n_calc, np.int64), np.arange(n_calc, dtype=np.int64), np.ones(n_calc))
    if W.shape[0] != n_calc:
        raise ValueError("weight matrix has %d rows for %d q_calc points" % (W.shape[0], n_calc))
    start, length, offset, weights = band(W)
    return W.shape[1], start, length, np.ones(W.shape[1], np.int64), offset, weights
