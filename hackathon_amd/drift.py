"""Uniform drift velocity from the radial record's phase sums.

`estimate_drift` is the reference's estimateDrift (motility-ddm
src/motility_fit.js:658-725): a weighted least-squares fit of the Fourier phase
increments phi = arg[C(q, tau+1) conj C(q, tau)] = -q.V dt over every binned
mode and lag pair, from the 8 sums per (bin, lag) that `DdmEngine.readRadial`
returns in `Radial.drift` (computed on the GPU, src/ddm_ref.js:360-381).  When a
swimmer component is given, only the lag pairs where its self-ISF has decayed
(|g(q vbar tau)| < 0.2) are used, so the remaining coherent phase rotation is
advection.  Host FP64 arithmetic: the inputs are NB x L x 8 sums.

`swimmer_isf` is the kernel of the gating: the 3D Schulz-averaged swimmer ISF
(isf3d, src/motility_fit.js:135-144) or the in-plane one (isf2d, :146-159,
evaluated with the reference's Legendre quadrature legendreP, :120-132; the
reference gates 2D fits with its tabulated isf2dFast, whose values agree with
isf2d to the tolerance of its own tests).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .fit import Prepared

DRIFT_ISF_MAX = 0.2   # GATES.DRIFT_ISF_MAX (src/motility_fit.js:34)
FLOW_SIGMA = 3.0      # GATES.FLOW_SIGMA (:30)
FLOW_MIN_UM_S = 2.0   # GATES.FLOW_MIN_UM_S (:31)


def isf3d(x: float, Z: float) -> float:
    """sin(Z atan u) / (Z u (1 + u^2)^(Z/2)), u = x/(Z+1); series below u = 1e-4."""
    u = x / (Z + 1)
    if u < 1e-4:
        return 1 - x * x * (Z + 2) / (6 * (Z + 1))
    return math.sin(Z * math.atan(u)) / (Z * u * math.exp(0.5 * Z * math.log1p(u * u)))


def _legendre_p(nu: float, c: float) -> float:
    """P_nu(c) for |c| <= 1 by the trapezoid rule on Laplace's integral."""
    s = math.sqrt(max(0.0, 1 - c * c))
    n = 16 + math.ceil(nu * math.atan2(s, c))
    acc = 0.0
    for i in range(n + 1):
        phi = math.pi * i / n
        im = s * math.cos(phi)
        val = math.exp(0.5 * nu * math.log(c * c + im * im)) * math.cos(nu * math.atan2(im, c))
        acc += 0.5 * val if i in (0, n) else val
    return acc / n


def isf2d(x: float, Z: float) -> float:
    """P_Z(1/sqrt(1+u^2)) / (1+u^2)^((Z+1)/2), u = x/(Z+1); series below u = 1e-4."""
    u = x / (Z + 1)
    if u < 1e-4:
        return 1 - x * x * (Z + 2) / (4 * (Z + 1))
    w = 1 + u * u
    ln_pre = -0.5 * (Z + 1) * math.log(w)
    if ln_pre < -25:
        return 0.0
    return _legendre_p(Z, 1 / math.sqrt(w)) * math.exp(ln_pre)


def swimmer_isf(geometry: str):
    return isf2d if geometry == "2d" else isf3d


@dataclass
class DriftEstimate:
    Vx: float
    Vy: float
    mag: float
    sigma: float
    angleDeg: float
    significant: bool
    pairs: int
    samples: float


def estimate_drift(radial, pr: Prepared, dxUm: float, geometry: str = "3d", swimmer=None) -> DriftEstimate:
    """estimateDrift (src/motility_fit.js:671-725).

    radial: the Radial record with drift sums (readRadial(drift=True));
    pr: `fit.prepare(radial, dxUm, fps)`; swimmer: None or (vbar, Z) of the
    fitted swimmer component, which turns on the lag-pair gating.
    """
    if radial.drift is None or len(radial.drift) == 0:
        raise ValueError("estimate_drift: the Radial record carries no drift sums (readRadial(drift=True))")
    L, N = radial.L, radial.N
    drift = np.asarray(radial.drift, dtype=np.float64).reshape(radial.NB, L, 8)
    kernel = swimmer_isf(geometry)
    S = np.zeros(8)
    modes_used = 0
    pairs = 0
    for b in pr.bins:
        used = False
        for tau in range(1, L - 1):
            if swimmer is not None:
                vbar, Z = swimmer
                g = kernel(b.q * vbar * tau * pr.dt, Z)
                if not (abs(g) < DRIFT_ISF_MAX):
                    continue
            o = drift[b.b, tau]
            if not (o[0] > 0):
                continue
            S += o
            used = True
            pairs += 1
        if used:
            modes_used += int(radial.nModes[b.b])
    Sw, Sxx, Sxy, Syy, Rx, Ry, Rpp, n = (float(v) for v in S)
    det = Sxx * Syy - Sxy * Sxy
    if n < 10 or not (det > 1e-12 * Sxx * Syy) or not (Sw > 0):
        return DriftEstimate(0.0, 0.0, 0.0, math.inf, 0.0, False, pairs, n)
    # beta solves [Sxx Sxy; Sxy Syy] beta = [Rx; Ry], phi ~ kx bx + ky by
    bx = (Syy * Rx - Sxy * Ry) / det
    by = (Sxx * Ry - Sxy * Rx) / det
    rss = max(Rpp - bx * Rx - by * Ry, 0.0)
    # lag pairs of one mode are correlated: inflate by the mean pairs per mode
    inflate = max(1.0, n / max(modes_used, 1))
    s2 = rss / (n - 2) * inflate
    cxx, cyy, cxy = s2 * Syy / det, s2 * Sxx / det, -s2 * Sxy / det
    conv = -(N * dxUm) / (2 * math.pi * pr.dt)          # beta -> V [um/s]
    Vx, Vy = conv * bx, conv * by
    mag = math.hypot(Vx, Vy)
    if mag > 0:
        ux, uy = Vx / mag, Vy / mag
        var_along = conv * conv * (ux * ux * cxx + 2 * ux * uy * cxy + uy * uy * cyy)
    else:
        var_along = conv * conv * (cxx + cyy) / 2
    sigma = math.sqrt(max(var_along, 0.0))
    return DriftEstimate(Vx, Vy, mag, sigma, math.atan2(Vy, Vx) * 180 / math.pi, mag > FLOW_SIGMA * sigma,
                         pairs, n)
