"""CPU restatement of the GPU video generator (test infrastructure only).

Follows hackathon_amd/csrc/ddm_synth.hip draw for draw: Philox4x32-10 with
counter (index, frame, stream, draw) and key = the 64-bit seed, 53-bit
uniforms, Box-Muller normals, Marsaglia-Tsang gamma speeds, Poisson by
sequential products (lambda < 10) or PTRS (Hoermann 1993), spot rendering with
the same truncation and minimum-image wrap.  The movie model is the
reference's simulator (motility-ddm test/sim.mjs:90-186) in pixel units.
Philox itself is pinned by the published known-answer vectors of the Random123
distribution (tests/test_synth_oracle.py).
"""
from __future__ import annotations

import math

import numpy as np

S_INIT, S_STEP, S_NOISE = 0, 1, 2
_M0, _M1, _W0, _W1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85
_MASK = 0xFFFFFFFF


def philox4x32_10(ctr, key):
    """One block: ctr 4 x u32, key 2 x u32 -> 4 x u32 (pure Python ints)."""
    x0, x1, x2, x3 = ctr
    k0, k1 = key
    for _ in range(10):
        p0 = _M0 * x0
        p1 = _M1 * x2
        hi0, lo0 = p0 >> 32, p0 & _MASK
        hi1, lo1 = p1 >> 32, p1 & _MASK
        x0, x1, x2, x3 = (hi1 ^ x1 ^ k0) & _MASK, lo1, (hi0 ^ x3 ^ k1) & _MASK, lo0
        k0 = (k0 + _W0) & _MASK
        k1 = (k1 + _W1) & _MASK
    return [x0, x1, x2, x3]


class Stream:
    """Draws of one (index, frame, stream) counter, in the device's order."""

    def __init__(self, idx, frame, stream, seed):
        self.c = [idx & _MASK, frame & _MASK, stream & _MASK, 0]
        self.key = [seed & _MASK, (seed >> 32) & _MASK]
        self.buf = []

    def next(self):
        if not self.buf:
            self.buf = philox4x32_10(self.c, self.key)
            self.c[3] = (self.c[3] + 1) & _MASK
        return self.buf.pop(0)

    def uniform(self):
        a = self.next() >> 5
        b = self.next() >> 6
        return (a * 67108864.0 + b + 0.5) * (1.0 / 9007199254740992.0)

    def normal2(self):
        u1, u2 = self.uniform(), self.uniform()
        r = math.sqrt(-2.0 * math.log(u1))
        th = 6.283185307179586 * u2
        return r * math.cos(th), r * math.sin(th)


def gamma_mt(g: Stream, shape, scale):
    d = shape - 1.0 / 3.0
    c = 1.0 / math.sqrt(9.0 * d)
    while True:
        x, _ = g.normal2()
        v = 1.0 + c * x
        if v <= 0.0:
            continue
        v = v * v * v
        u = g.uniform()
        if u < 1.0 - 0.0331 * (x * x) * (x * x):
            return d * v * scale
        if math.log(u) < 0.5 * x * x + d * (1.0 - v + math.log(v)):
            return d * v * scale


def poisson(g: Stream, lam):
    if not lam > 0.0:
        return 0.0
    if lam < 10.0:
        L = math.exp(-lam)
        p, k = 1.0, 0
        while True:
            p *= g.uniform()
            if p <= L:
                return float(k)
            k += 1
    slam, loglam = math.sqrt(lam), math.log(lam)
    b = 0.931 + 2.53 * slam
    a = -0.059 + 0.02483 * b
    invalpha = 1.1239 + 1.1328 / (b - 3.4)
    vr = 0.9277 - 3.6224 / (b - 2.0)
    while True:
        U = g.uniform() - 0.5
        V = g.uniform()
        us = 0.5 - abs(U)
        k = math.floor((2.0 * a / us + b) * U + lam + 0.43)
        if us >= 0.07 and V <= vr:
            return float(k)
        if k < 0 or (us < 0.013 and V > us):
            continue
        if math.log(V) + math.log(invalpha) - math.log(a / (us * us) + b) <= -lam + k * loglam - math.lgamma(k + 1.0):
            return float(k)


def n_swimmers(alpha, n):
    return int(math.floor(alpha * n + 0.5))


def particles(N, n, alpha, vbar, Z, seed, heading3d=False):
    """(n, 4): x0, y0, vx, vy (draw order of test/sim.mjs:109-121)."""
    n_sw = n_swimmers(alpha, n)
    out = np.zeros((n, 4))
    for p in range(n):
        g = Stream(p, 0, S_INIT, seed)
        x0, y0 = g.uniform() * N, g.uniform() * N
        vx = vy = 0.0
        if p < n_sw:
            phi = 6.283185307179586 * g.uniform()
            s = gamma_mt(g, Z + 1.0, vbar / (Z + 1.0))
            if heading3d:
                cz = 2.0 * g.uniform() - 1.0
                s *= math.sqrt(1.0 - cz * cz)
            vx, vy = s * math.cos(phi), s * math.sin(phi)
        out[p] = (x0, y0, vx, vy)
    return out


def tracks(N, T, n, alpha, vbar, Z, D, seed, heading3d=False, swim_diffuse=False):
    """(T, n, 2) wrapped positions, FP64."""
    part = particles(N, n, alpha, vbar, Z, seed, heading3d)
    n_sw = n_swimmers(alpha, n)
    sd = math.sqrt(2.0 * D)
    pos = np.zeros((T, n, 2))
    for p in range(n):
        bx = by = 0.0
        diff = (p >= n_sw or swim_diffuse) and D > 0.0
        for t in range(T):
            if diff and t > 0:
                n1, n2 = Stream(p, t, S_STEP, seed).normal2()
                bx += sd * n1
                by += sd * n2
            x = part[p, 0] + t * part[p, 2] + bx
            y = part[p, 1] + t * part[p, 3] + by
            pos[t, p] = (x - math.floor(x / N) * N, y - math.floor(y / N) * N)
    return pos


def _wrap(d, N):
    return np.where(d >= 0.5 * N, d - N, np.where(d < -0.5 * N, d + N, d))


def intensity(N, pos_t, sigma, amp, bg):
    """Expected counts of one frame (N, N) from positions (n, 2), FP64."""
    W = math.ceil(4.0 * sigma) + 1
    x = np.arange(N)[None, :]
    y = np.arange(N)[:, None]
    acc = np.zeros((N, N))
    for px, py in np.asarray(pos_t, dtype=np.float64):
        dx = _wrap(x - px, N)
        dy = _wrap(y - py, N)
        m = (np.abs(dx) <= W) & (np.abs(dy) <= W)
        acc += np.where(m, np.exp(-(dx * dx + dy * dy) / (2.0 * sigma * sigma)), 0.0)
    return bg + amp * acc


def noisy_frame(lam, t, seed, vmax=None):
    """Poisson counts of one frame from its expected counts (N, N)."""
    N = lam.shape[0]
    out = np.zeros_like(lam)
    for i in range(N * N):
        out.flat[i] = poisson(Stream(i, t, S_NOISE, seed), float(lam.flat[i]))
    if vmax is not None:
        out = np.clip(out, 0, vmax)
    return out
