"""Seeded synthetic DDM videos (Gaussian spots, Brownian and straight swimmers).

Generator of the benchmark workloads: spots of width sigma and peak amplitude
`amp` on a constant background, periodic wrap, Poisson photon noise, clipped to
the pixel type.  Swimmers move in straight lines with speeds drawn from a Schulz
(Gamma(Z+1, vbar/(Z+1))) distribution and a uniform heading; diffusers take
N(0, 2D) steps per axis per frame.  The structure follows the reference's movie
simulator (test/sim.mjs:90-186) with the workload's units (px, px/frame) and
Poisson instead of Gaussian camera noise.

`video_numpy` runs on the CPU (tests); `video_torch` renders the same model on
the GPU with torch (bench input only; the rendering is not part of the timed
path).
"""
from __future__ import annotations

import numpy as np


def _tracks(rng, T, N, n, alpha, vbar_px_frame, Z, D):
    """Positions (T, n, 2) of n particles: a fraction alpha swims, the rest diffuse."""
    n_sw = int(round(alpha * n))
    pos0 = rng.uniform(0, N, size=(n, 2))
    vel = np.zeros((n, 2))
    if n_sw:
        speed = rng.gamma(Z + 1.0, vbar_px_frame / (Z + 1.0), size=n_sw)
        phi = rng.uniform(0, 2 * np.pi, size=n_sw)
        vel[:n_sw, 0] = speed * np.cos(phi)
        vel[:n_sw, 1] = speed * np.sin(phi)
    steps = np.zeros((T, n, 2))
    if D > 0 and n > n_sw:
        steps[1:, n_sw:, :] = rng.normal(0.0, np.sqrt(2 * D), size=(T - 1, n - n_sw, 2))
    drift = np.arange(T)[:, None, None] * vel[None, :, :]
    return np.mod(pos0[None] + drift + np.cumsum(steps, axis=0), N)


def video_numpy(T, N, n_spots, sigma, amp, background, D=0.0, alpha=0.0, vbar_px_frame=0.0, Z=3.0,
                dtype=np.uint8, seed=1):
    """(T, N, N) frame stack; CPU rendering with separable Gaussians."""
    rng = np.random.default_rng(seed)
    pos = _tracks(rng, T, N, n_spots, alpha, vbar_px_frame, Z, D)
    R = int(np.ceil(4 * sigma))
    off = np.arange(-R, R + 2)
    out = np.empty((T, N, N), dtype=dtype)
    vmax = np.iinfo(dtype).max if np.issubdtype(dtype, np.integer) else None
    for t in range(T):
        px, py = pos[t, :, 0], pos[t, :, 1]
        x0 = np.floor(px).astype(np.int64)[:, None] + off[None, :]
        y0 = np.floor(py).astype(np.int64)[:, None] + off[None, :]
        gx = np.exp(-((x0 - px[:, None]) ** 2) / (2 * sigma * sigma))
        gy = amp * np.exp(-((y0 - py[:, None]) ** 2) / (2 * sigma * sigma))
        idx = (np.mod(y0, N)[:, :, None] * N + np.mod(x0, N)[:, None, :]).ravel()
        w = (gy[:, :, None] * gx[:, None, :]).ravel()
        img = background + np.bincount(idx, weights=w, minlength=N * N)
        img = rng.poisson(img).astype(np.float64)
        if vmax is not None:
            img = np.clip(img, 0, vmax)
        out[t] = img.reshape(N, N).astype(dtype)
    return out


def video_torch(T, N, n_spots, sigma, amp, background, D=0.0, alpha=0.0, vbar_px_frame=0.0, Z=3.0,
                dtype="uint16", seed=2, device="cuda", chunk=64):
    """(T, N, N) frame stack rendered on the GPU (torch); same model as video_numpy."""
    import torch
    rng = np.random.default_rng(seed)
    pos = torch.from_numpy(_tracks(rng, T, N, n_spots, alpha, vbar_px_frame, Z, D)).to(device, torch.float32)
    # uint16 frames are stored in an int16 tensor (same bits; torch's uint16
    # support on the GPU is limited): values above 32767 wrap in two's complement
    tdt = {"uint8": torch.uint8, "uint16": torch.int16, "float32": torch.float32}[dtype]
    vmax = {"uint8": 255.0, "uint16": 65535.0, "float32": None}[dtype]
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    R = int(np.ceil(4 * sigma))
    off = torch.arange(-R, R + 2, device=device, dtype=torch.float32)
    out = torch.empty((T, N, N), dtype=tdt, device=device)
    for t0 in range(0, T, chunk):
        t1 = min(T, t0 + chunk)
        p = pos[t0:t1]
        px, py = p[..., 0], p[..., 1]
        x0 = torch.floor(px)[..., None] + off
        y0 = torch.floor(py)[..., None] + off
        gx = torch.exp(-((x0 - px[..., None]) ** 2) / (2 * sigma * sigma))
        gy = amp * torch.exp(-((y0 - py[..., None]) ** 2) / (2 * sigma * sigma))
        xi = torch.remainder(x0, N).long()
        yi = torch.remainder(y0, N).long()
        idx = (yi[..., :, None] * N + xi[..., None, :]).reshape(t1 - t0, -1)
        w = (gy[..., :, None] * gx[..., None, :]).reshape(t1 - t0, -1)
        img = torch.full((t1 - t0, N * N), float(background), device=device, dtype=torch.float32)
        img.scatter_add_(1, idx, w)
        img = torch.poisson(img, generator=gen)
        if vmax is not None:
            img = img.clamp_(0, vmax)
        img = img.reshape(t1 - t0, N, N)
        if dtype == "uint16":
            img = img.to(torch.int32).to(torch.int16)
        out[t0:t1] = img.to(tdt)
    return out


# Workloads (sizes of the benchmark configurations).
def config1(seed=1):
    """64 frames of 128^2 u8: 200 spots, sigma 1.5, amp 100, bg 20, D = 0.5 px^2/frame."""
    return video_numpy(64, 128, 200, 1.5, 100.0, 20.0, D=0.5, dtype=np.uint8, seed=seed)


C2 = dict(T=1024, N=512, n_spots=4000, sigma=2.0, amp=100.0, background=200.0, D=0.3, alpha=0.6,
          vbar_px_frame=15.0 * 0.010, Z=3.0)
