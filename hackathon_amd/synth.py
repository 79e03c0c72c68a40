"""Seeded synthetic DDM videos (Gaussian spots, Brownian and straight swimmers).

Generator of the benchmark workloads: spots of width sigma and peak amplitude
`amp` on a constant background, periodic wrap, Poisson photon noise, clipped to
the pixel type.  Swimmers move in straight lines with speeds drawn from a Schulz
(Gamma(Z+1, vbar/(Z+1))) distribution and a uniform heading; diffusers take
N(0, 2D) steps per axis per frame.  The structure follows the reference's movie
simulator (test/sim.mjs:90-186) with the workload's units (px, px/frame) and
Poisson instead of Gaussian camera noise.

`GpuSynth` / `video_gpu` run the HIP generator of libddm.so (ddm_synth_video:
Philox4x32-10 streams, Marsaglia-Tsang speeds, PTRS Poisson noise) into device
memory; `video_numpy` is a small CPU renderer of the same movie model for
host-side tests, and `video_torch` a torch rendering kept for comparison.
"""
from __future__ import annotations

import ctypes as C

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


class VideoParams(C.Structure):
    _fields_ = [("N", C.c_int32), ("T", C.c_int32), ("n_spots", C.c_int32), ("dtype", C.c_int32),
                ("noise", C.c_int32), ("heading3d", C.c_int32),
                ("swim_diffuse", C.c_int32), ("pad_", C.c_int32), ("sigma", C.c_double), ("amp", C.c_double),
                ("background", C.c_double), ("alpha", C.c_double), ("vbar", C.c_double), ("Z", C.c_double),
                ("D", C.c_double), ("seed", C.c_uint64)]


_DT = {"uint8": 0, "uint16": 1, "float32": 2}


class GpuSynth:
    """The HIP video generator (ddm_synth_*); outputs go to torch device tensors."""

    def __init__(self):
        from .engine import DdmError, load_library
        self._err = DdmError
        L = load_library()
        if not getattr(L, "_synth_declared", False):
            vp = C.c_void_p
            L.ddm_synth_create.argtypes = [C.POINTER(vp)]
            L.ddm_synth_create.restype = C.c_int
            L.ddm_synth_destroy.argtypes = [vp]
            L.ddm_synth_destroy.restype = None
            L.ddm_synth_last_error.argtypes = [vp]
            L.ddm_synth_last_error.restype = C.c_char_p
            L.ddm_synth_video.argtypes = [vp, C.POINTER(VideoParams), vp]
            L.ddm_synth_video.restype = C.c_int
            L.ddm_synth_tracks.argtypes = [vp, C.POINTER(VideoParams), vp]
            L.ddm_synth_tracks.restype = C.c_int
            L._synth_declared = True
        self._lib = L
        h = C.c_void_p()
        if L.ddm_synth_create(C.byref(h)) != 0:
            raise DdmError("ddm_synth_create failed")
        self._h = h

    def close(self):
        if self._h:
            self._lib.ddm_synth_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def params(T, N, n_spots, sigma, amp, background, D=0.0, alpha=0.0, vbar_px_frame=0.0, Z=3.0,
               dtype="uint16", seed=1, noise=True, heading3d=False, swim_diffuse=False) -> VideoParams:
        return VideoParams(int(N), int(T), int(n_spots), _DT[dtype], 1 if noise else 0, 1 if heading3d else 0,
                           1 if swim_diffuse else 0, 0, float(sigma), float(amp),
                           float(background), float(alpha), float(vbar_px_frame), float(Z), float(D), int(seed))

    def video(self, p: VideoParams, out=None):
        """(T, N, N) device tensor (uint8 / int16 holding uint16 bits / float32)."""
        import torch
        tdt = {0: torch.uint8, 1: torch.int16, 2: torch.float32}[p.dtype]
        if out is None:
            out = torch.empty((p.T, p.N, p.N), dtype=tdt, device="cuda")
        elif out.dtype != tdt or out.numel() < p.T * p.N * p.N or not out.is_cuda:
            raise self._err("output tensor does not fit the video")
        rc = self._lib.ddm_synth_video(self._h, C.byref(p), C.c_void_p(out.data_ptr()))
        if rc != 0:
            raise self._err(self._lib.ddm_synth_last_error(self._h).decode())
        return out

    def tracks(self, p: VideoParams):
        """(T, n_spots, 2) device float32 positions."""
        import torch
        out = torch.empty((p.T, max(p.n_spots, 1), 2), dtype=torch.float32, device="cuda")
        rc = self._lib.ddm_synth_tracks(self._h, C.byref(p), C.c_void_p(out.data_ptr()))
        if rc != 0:
            raise self._err(self._lib.ddm_synth_last_error(self._h).decode())
        return out[:, :p.n_spots]


def video_gpu(T, N, n_spots, sigma, amp, background, D=0.0, alpha=0.0, vbar_px_frame=0.0, Z=3.0, dtype="uint16",
              seed=2, synth: GpuSynth | None = None):
    """HIP-rendered (T, N, N) frame stack in device memory."""
    g = synth or GpuSynth()
    return g.video(GpuSynth.params(T, N, n_spots, sigma, amp, background, D, alpha, vbar_px_frame, Z, dtype, seed))
