"""Device-side synthetic media for the large single-GPU workloads.

A seeded von Karman random field (correlation length a, Hurst exponent H,
unit variance, clipped to +-clip) built spectrally on the GPU, applied to a
linear depth gradient of Vs:

    Vs = (vs_top + (vs_bot - vs_top) * z/(nz-1)) * (1 + pert * xi)
    Vp = vpvs * Vs,  rho = 1741 * (Vp/1000)^0.25 (Gardner),
    Qs = 50 * Vs[km/s] (= 0.05 * Vs[m/s]),  Qp = 2 * Qs

written as the reference's media file layout float32 [nz][ny][nx][5]
(vp, vs, rho, qp, qs; AWP-ODC-SGT/src/structure.f:146-170) in device memory,
ready for awp_set_media_device().  torch is used only for device memory, the
random numbers and the FFT.
"""
from __future__ import annotations

import math

import torch


def von_karman_field(nz: int, ny: int, nx: int, dh: float, corr: float, hurst: float,
                     seed: int, device="cuda", clip: float = 3.0) -> torch.Tensor:
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    white = torch.randn((nz, ny, nx), generator=g, device=device, dtype=torch.float32)
    spec = torch.fft.rfftn(white)
    del white
    kz = torch.fft.fftfreq(nz, d=dh, device=device) * 2 * math.pi
    ky = torch.fft.fftfreq(ny, d=dh, device=device) * 2 * math.pi
    kx = torch.fft.rfftfreq(nx, d=dh, device=device) * 2 * math.pi
    # amplitude filter sqrt(P(k)), P(k) ~ (1 + k^2 a^2)^-(H + 3/2) in 3-D
    e = -(hurst + 1.5) / 2.0
    a2 = corr * corr
    for iz in range(nz):  # plane by plane keeps the temporary small
        k2 = kz[iz] ** 2 + ky[:, None] ** 2 + kx[None, :] ** 2
        spec[iz] *= torch.pow(1.0 + k2 * a2, e)
    field = torch.fft.irfftn(spec, s=(nz, ny, nx))
    del spec
    field -= field.mean()
    field /= field.std().clamp_min(1e-30)
    field.clamp_(-clip, clip)
    return field


def c2_media(nx: int, ny: int, nz: int, dh: float, seed: int, vs_top=500.0, vs_bot=4500.0,
             pert=0.05, vpvs=1.73, corr=5000.0, hurst=0.1, device="cuda") -> torch.Tensor:
    """float32 [nz][ny][nx][5] media on the device (see module docstring)."""
    xi = von_karman_field(nz, ny, nx, dh, corr, hurst, seed, device)
    out = torch.empty((nz, ny, nx, 5), device=device, dtype=torch.float32)
    zz = torch.arange(nz, device=device, dtype=torch.float32)
    vs0 = vs_top + (vs_bot - vs_top) * zz / max(nz - 1, 1)
    for iz in range(nz):
        vs = vs0[iz] * (1.0 + pert * xi[iz])
        vp = vpvs * vs
        out[iz, :, :, 0] = vp
        out[iz, :, :, 1] = vs
        out[iz, :, :, 2] = 1741.0 * torch.pow(vp / 1000.0, 0.25)
        out[iz, :, :, 4] = 0.05 * vs
        out[iz, :, :, 3] = 0.1 * vs
    del xi
    return out
