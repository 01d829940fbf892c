"""Seeded synthetic inputs generated directly on the GPU (trajectories, coil maps,
phantoms) for the benchmark and the serving driver."""
from __future__ import annotations

import math

import torch

PHI1 = 0.465571231876768  # 3-D golden means
PHI2 = 0.682327803828019


def kooshball_torch(nspokes: int, nsamp: int, offset: int, device, dtype=torch.float32) -> torch.Tensor:
    """3-D radial kooshball [3, nspokes*nsamp]: spoke n has u_z = 2 frac(n phi1) - 1,
    theta = 2 pi frac(n phi2); samples -pi + 2 pi m / nsamp along the direction."""
    n = torch.arange(offset, offset + nspokes, dtype=torch.float64, device=device)
    uz = torch.remainder(n * PHI1, 1.0) * 2.0 - 1.0
    th = 2.0 * math.pi * torch.remainder(n * PHI2, 1.0)
    r = torch.sqrt(torch.clamp(1.0 - uz * uz, min=0.0))
    dirs = torch.stack([uz, r * torch.cos(th), r * torch.sin(th)])
    t = -math.pi + 2.0 * math.pi * torch.arange(nsamp, dtype=torch.float64, device=device) / nsamp
    return (dirs[:, :, None] * t[None, None, :]).reshape(3, -1).to(dtype).contiguous()


def coil_maps_torch(ncoils: int, n: int, device, seed: int = 0) -> torch.Tensor:
    """Smooth Gaussian-magnitude x linear-phase maps with sum_c |S_c|^2 = 1, complex64."""
    g = torch.Generator(device="cpu")
    g.manual_seed(seed)
    ax = (torch.arange(n, dtype=torch.float32, device=device) - n / 2) / n
    Z, Y, X = torch.meshgrid(ax, ax, ax, indexing="ij")
    maps = torch.empty((ncoils, n, n, n), dtype=torch.complex64, device=device)
    for c in range(ncoils):
        ctr = (torch.rand(3, generator=g) - 0.5).tolist()
        w = 0.3 + 0.3 * float(torch.rand(1, generator=g))
        kv = ((torch.rand(3, generator=g) - 0.5) * 6).tolist()
        mag = torch.exp(-((Z - ctr[0]) ** 2 + (Y - ctr[1]) ** 2 + (X - ctr[2]) ** 2) / (2 * w * w))
        ph = kv[0] * Z + kv[1] * Y + kv[2] * X
        maps[c] = torch.polar(mag, ph)
    norm = torch.sqrt((maps.abs() ** 2).sum(0, keepdim=True))
    maps /= norm
    return maps


def ellipsoid_phantom_torch(n: int, device, seed: int = 0, nell: int = 6) -> torch.Tensor:
    """Random-ellipsoid complex phantom [n, n, n] (complex64)."""
    g = torch.Generator(device="cpu")
    g.manual_seed(seed)
    ax = (torch.arange(n, dtype=torch.float32, device=device) - n / 2) / (n / 2)
    Z, Y, X = torch.meshgrid(ax, ax, ax, indexing="ij")
    img = torch.zeros((n, n, n), dtype=torch.complex64, device=device)
    for _ in range(nell):
        c = ((torch.rand(3, generator=g) - 0.5) * 0.8).tolist()
        a = (0.1 + 0.4 * torch.rand(3, generator=g)).tolist()
        v = complex(torch.polar(torch.tensor(0.2 + 0.8 * float(torch.rand(1, generator=g))),
                                torch.tensor(2 * math.pi * float(torch.rand(1, generator=g)))))
        mask = ((Z - c[0]) / a[0]) ** 2 + ((Y - c[1]) / a[1]) ** 2 + ((X - c[2]) / a[2]) ** 2 <= 1.0
        img[mask] += v
    return img
