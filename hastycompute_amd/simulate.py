"""Trajectory simulation on the GPU: the PC-VIPR spoke generator of
python/python/hastypy/util/image_creation.py:13-57 (create_spoke), vectorised over many
spokes so a whole trajectory is generated in HBM.

A spoke is the line k_z = pi * linspace(-1/3, 1, S) with Gaussian jitter (std `noise`)
on all three components, rotated by Rz(zangle) Rx(xangle); missing angles are drawn
uniformly (xangle in [0, pi), zangle in [0, 2 pi scale_factor)). With crop_factor != 1
the spoke is scaled and the samples outside [-pi, pi)^3 are dropped (so the result is
ragged: a list per spoke). Random draws use a torch.Generator, not numpy's global state.
"""
from __future__ import annotations

import math
from typing import List, Optional, Sequence, Union

import torch


def _rot(xangle: torch.Tensor, zangle: torch.Tensor) -> torch.Tensor:
    """Rz(z) @ Rx(x), [nspokes, 3, 3]."""
    cx, sx = torch.cos(xangle), torch.sin(xangle)
    cz, sz = torch.cos(zangle), torch.sin(zangle)
    one, zero = torch.ones_like(cx), torch.zeros_like(cx)
    rx = torch.stack([torch.stack([one, zero, zero], -1), torch.stack([zero, cx, -sx], -1),
                      torch.stack([zero, sx, cx], -1)], -2)
    rz = torch.stack([torch.stack([cz, -sz, zero], -1), torch.stack([sz, cz, zero], -1),
                      torch.stack([zero, zero, one], -1)], -2)
    return rz @ rx


def create_spokes(nspokes: int, samp_per_spoke: int, method: str = "PCVIPR", noise: float = 0.005,
                  scale_factor: float = 1.0, crop_factor: float = 1.0,
                  xangles: Optional[Union[Sequence[float], torch.Tensor]] = None,
                  zangles: Optional[Union[Sequence[float], torch.Tensor]] = None, device="cuda",
                  generator: Optional[torch.Generator] = None) -> Union[torch.Tensor, List[torch.Tensor]]:
    """nspokes spokes of create_spoke: [3, nspokes * samp_per_spoke] float32 on `device`
    (crop_factor == 1), or a list of [3, M_i] tensors (crop_factor != 1)."""
    if method != "PCVIPR":
        raise RuntimeError("Not a valid method")
    S = samp_per_spoke
    t = torch.linspace(-1.0 / 3.0, 1.0, S, dtype=torch.float32, device=device) * math.pi
    spokes = torch.zeros((nspokes, 3, S), dtype=torch.float32, device=device)
    if noise != 0.0:
        spokes.normal_(0.0, noise, generator=generator)
    spokes[:, 2, :] += t
    if xangles is None:
        xangles = math.pi * torch.rand(nspokes, device=device, generator=generator)
    if zangles is None:
        zangles = scale_factor * 2 * math.pi * torch.rand(nspokes, device=device, generator=generator)
    xa = torch.as_tensor(xangles, dtype=torch.float32, device=device).reshape(nspokes)
    za = torch.as_tensor(zangles, dtype=torch.float32, device=device).reshape(nspokes)
    spokes = _rot(xa, za) @ spokes
    if crop_factor == 1.0:
        return spokes.permute(1, 0, 2).reshape(3, nspokes * S).contiguous()
    spokes = spokes * crop_factor
    keep = (spokes.abs() < math.pi).all(dim=1)  # [nspokes, S]
    return [spokes[i][:, keep[i]].contiguous() for i in range(nspokes)]


def create_spoke(samp_per_spoke: int, method: str = "PCVIPR", noise: float = 0.005, scale_factor: float = 1.0,
                 crop_factor: float = 1.0, xangle: Optional[float] = None, zangle: Optional[float] = None,
                 device="cuda", generator: Optional[torch.Generator] = None) -> torch.Tensor:
    """One spoke, image_creation.py:13-57: [3, samp_per_spoke] (fewer when cropped)."""
    out = create_spokes(1, samp_per_spoke, method, noise, scale_factor, crop_factor,
                        None if xangle is None else [xangle], None if zangle is None else [zangle], device, generator)
    return out if isinstance(out, torch.Tensor) else out[0]
