"""Receive coil sensitivities (reference: ``src/blochparty/coils.py:12-43``)."""

from __future__ import annotations

import math

import torch
from torch import Tensor


def birdcage_sensitivities(pos: Tensor, n_coils: int = 8, radius: float = 0.18, normalize: bool = False,
                           radius_y: float | None = None, element_radius: float | None = None) -> Tensor:
    """(N, C) complex sensitivities of ``n_coils`` elements evenly spaced on a circle (an ellipse with ``radius_y``).
    Magnitude radius / distance (distance clamped at 0.05 radius), or a loop's on-axis law with ``element_radius``;
    the phase winds around each element: atan2 of the offset minus the element's angle."""
    ang = 2 * math.pi * torch.arange(n_coils, dtype=pos.dtype, device=pos.device) / n_coils
    ry = radius if radius_y is None else radius_y
    dx = pos[:, :1] - radius * torch.cos(ang)
    dy = pos[:, 1:2] - ry * torch.sin(ang)
    dist = torch.sqrt(dx**2 + dy**2).clamp_min(0.05 * radius)
    if element_radius is None:
        mag = radius / dist
    else:
        a = element_radius
        mag = radius * a**2 / (a**2 + dist**2) ** 1.5 * (a**2 + radius**2) ** 1.5 / a**2
    sens = torch.polar(mag, torch.atan2(dy, dx) - ang)
    if normalize:
        sens = sens / torch.sqrt((sens.abs() ** 2).sum(dim=1, keepdim=True))
    return sens
