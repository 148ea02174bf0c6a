"""Phantoms: isochromats with tissue properties, and the map-based builders the workloads use.

Reference: ``src/blochparty/phantom.py`` — ``Phantom`` (:17-51), ``grid_positions`` (:54-66), ``phantom_from_maps``
(:90-205), ``TISSUES`` and the Shepp-Logan geometry (:209-247), ``shepp_logan_maps`` / ``shepp_logan`` (:250-279).
``three_class_maps`` has no reference counterpart: it is the seeded head-like phantom of the 3D workloads (three
tissue classes from a thresholded smooth Gaussian random field, after ``anatomy/phantom.py:109-127``).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, replace

import torch
from torch import Tensor

GAMMA_BAR = 42.577478518e6  # 1H gyromagnetic ratio / 2 pi [Hz/T]


@dataclass
class Phantom:
    pos: Tensor  # (N, 3) [m]
    pd: Tensor  # (N,) equilibrium magnetization
    t1: Tensor  # (N,) [s]
    t2: Tensor  # (N,) [s]
    db0: Tensor  # (N,) off-resonance [Hz]
    b1: Tensor | None = None  # (N,) transmit scale of every RF pulse; None = 1
    bound: object | None = None  # magnetization transfer pool (refused by this engine)
    pools: tuple | None = None  # free exchanging pools (refused)
    diffusion: Tensor | None = None  # (refused)
    restriction: object | None = None  # (refused)
    motion: object | None = None  # (refused)

    def __len__(self) -> int:
        return int(self.pd.shape[0])

    def to(self, dtype=None, device=None) -> Phantom:
        conv = lambda t: None if t is None else t.to(dtype=dtype, device=device)
        return replace(self, pos=conv(self.pos), pd=conv(self.pd), t1=conv(self.t1), t2=conv(self.t2),
                       db0=conv(self.db0), b1=conv(self.b1), diffusion=conv(self.diffusion))


def grid_positions(shape: int | tuple[int, ...], fov: float, dtype=torch.float64) -> Tensor:
    """(*shape, 3) voxel centres; ``shape`` n (n x n at z = 0), (ny, nx) or (nz, ny, nx); index i of an axis of n
    sits at (i - n//2) fov / n, the Cartesian k-space convention."""
    shape = (shape, shape) if isinstance(shape, int) else tuple(shape)
    axes = [(torch.arange(n, dtype=dtype) - n // 2) * (fov / n) for n in shape]
    xyz = list(reversed(torch.meshgrid(*axes, indexing="ij")))
    if len(xyz) == 2:
        xyz.append(torch.zeros_like(xyz[0]))
    return torch.stack(xyz, dim=-1)


FAT_PEAKS = ((-3.80, 0.087), (-3.40, 0.693), (-2.60, 0.128), (-1.94, 0.004), (-0.39, 0.039), (0.60, 0.048))

TISSUES = {
    "fat": (0.9, 0.38, 0.10),
    "wm": (0.69, 0.85, 0.07),
    "gm": (0.8, 1.4, 0.10),
    "csf": (1.0, 4.0, 2.0),
    "lesion_t2": (0.9, 1.6, 0.25),
    "lesion_t1": (0.8, 0.4, 0.08),
}
LABELS = ["background", *TISSUES]


def _per_axis_counts(spins_per_voxel, ndim: int) -> tuple[int, ...]:
    if isinstance(spins_per_voxel, int):
        side = round(spins_per_voxel ** (1 / ndim))
        if side**ndim != spins_per_voxel:
            raise ValueError(f"spins_per_voxel={spins_per_voxel} is not a perfect {ndim}-th power; "
                             f"give spins per side along each axis as a tuple instead")
        return (side,) * ndim
    counts = tuple(int(k) for k in spins_per_voxel)
    if len(counts) != ndim or min(counts) < 1:
        raise ValueError(f"spins_per_voxel needs {ndim} positive counts (map axis order), got {spins_per_voxel}")
    return counts


def phantom_from_maps(pd: Tensor, t1: Tensor, t2: Tensor, fov: float, db0: Tensor | None = None, *,
                      spins_per_voxel: int | tuple[int, ...] = 1, jitter: str | None = None, seed: int = 0,
                      t2prime: Tensor | None = None, t2prime_sampling: str = "quantile", b1: Tensor | None = None,
                      fat_fraction: Tensor | None = None, field_strength: float = 3.0, fat_t1=None, fat_t2=None,
                      fat_peaks=FAT_PEAKS) -> Phantom:
    """Spins from (z, y, x)-ordered maps; PD == 0 voxels are dropped. ``spins_per_voxel`` sub-voxel isochromats
    carry PD / K each, on a regular sub-grid or (``jitter="stratified"``) at a seeded uniform point per sub-cell.
    ``t2prime`` adds Lorentzian (Cauchy-quantile) offsets per voxel; ``fat_fraction`` adds multi-peak fat spins."""
    mask = pd > 0
    pos = grid_positions(pd.shape, fov, dtype=pd.dtype)[mask]
    db0 = torch.zeros_like(pd) if db0 is None else db0
    pick = lambda m: None if m is None else m[mask]
    pd_v, t1_v, t2_v, db0_v, b1_v = pd[mask], t1[mask], t2[mask], db0[mask], pick(b1)
    counts = _per_axis_counts(spins_per_voxel, pd.ndim)
    k = math.prod(counts)
    if jitter not in (None, "stratified"):
        raise ValueError(f"jitter must be None or 'stratified', got {jitter!r}")
    if t2prime_sampling not in ("quantile", "random"):
        raise ValueError(f"t2prime_sampling must be 'quantile' or 'random', got {t2prime_sampling!r}")
    if k == 1 and jitter is None and t2prime is None and fat_fraction is None:
        return Phantom(pos=pos, pd=pd_v, t1=t1_v, t2=t2_v, db0=db0_v, b1=b1_v)
    gen = torch.Generator().manual_seed(seed)
    n_vox = pos.shape[0]
    if k > 1 or jitter is not None:
        frac = torch.zeros(n_vox, k, 3, dtype=pd.dtype)
        cells = torch.cartesian_prod(*(torch.arange(c) for c in counts)).reshape(k, len(counts))
        u = (torch.rand(n_vox, k, len(counts), generator=gen, dtype=pd.dtype) if jitter == "stratified"
             else torch.full((n_vox, k, len(counts)), 0.5, dtype=pd.dtype))
        sides = torch.tensor(counts, dtype=pd.dtype)
        frac[..., : len(counts)] = ((cells + u) / sides - 0.5).flip(-1)  # (z, y, x) map axes -> (x, y, z)
        voxel = torch.tensor([fov / n for n in reversed(pd.shape)] + [0.0] * (3 - pd.ndim), dtype=pd.dtype)
        pos = (pos[:, None, :] + frac * voxel).reshape(-1, 3)
    rep = lambda v: None if v is None else v.repeat_interleave(k)
    pd_s, t1_s, t2_s, db0_s, b1_s = rep(pd_v / k if k > 1 else pd_v), rep(t1_v), rep(t2_v), rep(db0_v), rep(b1_v)
    if t2prime is not None:
        if t2prime_sampling == "quantile":
            q = (torch.arange(k, dtype=pd.dtype) + 0.5) / k
            q = q[torch.argsort(torch.rand(n_vox, k, generator=gen), dim=1)]
        else:
            q = torch.rand(n_vox, k, generator=gen, dtype=pd.dtype)
        half_width = 1.0 / (2 * math.pi * t2prime[mask])
        db0_s = db0_s + (half_width[:, None] * torch.tan(math.pi * (q - 0.5))).reshape(-1)
    if fat_fraction is None:
        return Phantom(pos=pos, pd=pd_s, t1=t1_s, t2=t2_s, db0=db0_s, b1=b1_s)
    ff = rep(fat_fraction[mask])
    fat = ff > 0
    peaks = torch.tensor(fat_peaks, dtype=pd.dtype).reshape(-1, 2)
    shift_hz, amp = peaks[:, 0] * (1e-6 * GAMMA_BAR * field_strength), peaks[:, 1] / peaks[:, 1].sum()
    n_peaks, n_fat = peaks.shape[0], int(fat.sum())

    def fat_param(v, default):
        v = torch.as_tensor(default if v is None else v, dtype=pd.dtype)
        return v.expand(n_fat) if v.ndim == 0 else rep(v[mask])[fat]

    tile = lambda v: v.repeat(n_peaks)
    return Phantom(
        pos=torch.cat([pos, pos[fat].repeat(n_peaks, 1)]),
        pd=torch.cat([pd_s * (1 - ff), ((pd_s * ff)[fat][None, :] * amp[:, None]).reshape(-1)]),
        t1=torch.cat([t1_s, tile(fat_param(fat_t1, TISSUES["fat"][1]))]),
        t2=torch.cat([t2_s, tile(fat_param(fat_t2, TISSUES["fat"][2]))]),
        db0=torch.cat([db0_s, (db0_s[fat][None, :] + shift_hz[:, None]).reshape(-1)]),
        b1=None if b1_s is None else torch.cat([b1_s, tile(b1_s[fat])]),
    )


# Modified Shepp-Logan (Toft): centre, semi-axes, rotation [deg] in [-1, 1] coordinates; later ellipses win.
_SL_2D = [
    (0.0, 0.0, 0.69, 0.92, 0.0, "fat"),
    (0.0, -0.0184, 0.6624, 0.874, 0.0, "wm"),
    (0.22, 0.0, 0.11, 0.31, -18.0, "csf"),
    (-0.22, 0.0, 0.16, 0.41, 18.0, "csf"),
    (0.0, 0.35, 0.21, 0.25, 0.0, "gm"),
    (0.0, 0.1, 0.046, 0.046, 0.0, "lesion_t2"),
    (0.0, -0.1, 0.046, 0.046, 0.0, "lesion_t2"),
    (-0.08, -0.605, 0.046, 0.023, 0.0, "lesion_t1"),
    (0.0, -0.606, 0.023, 0.023, 0.0, "lesion_t1"),
    (0.06, -0.605, 0.023, 0.046, 0.0, "lesion_t1"),
]
# 3D (Kak & Slaney ellipsoids without the out-of-plane tilts): centre, semi-axes, rotation about z [deg].
_SL_3D = [
    (0.0, 0.0, 0.0, 0.69, 0.92, 0.81, 0.0, "fat"),
    (0.0, -0.0184, 0.0, 0.6624, 0.874, 0.78, 0.0, "wm"),
    (0.22, 0.0, 0.0, 0.11, 0.31, 0.22, -18.0, "csf"),
    (-0.22, 0.0, 0.0, 0.16, 0.41, 0.28, 18.0, "csf"),
    (0.0, 0.35, -0.15, 0.21, 0.25, 0.41, 0.0, "gm"),
    (0.0, 0.1, 0.25, 0.046, 0.046, 0.05, 0.0, "lesion_t2"),
    (0.0, -0.1, 0.25, 0.046, 0.046, 0.05, 0.0, "lesion_t2"),
    (-0.08, -0.605, 0.0, 0.046, 0.023, 0.05, 0.0, "lesion_t1"),
    (0.0, -0.606, 0.0, 0.023, 0.023, 0.02, 0.0, "lesion_t1"),
    (0.06, -0.605, 0.0, 0.023, 0.046, 0.02, 0.0, "lesion_t1"),
]


def shepp_logan_maps(n: int, ndim: int = 2, dtype=torch.float64) -> dict[str, Tensor]:
    """Label, PD, T1, T2 maps (n,)*ndim ordered (z, y, x); background PD 0, T1 = T2 = 1 s."""
    x, y, z = grid_positions((n,) * ndim, 2.0, dtype=dtype).unbind(-1)
    shapes = ([(x0, y0, 0.0, a, b, math.inf, deg, t) for x0, y0, a, b, deg, t in _SL_2D] if ndim == 2 else _SL_3D)
    label = torch.zeros(x.shape, dtype=torch.long)
    for x0, y0, z0, a, b, c, deg, tissue in shapes:
        th = math.radians(deg)
        u = (x - x0) * math.cos(th) + (y - y0) * math.sin(th)
        v = -(x - x0) * math.sin(th) + (y - y0) * math.cos(th)
        label[(u / a) ** 2 + (v / b) ** 2 + ((z - z0) / c) ** 2 <= 1.0] = LABELS.index(tissue)
    table = torch.tensor([(0.0, 1.0, 1.0), *TISSUES.values()], dtype=dtype)
    props = table[label]
    return {"label": label, "pd": props[..., 0], "t1": props[..., 1], "t2": props[..., 2]}


def shepp_logan(n: int = 64, fov: float = 0.22, ndim: int = 2, dtype=torch.float64, *, chemical_shift: bool = False,
                **kwargs) -> Phantom:
    maps = shepp_logan_maps(n, ndim=ndim, dtype=dtype)
    if chemical_shift:
        kwargs["fat_fraction"] = (maps["label"] == LABELS.index("fat")).to(dtype)
    return phantom_from_maps(maps["pd"], maps["t1"], maps["t2"], fov, **kwargs)


# three tissue classes of the 3D workloads: (PD, T1 [s], T2 [s])
THREE_CLASSES = ((0.7, 0.8, 0.080), (0.8, 1.3, 0.110), (1.0, 4.0, 2.0))


def smooth_field(shape: tuple[int, ...], corr: float, seed: int, dtype=torch.float64) -> Tensor:
    """A zero-mean, unit-variance Gaussian random field on ``shape`` with a Gaussian correlation of ``corr`` voxels
    (white noise filtered in Fourier space), seeded."""
    gen = torch.Generator().manual_seed(seed)
    noise = torch.randn(shape, generator=gen, dtype=torch.float64)
    freqs = torch.meshgrid(*(torch.fft.fftfreq(n, dtype=torch.float64) for n in shape), indexing="ij")
    k2 = sum(f * f for f in freqs)
    f = torch.fft.ifftn(torch.fft.fftn(noise) * torch.exp(-2 * math.pi**2 * corr**2 * k2)).real
    f = f - f.mean()
    return (f / f.std()).to(dtype)


def three_class_maps(n: int, ndim: int = 3, *, seed: int = 1, db0_seed: int = 2, db0_max: float = 100.0,
                     corr: float = 4.0, dtype=torch.float64) -> dict[str, Tensor]:
    """PD / T1 / T2 maps of three tissue classes (``THREE_CLASSES``) drawn by thresholding a smooth random field
    at its terciles (seed ``seed``), and a smooth off-resonance map scaled to +-``db0_max`` Hz (seed ``db0_seed``)."""
    shape = (n,) * ndim
    g = smooth_field(shape, corr, seed)
    lo, hi = torch.quantile(g.flatten()[:: max(1, g.numel() // 65536)], torch.tensor([1 / 3, 2 / 3],
                                                                                      dtype=torch.float64))
    label = (g > lo).long() + (g > hi).long()
    table = torch.tensor(THREE_CLASSES, dtype=dtype)
    props = table[label]
    b = smooth_field(shape, corr, db0_seed)
    db0 = (b / b.abs().max() * db0_max).to(dtype)
    return {"label": label, "pd": props[..., 0], "t1": props[..., 1], "t2": props[..., 2], "db0": db0}
