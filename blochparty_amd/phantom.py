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
    bound: object | None = None  # magnetization transfer: a bound pool per spin (exchange.BoundPool)
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


def _axis_counts(spins_per_voxel, ndim: int) -> tuple[int, ...]:
    """Isochromats per voxel along each map axis: a tuple as given, or an int that is a perfect ndim-th power."""
    if not isinstance(spins_per_voxel, int):
        counts = tuple(int(k) for k in spins_per_voxel)
        if len(counts) != ndim or min(counts) < 1:
            raise ValueError(f"spins_per_voxel needs {ndim} positive counts (map axis order), got {spins_per_voxel}")
        return counts
    side = round(spins_per_voxel ** (1 / ndim))
    if side**ndim != spins_per_voxel:
        raise ValueError(f"spins_per_voxel={spins_per_voxel} is not a perfect {ndim}-th power; "
                         f"give spins per side along each axis as a tuple instead")
    return (side,) * ndim


def _subcell_offsets(counts: tuple[int, ...], n_vox: int, stratified: bool, gen, dtype) -> Tensor:
    """(n_vox, k, 3) offsets from the voxel centre in voxel units, (x, y, z) order: sub-cell c of count m along a map
    axis sits at (c + u) / m - 1/2, u = 1/2 (the sub-cell's centre) or, stratified, a seeded uniform draw per spin
    and axis. Sub-cells are numbered with the last map axis fastest."""
    ndim, k = len(counts), math.prod(counts)
    index = torch.stack(torch.meshgrid(*(torch.arange(m, dtype=dtype) for m in counts), indexing="ij"), -1)
    index = index.reshape(k, ndim)
    u = (torch.rand(n_vox, k, ndim, generator=gen, dtype=dtype) if stratified
         else torch.full((n_vox, k, ndim), 0.5, dtype=dtype))
    per_axis = (index + u) / torch.tensor(counts, dtype=dtype) - 0.5  # map axis order (z, y, x) / (y, x)
    out = torch.zeros(n_vox, k, 3, dtype=dtype)
    out[..., :ndim] = per_axis.flip(-1)
    return out


def _lorentzian_offsets(hz_half_width: Tensor, k: int, quantiles: bool, gen, dtype) -> Tensor:
    """(n_vox * k,) off-resonance offsets of k isochromats per voxel drawn from a Cauchy (Lorentzian) line of the
    given half width: at the k equal-probability quantile midpoints in a seeded random order per voxel, or at
    independent uniform probabilities. F^-1(p) = w tan(pi (p - 1/2))."""
    n_vox = hz_half_width.shape[0]
    draw = torch.rand(n_vox, k, generator=gen, dtype=dtype)
    if quantiles:
        rank = draw.argsort(dim=1).argsort(dim=1)
        draw = (rank.to(dtype) + 0.5) / k
    return (hz_half_width[:, None] * torch.tan(math.pi * (draw - 0.5))).reshape(-1)


def phantom_from_maps(pd: Tensor, t1: Tensor, t2: Tensor, fov: float, db0: Tensor | None = None, *,
                      spins_per_voxel: int | tuple[int, ...] = 1, jitter: str | None = None, seed: int = 0,
                      t2prime: Tensor | None = None, t2prime_sampling: str = "quantile", b1: Tensor | None = None,
                      fat_fraction: Tensor | None = None, field_strength: float = 3.0, fat_t1=None, fat_t2=None,
                      fat_peaks=FAT_PEAKS) -> Phantom:
    """Spins from maps ordered (z, y, x) or (y, x) on ``fov``; voxels with PD 0 hold none.

    - ``spins_per_voxel`` isochromats per voxel (int or per-axis counts), each with PD / k, spin (voxel v, sub-cell
      c) at index v k + c; on the sub-cells' centres, or (``jitter="stratified"``) at a seeded uniform point in
      each sub-cell;
    - ``t2prime``: each voxel's isochromats get Lorentzian off-resonance offsets of half width 1 / (2 pi T2'),
      ``t2prime_sampling`` "quantile" (stratified, shuffled) or "random";
    - ``fat_fraction``: the water spins keep PD (1 - ff); after them come the fat spins of every peak of
      ``fat_peaks`` (ppm, relative amplitude) in turn, PD ff x amplitude, shifted by ppm x gamma_bar x B0, with fat
      T1 / T2 (``fat_t1``, ``fat_t2``, default the ``TISSUES`` fat values).
    """
    if jitter not in (None, "stratified"):
        raise ValueError(f"jitter must be None or 'stratified', got {jitter!r}")
    if t2prime_sampling not in ("quantile", "random"):
        raise ValueError(f"t2prime_sampling must be 'quantile' or 'random', got {t2prime_sampling!r}")
    dtype = pd.dtype
    occupied = pd > 0
    centres = grid_positions(pd.shape, fov, dtype=dtype)[occupied]
    per_voxel = lambda m: None if m is None else m[occupied]
    water = {"pd": pd[occupied], "t1": t1[occupied], "t2": t2[occupied],
             "db0": torch.zeros_like(pd)[occupied] if db0 is None else db0[occupied], "b1": per_voxel(b1)}
    counts = _axis_counts(spins_per_voxel, pd.ndim)
    k = math.prod(counts)
    gen = torch.Generator().manual_seed(seed)
    if k > 1 or jitter is not None:
        voxel = torch.tensor([fov / m for m in reversed(pd.shape)] + [0.0] * (3 - pd.ndim), dtype=dtype)
        offsets = _subcell_offsets(counts, centres.shape[0], jitter == "stratified", gen, dtype)
        pos = (centres[:, None, :] + offsets * voxel).reshape(-1, 3)
    else:
        pos = centres
    spins = {name: None if v is None else v.repeat_interleave(k) for name, v in water.items()}
    if k > 1:
        spins["pd"] = spins["pd"] / k
    if t2prime is not None:
        half_width = 1.0 / (2 * math.pi * t2prime[occupied])
        spins["db0"] = spins["db0"] + _lorentzian_offsets(half_width, k, t2prime_sampling == "quantile", gen, dtype)
    if fat_fraction is None:
        return Phantom(pos=pos, **spins)
    ff = fat_fraction[occupied].repeat_interleave(k)
    fatty = ff > 0
    table = torch.tensor(fat_peaks, dtype=dtype).reshape(-1, 2)
    shift = table[:, 0] * (1e-6 * GAMMA_BAR * field_strength)
    weight = table[:, 1] / table[:, 1].sum()
    n_peaks = table.shape[0]

    def fat_value(given, tissue_default):
        v = torch.as_tensor(tissue_default if given is None else given, dtype=dtype)
        v = v.expand(int(fatty.sum())) if v.ndim == 0 else v[occupied].repeat_interleave(k)[fatty]
        return v.repeat(n_peaks)

    fat_pd = (spins["pd"] * ff)[fatty]
    return Phantom(
        pos=torch.cat([pos, pos[fatty].repeat(n_peaks, 1)]),
        pd=torch.cat([spins["pd"] * (1 - ff), (weight[:, None] * fat_pd[None, :]).reshape(-1)]),
        t1=torch.cat([spins["t1"], fat_value(fat_t1, TISSUES["fat"][1])]),
        t2=torch.cat([spins["t2"], fat_value(fat_t2, TISSUES["fat"][2])]),
        db0=torch.cat([spins["db0"], (spins["db0"][fatty][None, :] + shift[:, None]).reshape(-1)]),
        b1=None if spins["b1"] is None else torch.cat([spins["b1"], spins["b1"][fatty].repeat(n_peaks)]),
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
