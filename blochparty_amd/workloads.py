"""The benchmark workloads, built from the package's own builders with seeded synthetic inputs.

- ``gre2d``: 2D spoiled GRE 64 x 64 on a Shepp-Logan phantom (one spin per voxel, 2,039 spins), 8 birdcage coils,
  15 deg hard pulse, TR 10 ms, TE 5 ms, 200 dummy TRs.
- ``gre3d``: 3D spoiled GRE n^3 (n = 64: 262,144 voxels x 4 sub-voxel isochromats = 1,048,576 spins) on a three-class
  smooth random phantom (T1 0.8 / 1.3 / 4.0 s, T2 80 / 110 / 2000 ms, PD 0.7 / 0.8 / 1.0) with a smooth +-100 Hz
  off-resonance, 8 coils, a 2 ms windowed-sinc slab pulse (TBW 4, 200 substeps on a 10 us RF raster) of 10 deg over
  the field of view, TR 8 ms, n x n phase encodes x n readout samples.
- ``multislice``: multi-slice 2D spoiled GRE, 5 slices of 3 mm (centres 6 mm apart) acquired one after another,
  256 x 256 in-plane, 8 isochromats through each slice's thickness (2,621,440 spins), three tissue classes per
  slice, B1 drawn uniformly from 0.7-1.3 per voxel, a smooth +-200 Hz off-resonance, 8 coils, a 3 ms windowed-sinc
  slice pulse (TBW 4, 300 substeps on a 10 us raster) of 30 deg shifted to each slice, TR 10 ms: the learned RF
  surrogate's workload (five distinct reused pulses, one per slice offset).
- ``mt_gre2d``: the full variant's workload: one slice of the multi-slice phantom (n x n, 4 isochromats through its
  3 mm), a super-Lorentzian bound pool by tissue class (f 0.15 / 0.10 / 0, k 30 Hz, T1b 1 s, T2b 12 us, dw -200 Hz),
  a 0.55 T scanner's leading-order concomitant fields (z asymmetry 0.5), 8 coils, a spoiled GRE with a 2 ms sinc
  slice pulse (TBW 4, 100 substeps on a 20 us raster) of 20 deg, TR 10 ms, RF spoiling, n x n samples.
"""

from __future__ import annotations

import torch

from .builder import System
from .coils import birdcage_sensitivities
from .library import gre_2d, gre_cartesian
from .phantom import THREE_CLASSES, Phantom, grid_positions, phantom_from_maps, shepp_logan, smooth_field, \
    three_class_maps
from .sequence import Sequence

FOV = 0.22


def gre2d(n: int = 64, n_dummy: int = 200, n_coils: int = 8):
    ph = shepp_logan(n, FOV)
    seq = gre_2d(n, FOV, n_dummy=n_dummy, te=5e-3, flip_deg=15.0, tr=10e-3)
    return ph, seq, birdcage_sensitivities(ph.pos, n_coils)


def gre3d_phantom(n: int = 64, spins_per_voxel=(1, 2, 2)) -> Phantom:
    maps = three_class_maps(n, 3, seed=1, db0_seed=2, db0_max=100.0)
    return phantom_from_maps(maps["pd"], maps["t1"], maps["t2"], FOV, maps["db0"], spins_per_voxel=spins_per_voxel)


def gre3d_sequence(n: int = 64, n_dummy: int = 0, flip_deg=10.0):
    """``flip_deg`` may be a tensor that requires grad (the RF waveform is then flip_deg x the shape per degree)."""
    return gre_cartesian((n, n, n), FOV, flip_deg=flip_deg, tr=8e-3, n_dummy=n_dummy, slice_thickness=FOV,
                         rf_duration=2e-3, time_bw_product=4.0, system=System(rf_raster=10e-6))


def gre3d(n: int = 64, n_coils: int = 8, spins_per_voxel=(1, 2, 2)):
    ph = gre3d_phantom(n, spins_per_voxel)
    return ph, gre3d_sequence(n), birdcage_sensitivities(ph.pos, n_coils)


def subset(ph: Phantom, coils: torch.Tensor, count: int, seed: int = 0):
    """A seeded subset of ``count`` spins (spins are independent: per-spin parity is exact on any subset)."""
    g = torch.Generator().manual_seed(seed)
    idx = torch.randperm(len(ph), generator=g)[:count].sort().values
    pick = lambda t: None if t is None else t[idx]
    return Phantom(ph.pos[idx], ph.pd[idx], ph.t1[idx], ph.t2[idx], ph.db0[idx], b1=pick(ph.b1)), coils[idx]


def first_trs(seq, n_tr: int):
    """The first ``n_tr`` spoiled TRs of a sequence (cut after the n_tr-th spoiled segment)."""
    import numpy as np

    from .sequence import Sequence

    ends = np.flatnonzero(np.asarray(seq.spoil, dtype=bool))
    e = int(ends[n_tr - 1]) + 1 if len(ends) >= n_tr else len(seq)
    return Sequence(seq.dt[:e], seq.rf[:e], seq.grad[:e], seq.adc_phase[:e], seq.is_rf[:e], seq.adc[:e],
                    seq.spoil[:e], dict(seq.meta), None if seq.b0_offset is None else seq.b0_offset[:e])


def concat(seqs: list[Sequence]) -> Sequence:
    """The sequences played one after another (their meta from the first, with each part's segment count)."""
    cat = lambda name: torch.cat([getattr(q, name) for q in seqs])
    flags = lambda name: [v for q in seqs for v in getattr(q, name)]
    meta = dict(seqs[0].meta)
    meta["parts"] = [len(q) for q in seqs]
    return Sequence(cat("dt"), cat("rf"), cat("grad"), cat("adc_phase"), flags("is_rf"), flags("adc"),
                    flags("spoil"), meta)


SLICES, SLICE_THICKNESS, SLICE_SPACING, SLICE_ISO = 5, 3e-3, 6e-3, 8


def slice_centres(slices: int = SLICES, spacing: float = SLICE_SPACING) -> list[float]:
    return [(s - (slices - 1) / 2) * spacing for s in range(slices)]


def multislice_sequence(n: int = 256, lines: int | None = None, slices: int = SLICES, flip_deg=30.0,
                        tr: float = 10e-3) -> Sequence:
    """Each slice in turn: a spoiled GRE of ``lines`` (default n) phase encodes x n samples with the 3 ms sinc
    (TBW 4) shifted to the slice's centre."""
    parts = [gre_cartesian((lines or n, n), FOV, flip_deg=flip_deg, tr=tr, slice_thickness=SLICE_THICKNESS,
                           slice_offset=z, rf_duration=3e-3, time_bw_product=4.0, system=System(rf_raster=10e-6))
             for z in slice_centres(slices)]
    return concat(parts)


def multislice_phantom(n: int = 256, slices: int = SLICES, iso: int = SLICE_ISO, seed: int = 1) -> Phantom:
    """Per slice an n x n grid of three-class voxels (seeded smooth field thresholded at its terciles), each voxel
    ``iso`` isochromats at equal steps through the slice's thickness, PD / iso each; B1 ~ U(0.7, 1.3) per voxel
    (seed 4), off-resonance a smooth field scaled to +-200 Hz (seed 2). Spins slice by slice, voxel by voxel."""
    shape = (slices, n, n)
    field = smooth_field(shape, 4.0, seed)
    lo, hi = torch.quantile(field.flatten()[:: max(1, field.numel() // 65536)],
                            torch.tensor([1 / 3, 2 / 3], dtype=torch.float64))
    label = ((field > lo).long() + (field > hi).long()).reshape(-1)
    props = torch.tensor(THREE_CLASSES, dtype=torch.float64)[label]
    db0 = smooth_field(shape, 8.0, 2)
    db0 = (db0 / db0.abs().max() * 200.0).reshape(-1)
    b1 = 0.7 + 0.6 * torch.rand(slices * n * n, generator=torch.Generator().manual_seed(4), dtype=torch.float64)
    plane = grid_positions(n, FOV).reshape(-1, 3)  # z = 0
    steps = ((torch.arange(iso, dtype=torch.float64) + 0.5) / iso - 0.5) * SLICE_THICKNESS
    z = torch.tensor(slice_centres(slices), dtype=torch.float64)[:, None, None] + steps[None, None, :]
    pos = plane[None, :, None, :].expand(slices, n * n, iso, 3).clone()
    pos[..., 2] = z.expand(slices, n * n, iso)
    rep = lambda v: v.repeat_interleave(iso)
    return Phantom(pos=pos.reshape(-1, 3), pd=rep(props[:, 0]) / iso, t1=rep(props[:, 1]), t2=rep(props[:, 2]),
                   db0=rep(db0), b1=rep(b1))


MT_POOL_F = (0.15, 0.10, 0.0)  # bound-pool fraction per tissue class (the third, fluid, has none)


def mt_gre2d(n: int = 128, n_coils: int = 8, iso: int = 4, field_strength: float = 0.55):
    """(phantom with a bound pool, sequence, coils, scanner) of the full variant's workload (module docstring)."""
    from dataclasses import replace

    from .exchange import BoundPool
    from .fields import Scanner

    ph = multislice_phantom(n, slices=1, iso=iso)
    t1_class = torch.tensor([c[1] for c in THREE_CLASSES], dtype=torch.float64)
    label = (ph.t1[:, None] - t1_class[None]).abs().argmin(1)
    pool = BoundPool(f=torch.tensor(MT_POOL_F, dtype=torch.float64)[label], k=30.0, t1=1.0, t2=12e-6, dw=-200.0)
    seq = gre_cartesian((n, n), FOV, flip_deg=20.0, tr=10e-3, slice_thickness=SLICE_THICKNESS, rf_duration=2e-3,
                        time_bw_product=4.0, spoiling="rf", system=System(rf_raster=20e-6))
    return (replace(ph, bound=pool), seq, birdcage_sensitivities(ph.pos, n_coils),
            Scanner(field_strength=field_strength))


def multislice(n: int = 256, lines: int | None = None, n_coils: int = 8, slices: int = SLICES):
    ph = multislice_phantom(n, slices)
    return ph, multislice_sequence(n, lines, slices), birdcage_sensitivities(ph.pos, n_coils)
