"""The benchmark workloads, built from the package's own builders with seeded synthetic inputs.

- ``gre2d``: 2D spoiled GRE 64 x 64 on a Shepp-Logan phantom (one spin per voxel, 2,039 spins), 8 birdcage coils,
  15 deg hard pulse, TR 10 ms, TE 5 ms, 200 dummy TRs.
- ``gre3d``: 3D spoiled GRE n^3 (n = 64: 262,144 voxels x 4 sub-voxel isochromats = 1,048,576 spins) on a three-class
  smooth random phantom (T1 0.8 / 1.3 / 4.0 s, T2 80 / 110 / 2000 ms, PD 0.7 / 0.8 / 1.0) with a smooth +-100 Hz
  off-resonance, 8 coils, a 2 ms windowed-sinc slab pulse (TBW 4, 200 substeps on a 10 us RF raster) of 10 deg over
  the field of view, TR 8 ms, n x n phase encodes x n readout samples.
"""

from __future__ import annotations

import torch

from .builder import System
from .coils import birdcage_sensitivities
from .library import gre_2d, gre_cartesian
from .phantom import Phantom, phantom_from_maps, shepp_logan, three_class_maps

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
