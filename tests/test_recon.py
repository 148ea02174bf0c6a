"""Downstream recon (``blochparty_amd.recon``) on the FP64 oracle's signals, CPU: a spoiled GRE's image against the
steady-state closed form (the reference's tests/sequences/test_pulseq.py:393-416 pattern), the k-space trajectory on
the GRE's and the TSE's Cartesian grids (refocusing negates it), coil combination, and the ISMRMRD acquisitions."""

import math

import numpy as np
import pytest
import torch

import blochparty_amd as bp
from blochparty_amd import recon
from oracle import core

f64 = torch.float64


def tissue_maps(n, seed=0):
    """PD / T1 / T2 maps with short T1 (a few hundred dummy TRs reach the steady state) and no background."""
    g = torch.Generator().manual_seed(seed)
    u = lambda lo, hi: lo + (hi - lo) * torch.rand(n, n, generator=g, dtype=f64)
    return u(0.5, 1.0), u(0.15, 0.4), u(0.05, 0.2)


def ernst(pd, t1, t2, tr, te, alpha):
    e1 = torch.exp(-tr / t1)
    return pd * math.sin(alpha) * (1 - e1) / (1 - math.cos(alpha) * e1) * torch.exp(-te / t2)


def test_gre_image_is_the_steady_state_closed_form():
    n, fov, flip, tr = 16, 0.22, 20.0, 10e-3
    pd, t1, t2 = tissue_maps(n)
    ph = bp.phantom_from_maps(pd, t1, t2, fov)
    seq = bp.gre_2d(n, fov, flip_deg=flip, tr=tr, te=2e-3, dwell=1e-6, n_dummy=300)
    coils = bp.birdcage_sensitivities(ph.pos, 4)
    sig, _ = core.simulate(ph, seq, coils)
    images = recon.reconstruct(sig, seq)
    te = float(seq.meta["te"])
    expected = 1j * ernst(pd, t1, t2, tr, te, math.radians(flip))
    # coils: the sensitivities at the voxel centres combine the coil images back into one
    sens = coils.T.reshape(4, n, n)
    img = recon.coil_combine(images, sens)
    err = float((img - expected).abs().max() / expected.abs().max())
    assert err < 5e-4, err  # T2 decay over the 16 us readout blurs at the 1e-4 level
    assert float((recon.rss(images) - (sens.abs() ** 2).sum(0).sqrt() * expected.abs()).abs().max()) < 1e-3


def test_trajectory_on_the_gre_and_tse_grids():
    n, fov = 8, 0.2
    grid = (np.arange(n) - n // 2) / fov
    gre = bp.gre_2d(n, fov, ramps=True, spoiling="rf", slice_thickness=5e-3, n_dummy=2)
    k = recon.kspace_trajectory(gre).numpy().reshape(n, n, 3)
    assert np.abs(k[..., 0] - grid[None, :]).max() < 1e-6 and np.abs(k[..., 1] - grid[:, None]).max() < 1e-6
    tse = bp.tse_2d(n, fov, etl=4, tr=0.3, ordering="centric", slice_thickness=5e-3, rf_duration=1e-3,
                    system=bp.System(rf_raster=20e-6))
    kt = recon.kspace_trajectory(tse).numpy()
    flat = np.asarray(tse.meta["kspace_index"])
    line, col = flat // n, flat % n
    assert np.abs(kt[:, 0] - grid[col]).max() < 1e-6  # every echo on its grid point, after the refocusings
    assert np.abs(kt[:, 1] - grid[line]).max() < 1e-6
    assert recon.flip_angle(np.full(50, 90.0 / 360 / 1e-3), 2e-5) == pytest.approx(90.0, abs=1e-9)


def test_tse_kspace_grid_and_acquisitions():
    n = 8
    ph = bp.shepp_logan(n)
    seq = bp.tse_2d(n, etl=4, tr=0.3)
    coils = bp.birdcage_sensitivities(ph.pos, 2)
    sig, _ = core.simulate(ph, seq, coils)
    k = recon.kspace_grid(sig, seq)
    flat = torch.as_tensor(seq.meta["kspace_index"])
    assert torch.equal(k.reshape(2, -1)[:, flat], sig)
    acqs = recon.acquisitions(sig, seq)
    assert len(acqs) == n and all(a["number_of_samples"] == n and a["data"].shape == (2, n) for a in acqs)
    lines = sorted(a["kspace_encode_step_1"] for a in acqs)
    assert lines == list(range(n))
    head = recon.ismrmrd_header(seq, 2)
    assert "<receiverChannels>2</receiverChannels>" in head and "<x>8</x>" in head
    try:
        import h5py  # noqa: F401
    except ImportError:
        with pytest.raises(ImportError, match="h5py"):
            recon.write_dataset("unused.h5", sig, seq)
