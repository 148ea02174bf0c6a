"""NUFFT readouts on the GPU (readouts="nufft") against the per-sample path and the FP64 oracle (the reference's
tests/engines/test_gpu_nufft.py:52-76 checks NUFFT against per-sample at 1e-4, with dispatch splits)."""

import dataclasses

import pytest
import torch

import blochparty_amd as bp
from blochparty_amd import workloads as W
from blochparty_amd.engines import driver
from cases import rel
from oracle import core

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gre32():
    ph, seq, coils = W.gre3d(32, spins_per_voxel=(1, 1, 1))
    return ph, W.first_trs(seq, 6), coils


def test_nufft_matches_per_sample_and_oracle(gpu, gre32):
    ph, seq, coils = gre32
    direct = bp.simulate(ph, seq, coils, engine="gpu")
    nu = bp.simulate(ph, seq, coils, engine="gpu", readouts="nufft")
    e = rel(nu.signal, direct.signal)
    print(f"nufft vs per-sample: signal {e:.2e}, m {rel(nu.m, direct.m):.2e}")
    assert e < 1e-4
    assert rel(nu.m, direct.m) < 1e-5  # the spins come back in the caller's order
    ref_sig, ref_m = core.simulate(ph, seq, coils)
    assert rel(nu.signal, ref_sig) < 1e-4
    assert rel(nu.m, ref_m) < 1e-5


@pytest.mark.parametrize("n_coils", [16, 3])
def test_nufft_coil_counts(gpu, gre32, n_coils):
    """Two coil tiles (16 coils) and a partial tile (3): the window columns' contraction and the grid's values."""
    ph, seq, _ = gre32
    coils = bp.birdcage_sensitivities(ph.pos, n_coils)
    direct = bp.simulate(ph, seq, coils, engine="gpu")
    nu = bp.simulate(ph, seq, coils, engine="gpu", readouts="nufft")
    ref_sig, _ = core.simulate(ph, seq, coils)
    print(f"{n_coils} coils: nufft vs per-sample {rel(nu.signal, direct.signal):.2e}, vs oracle "
          f"{rel(nu.signal, ref_sig):.2e}")
    assert rel(nu.signal, direct.signal) < 1e-4 and rel(nu.signal, ref_sig) < 1e-4


def test_nufft_width_terms_and_dispatch_splits(gpu, gre32):
    ph, seq, coils = gre32
    direct = bp.simulate(ph, seq, coils, engine="gpu")
    wide = bp.simulate(ph, seq, coils, engine="gpu", readouts="nufft", nufft_width=8, nufft_terms=4)
    print(f"width 8, 4 terms: {rel(wide.signal, direct.signal):.2e}")
    assert rel(wide.signal, direct.signal) < 1e-4
    split = bp.simulate(ph, seq, coils, engine="gpu", readouts="nufft", chunk=301)
    base = bp.simulate(ph, seq, coils, engine="gpu", readouts="nufft")
    assert rel(split.signal, base.signal) < 1e-6
    again = bp.simulate(ph, seq, coils, engine="gpu", readouts="nufft")
    assert torch.equal(again.signal, base.signal)  # deterministic: no atomics


def test_nufft_with_reused_pulses_and_b0_offset(gpu):
    ph, seq, coils = W.gre3d(32, spins_per_voxel=(1, 1, 1))
    seq = W.first_trs(seq, 4)
    seq = dataclasses.replace(seq, b0_offset=torch.full((len(seq),), 37.0, dtype=torch.float64))
    direct = bp.simulate(ph, seq, coils, engine="gpu", pulses="reuse")
    nu = bp.simulate(ph, seq, coils, engine="gpu", pulses="reuse", readouts="nufft")
    assert rel(nu.signal, direct.signal) < 1e-4


def test_nufft_refuses_timing_gradients_and_falls_back_without_runs(gpu):
    ph, seq, coils = bp.Phantom(pos=torch.zeros(4, 3, dtype=torch.float64), pd=torch.ones(4, dtype=torch.float64),
                                t1=torch.ones(4, dtype=torch.float64), t2=torch.full((4,), 0.1, dtype=torch.float64),
                                db0=torch.zeros(4, dtype=torch.float64)), W.first_trs(W.gre3d_sequence(32), 2), None
    seq.dt.requires_grad_(True)
    with pytest.raises(ValueError, match="readouts='direct'"):
        bp.simulate(ph, seq, engine="gpu", readouts="nufft")
    seq.dt.requires_grad_(False)
    short = W.first_trs(W.gre3d_sequence(16), 2)  # readouts too short: the per-sample path, same result
    case = driver.prepare_case(ph, short, None, nufft=(6, 3))
    assert case.nu is None
    assert rel(bp.simulate(ph, short, engine="gpu", readouts="nufft").signal,
               bp.simulate(ph, short, engine="gpu").signal) == 0
