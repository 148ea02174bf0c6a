"""The bound-pool (magnetization transfer, rate model) forward kernel against the FP64 oracle (oracle/exchange.py),
GPU: a slice-selective spoiled GRE with a super-Lorentzian pool (the RF's offset from each spin's resonance sets its
saturation: incidental MT), a Lorentzian pool with an off-resonant saturation train that only the bound pool feels,
and an initial magnetization with the bound pool's Mzb."""

from dataclasses import replace

import pytest
import torch

import blochparty_amd as bp
from cases import rel
from oracle import exchange as E

pytestmark = pytest.mark.gpu
f64 = torch.float64


def mt_phantom(n=12, seed=0, lineshape="superlorentzian"):
    ph = bp.shepp_logan(n)
    g = torch.Generator().manual_seed(seed)
    N = len(ph)
    pos = ph.pos.clone()
    pos[:, 2] = (torch.rand(N, generator=g, dtype=f64) - 0.5) * 8e-3
    pool = bp.BoundPool(f=0.05 + 0.15 * torch.rand(N, generator=g, dtype=f64),
                        k=10 + 30 * torch.rand(N, generator=g, dtype=f64), t1=1.0,
                        t2=8e-6 + 6e-6 * torch.rand(N, generator=g, dtype=f64), dw=-200.0, lineshape=lineshape)
    return replace(ph, pos=pos, b1=torch.linspace(0.85, 1.15, N, dtype=f64),
                   db0=(torch.rand(N, generator=g, dtype=f64) - 0.5) * 80, bound=pool)


def compare(ph, seq, coils, m_init=None, tol=1e-5):
    ref_s, ref_m = E.simulate_bound(ph, seq, coils, m_init=m_init)
    out = bp.simulate(ph, seq, coils, engine="gpu", m_init=m_init)
    assert out.m.shape == (len(ph), 3) and out.m_bound.shape == (len(ph),)  # the reference's SimResult fields
    m4 = torch.cat([out.m, out.m_bound[:, None]], 1).double()
    e_s, e_m = rel(out.signal.to(torch.complex128), ref_s), rel(m4, ref_m)
    print(f"bound pool vs oracle: signal {e_s:.1e}, m {e_m:.1e}", flush=True)
    assert e_s < tol and e_m < tol, (e_s, e_m)
    return out, ref_m


def test_slice_selective_gre_with_a_superlorentzian_pool(gpu):
    ph = mt_phantom()
    seq = bp.gre_2d(12, ramps=True, spoiling="rf", slice_thickness=5e-3, rf_duration=1e-3, n_dummy=4,
                    flip_deg=25.0, system=bp.System(rf_raster=20e-6))
    out, ref_m = compare(ph, seq, bp.birdcage_sensitivities(ph.pos, 3))
    assert float(ref_m[:, 3].min()) < float((ph.bound.f * ph.pd).min())  # the bound pool really is saturated


def test_bound_only_saturation_with_a_lorentzian_pool(gpu):
    ph = mt_phantom(8, seed=1, lineshape="lorentzian")
    b = bp.SequenceBuilder()
    for i in range(6):  # off-resonant saturation that only the bound pool feels, then an excitation and a readout
        b.rf(2e-3, amplitude=300.0, n_steps=8)
        b.delay(1e-3)
        b.rf(40e-6, flip_deg=20.0, phase=0.3 * i, n_steps=2)
        b.delay(2e-3, grad=(300.0, 0.0, 0.0))
        b.readout(8, 20e-6, grad=(-400.0, 100.0, 0.0))
        b.delay(5e-3).spoil()
    seq = b.build()
    only = [False] * len(seq)
    pos = 0
    for _ in range(6):
        for s in range(8):
            only[pos + s] = True
        pos += 8 + 1 + 2 + 1 + 9 + 1
    seq.bound_only = only
    seq.rf_freq = torch.where(torch.tensor(only), torch.tensor(2000.0, dtype=f64), torch.zeros(len(seq), dtype=f64))
    compare(ph, seq, bp.birdcage_sensitivities(ph.pos, 2))


def test_initial_magnetization_with_the_bound_pool(gpu):
    ph = mt_phantom(8, seed=2, lineshape="lorentzian")
    seq = bp.gre_2d(8, n_dummy=2)
    g = torch.Generator().manual_seed(3)
    m0 = torch.randn(len(ph), 4, generator=g, dtype=f64) * 0.3
    compare(ph, seq, bp.birdcage_sensitivities(ph.pos, 2), m_init=m0)


def test_bound_pool_refusals(gpu):
    ph = mt_phantom(6)
    seq = bp.gre_2d(6)
    with pytest.raises(NotImplementedError):
        bp.simulate(ph, seq, engine="gpu", pulses="reuse")
    with pytest.raises(NotImplementedError):
        bp.simulate(replace(ph, bound=replace(ph.bound, cutoff=None)), seq, engine="gpu")
    t1 = ph.t1.clone().requires_grad_()
    with pytest.raises(NotImplementedError):
        bp.simulate(replace(ph, t1=t1), seq, engine="gpu", readouts="nufft")
