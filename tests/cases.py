"""Shared test inputs and the comparison helper.

``rel`` is the reference suite's comparison (tests/helpers.py:11-20): max |a - b| / max |b| after casting ``a`` to
``b``'s dtype. The cases exercise what the reference's GPU tests exercise (tests/engines/test_gpu.py,
tests/cases.py:151-198): shaped RF with gradients and off-resonance, readouts split over runs, spoiling, RF phase.
"""

from __future__ import annotations

import math
from dataclasses import replace

import torch

import blochparty_amd as bp

f64 = torch.float64


def rel(a, b, scale=None) -> float:
    if b.numel() == 0:
        return 0.0
    if a.dtype != b.dtype:
        a = a.to(b.dtype)
    diff = (a - b).abs().max()
    scale = b.abs().max() if scale is None else scale
    return float(diff / scale) if scale > 0 else float(diff)


def random_case(seed: int, n: int = 6):
    """Shaped RF on a 10 us raster with gradients, hard pulses of random flip and phase, gradient delays, a spoil,
    and a final excitation; random spins."""
    g = torch.Generator().manual_seed(seed)
    rnd = lambda *shape: torch.rand(*shape, generator=g, dtype=f64)
    phantom = bp.Phantom(pos=(rnd(n, 3) - 0.5) * 0.2, pd=0.5 + rnd(n), t1=0.2 + rnd(n), t2=0.02 + 0.1 * rnd(n),
                         db0=(rnd(n) - 0.5) * 200)
    b = bp.SequenceBuilder()
    for _ in range(4):
        b.rf_waveform((rnd(40, 2) - 0.5) * 2000, 10e-6, grad=(rnd(3) - 0.5) * 2000)
        b.delay(rnd(()) * 3e-3, grad=(rnd(3) - 0.5) * 4000)
        b.rf(100e-6, flip_deg=180 * rnd(()), phase=2 * math.pi * rnd(()))
        b.delay(rnd(()) * 3e-3, grad=(rnd(3) - 0.5) * 4000)
    b.spoil()
    b.rf(200e-6, flip_deg=90.0, n_steps=10)
    b.delay(1e-3)
    b.readout(16, 10e-6, grad=(800.0, -300.0, 150.0), phase=0.7)
    return phantom, b.build()


def random_phantom(n=50, seed=0):
    g = torch.Generator().manual_seed(seed)
    rnd = lambda *s: torch.rand(*s, generator=g, dtype=f64)
    return bp.Phantom(pos=(rnd(n, 3) - 0.5) * 0.2, pd=0.5 + rnd(n), t1=0.2 + rnd(n), t2=0.02 + 0.1 * rnd(n),
                      db0=(rnd(n) - 0.5) * 200)


def busy_sequence(amp=2500.0, gx=2000.0, seed=0):
    """RF of varying phase, readouts split over runs with changing gradients, a spoiled run without ADC, an ADC
    right after RF."""
    g = torch.Generator().manual_seed(seed)
    rnd = lambda *s: torch.rand(*s, generator=g, dtype=f64)
    b = bp.SequenceBuilder()
    for i in range(4):
        b.rf(100e-6, amplitude=amp, phase=1.3 * i, n_steps=1)
        b.delay(0.4e-3, grad=(gx, (rnd(()) - 0.5) * 3000, 0.0))
        b.readout(8, 20e-6, grad=(-gx, 0.0, 400.0), phase=0.2 * i)
        b.delay(0.2e-3, grad=(rnd(3) - 0.5) * 2000)
        b.readout(5, 30e-6, grad=(rnd(3) - 0.5) * 2000)
        b.delay(1e-3).spoil()
    b.rf(50e-6, amplitude=amp, n_steps=1).readout(3, 10e-6).delay(2e-3, grad=(1e3, 0.0, 0.0)).spoil()
    return b.build()


def c1_case(n: int = 64, n_dummy: int = 200):
    """The plain 2D workload: Shepp-Logan n x n, 8 birdcage coils, spoiled GRE, 15 deg hard pulse, TR 10 ms, TE 5 ms."""
    ph = bp.shepp_logan(n)
    seq = bp.gre_2d(n, n_dummy=n_dummy, te=5e-3, flip_deg=15.0, tr=10e-3)
    return ph, seq, bp.birdcage_sensitivities(ph.pos, 8)


def weighted_loss(signal, m):
    """Linear in the outputs with fixed seeded float64 weights (the same for the GPU and the oracle)."""
    g = torch.Generator().manual_seed(1)
    w = lambda x: torch.randn(x.shape, generator=g, dtype=f64)
    return ((signal.real.double() * w(signal.real)).sum() + (signal.imag.double() * w(signal.imag)).sum()
            + (m.double() * w(m)).sum())


def occurrence_index(seq) -> torch.Tensor:
    """Per segment, the number of RF runs begun up to it (an occurrence's segments share it)."""
    is_rf = torch.tensor(seq.is_rf)
    start = is_rf & ~torch.cat([torch.tensor([False]), is_rf[:-1]])
    return torch.cumsum(start.to(f64), 0)


def map_gradients(sim, ph, seq, coils, m_init, **kw):
    """d loss / d(spin inputs, coils, m_init) and d loss / d(flip, theta) with rf = flip rf0 e^{i theta n}, n the
    occurrence's index: a common amplitude and a per-occurrence phase, both through rf."""
    leaves = {}

    def leaf(name, t):
        t = t.detach().clone().requires_grad_()
        leaves[name] = t
        return t

    p = replace(ph, **{k: leaf(k, getattr(ph, k)) for k in ("pos", "pd", "t1", "t2", "db0", "b1")
                       if getattr(ph, k) is not None})
    flip, theta = leaf("flip", torch.tensor(1.0, dtype=f64)), leaf("theta", torch.tensor(0.0, dtype=f64))
    rf0 = torch.complex(seq.rf[:, 0], seq.rf[:, 1])
    rf = rf0 * flip * torch.exp(1j * theta * occurrence_index(seq))
    s = replace(seq, rf=torch.view_as_real(rf))
    sig, m = sim(p, s, leaf("coils", coils), None if m_init is None else leaf("m_init", m_init), **kw)
    grads = torch.autograd.grad(weighted_loss(sig, m), list(leaves.values()))
    return dict(zip(leaves, grads))
