"""Gradients through the full variant (hb_adjoint_full) against the FP64 oracle's autograd: a Lorentzian bound pool
with RF only the bound pool feels, every input of the reference's contract (spins, pool, coils, m_init with Mzb,
dt, rf, grad, b0_offset, the RF frequency); a super-Lorentzian pool under a slice-selective GRE (the oracle's F by
Gauss-Legendre quadrature in torch, so that its slope is differentiated too); exact concomitant fields with 18 coils
(two sweeps of 16); leading-order fields with a bound pool and a gradient nonlinearity; the checkpoint interval."""

import math
from dataclasses import replace

import numpy as np
import pytest
import torch

import blochparty_amd as bp
from blochparty_amd import fields as F
from cases import rel, weighted_loss
from oracle import core
from oracle import exchange as E
from oracle import fields as OF

pytestmark = pytest.mark.gpu
f64 = torch.float64

SPIN = ("pos", "pd", "t1", "t2", "db0", "b1")
POOL = ("f", "k", "t1", "t2", "dw")
SEQ = ("dt", "rf", "grad", "b0_offset", "rf_freq")


def leaves_of(ph, seq, coils, m_init, seq_keys=SEQ):
    """Copies of every input as leaves: (phantom, sequence, coils, m_init, {name: leaf})."""
    out = {}

    def leaf(name, t):
        t = torch.as_tensor(t)
        t = (t if t.is_complex() else t.to(f64)).detach().clone().requires_grad_()
        out[name] = t
        return t

    p = replace(ph, **{k: leaf(k, getattr(ph, k)) for k in SPIN if getattr(ph, k) is not None})
    if ph.bound is not None:
        b = ph.bound.expand(len(ph), ph.pd)
        p = replace(p, bound=replace(b, **{k: leaf("bound_" + k, getattr(b, k)) for k in POOL}))
    s = replace(seq, **{k: leaf(k, getattr(seq, k)) for k in seq_keys if getattr(seq, k) is not None})
    c = None if coils is None else leaf("coils", coils.to(torch.complex128))
    m0 = None if m_init is None else leaf("m_init", m_init)
    return p, s, c, m0, out


def gradients(run, ph, seq, coils=None, m_init=None, seq_keys=SEQ):
    p, s, c, m0, leaves = leaves_of(ph, seq, coils, m_init, seq_keys)
    sig, m = run(p, s, c, m0)
    grads = torch.autograd.grad(weighted_loss(sig, m), list(leaves.values()), allow_unused=True)
    return {k: g for k, g in zip(leaves, grads) if g is not None}


def gpu_run(**kw):
    def run(p, s, c, m0):
        r = bp.simulate(p, s, c, m_init=m0, engine="gpu", **kw)
        m = r.m if r.m_bound is None else torch.cat([r.m, r.m_bound[:, None]], 1)
        return r.signal, m

    return run


def compare(ref, out, tol, loose=(), loose_tol=None):
    errors = {k: rel(out[k], ref[k]) for k in ref if ref[k].abs().max() > 0}
    print({k: f"{v:.1e}" for k, v in errors.items()}, flush=True)
    assert set(ref) <= set(out), set(ref) - set(out)
    for k, e in errors.items():
        assert e < (loose_tol if k in loose else tol), (k, e)
    return errors


def pool_phantom(n, seed, lineshape, f0=0.1):
    ph = bp.shepp_logan(n)
    g = torch.Generator().manual_seed(seed)
    N = len(ph)
    pos = ph.pos.clone()
    pos[:, 2] = (torch.rand(N, generator=g, dtype=f64) - 0.5) * 8e-3
    pool = bp.BoundPool(f=f0 + 0.1 * torch.rand(N, generator=g, dtype=f64),
                        k=10 + 30 * torch.rand(N, generator=g, dtype=f64), t1=0.8 + 0.4 * torch.rand(N, generator=g,
                                                                                                     dtype=f64),
                        t2=(8e-6 if lineshape == "superlorentzian" else 2e-4) * (1 + 0.5 * torch.rand(N, generator=g,
                                                                                                       dtype=f64)),
                        dw=-200.0 + 50 * torch.rand(N, generator=g, dtype=f64), lineshape=lineshape)
    return replace(ph, pos=pos, b1=torch.linspace(0.85, 1.15, N, dtype=f64),
                   db0=(torch.rand(N, generator=g, dtype=f64) - 0.5) * 80, bound=pool)


def saturation_sequence():
    """Off-resonant saturation that only the bound pool feels (RF frequency 400 Hz), then an excitation, a readout
    and a spoil, six times; a b0_offset drift."""
    b = bp.SequenceBuilder()
    for i in range(6):
        b.rf(2e-3, amplitude=250.0, n_steps=8)
        b.delay(1e-3)
        b.rf(40e-6, flip_deg=25.0, phase=0.3 * i, n_steps=2)
        b.delay(2e-3, grad=(300.0, 0.0, 0.0))
        b.readout(8, 20e-6, grad=(-400.0, 100.0, 0.0))
        b.delay(5e-3).spoil()
    seq = b.build()
    # the builder's segment order: 8 saturation, 1 delay, 2 excitation, 1 delay, 9 readout, 1 delay per block
    mask, pos = [False] * len(seq), 0
    for _ in range(6):
        for s in range(8):
            mask[pos + s] = True
        pos += 8 + 1 + 2 + 1 + 9 + 1
    seq.bound_only = mask
    seq.rf_freq = torch.where(torch.tensor(mask), torch.tensor(400.0, dtype=f64), torch.zeros(len(seq), dtype=f64))
    seq.b0_offset = 3.0 * torch.sin(torch.arange(len(seq), dtype=f64) * 0.1)
    return seq


def test_lorentzian_pool_every_input(gpu):
    ph = pool_phantom(8, 0, "lorentzian")
    seq = saturation_sequence()
    coils = bp.birdcage_sensitivities(ph.pos, 2)
    m0 = torch.randn(len(ph), 4, generator=torch.Generator().manual_seed(3), dtype=f64) * 0.3
    m0[:, 2] += ph.pd
    m0[:, 3] = ph.bound.f * ph.pd * 0.7

    def oracle(p, s, c, m):
        return E.simulate_bound(p, s, c, m_init=m)

    ref = gradients(oracle, ph, seq, coils, m0)
    out = gradients(gpu_run(checkpoint_every=64), ph, seq, coils, m0)
    compare(ref, out, 1e-4)


ORACLE_G = E.lineshape_G


def torch_sl_lineshape(kind, delta, t2b, cutoff, nodes=256):
    """The oracle's lineshape with F(x) by Gauss-Legendre quadrature in torch (so that autograd sees its slope):
    F = sqrt(2 pi) u0 [int_0^{pi/2} exp(-2 x^2 / cos^4) / cos + int_0^{acosh sqrt 3} exp(-2 x^2 / sinh^4) / sinh]."""
    if kind == "lorentzian":
        return ORACLE_G(kind, delta, t2b, cutoff)
    x = 2 * math.pi * delta.abs().clamp_min(cutoff) * t2b
    t, w = (torch.as_tensor(v, dtype=f64) for v in np.polynomial.legendre.leggauss(nodes))
    u0 = 1.0 / math.sqrt(3.0)

    def integral(a, b, fn):
        y = 0.5 * (b - a) * (t + 1) + a
        return 0.5 * (b - a) * (w * fn(y)).sum(-1)

    xx = x[..., None] ** 2
    i1 = integral(0.0, math.pi / 2, lambda p: torch.exp(-2 * xx / torch.cos(p) ** 4) / torch.cos(p))
    i2 = integral(0.0, math.acosh(math.sqrt(3.0)), lambda p: torch.exp(-2 * xx / torch.sinh(p) ** 4) / torch.sinh(p))
    return t2b * math.sqrt(2 * math.pi) * u0 * (i1 + i2)


def test_superlorentzian_pool_slice_selective_gre(gpu, monkeypatch):
    """The saturation offset of each spin follows the slice gradient (incidental MT). Tolerances: 1e-4, and 3e-3 for
    T2b and dw, which act only through the lineshape: the kernels interpolate ln F linearly in ln x on steps of
    0.0031, so their slope of G differs from the exact one by up to about step / 2 |(ln F)''|."""
    x = torch.tensor([0.01, 0.1, 0.5, 1.0, 2.0], dtype=f64)
    F_quad = torch.as_tensor(E.sl_F(x.numpy()))
    F_gl = torch_sl_lineshape("superlorentzian", x / (2 * math.pi), torch.ones(5, dtype=f64), 0.0)
    assert rel(F_gl, F_quad) < 1e-10  # the quadrature pinned against the oracle's adaptive one
    ph = pool_phantom(8, 1, "superlorentzian")
    seq = bp.gre_2d(8, ramps=True, spoiling="rf", slice_thickness=5e-3, rf_duration=1e-3, n_dummy=3, flip_deg=25.0,
                    system=bp.System(rf_raster=40e-6))
    coils = bp.birdcage_sensitivities(ph.pos, 3)
    monkeypatch.setattr(E, "lineshape_G", torch_sl_lineshape)

    def oracle(p, s, c, m):
        return E.simulate_bound(p, s, c, m_init=m)

    keys = ("dt", "rf", "grad")
    ref = gradients(oracle, ph, seq, coils, None, keys)
    out = gradients(gpu_run(checkpoint_every=128), ph, seq, coils, None, keys)
    compare(ref, out, 1e-4, loose=("bound_t2", "bound_dw"), loose_tol=3e-3)


def field_case(n=8, seed=0):
    ph = bp.shepp_logan(n)
    g = torch.Generator().manual_seed(seed)
    N = len(ph)
    pos = ph.pos.clone()
    pos[:, 2] = (torch.rand(N, generator=g, dtype=f64) - 0.5) * 0.08
    ph = replace(ph, pos=pos, b1=torch.linspace(0.9, 1.1, N, dtype=f64),
                 db0=(torch.rand(N, generator=g, dtype=f64) - 0.5) * 60)
    seq = bp.gre_2d(n, ramps=True, spoiling="rf", slice_thickness=0.1, rf_duration=1e-3, n_dummy=2, flip_deg=30.0,
                    system=bp.System(rf_raster=40e-6))
    seq.b0_offset = 2.0 * torch.cos(torch.arange(len(seq), dtype=f64) * 0.03)
    return ph, seq


def test_exact_concomitant_fields_with_18_coils(gpu):
    """True positions enter twice (encoding, and the concomitant coefficients through fields.spin_fields); 18 coils
    take two sweeps of 16."""
    ph, seq = field_case()
    coils = bp.birdcage_sensitivities(ph.pos, 18)
    B0, a = 0.35, 0.45

    def oracle(p, s, c, m):
        return core.simulate(p, s, c, m_init=m, field=OF.Field(p.pos, F.GAMMA_BAR * B0, a, 2))

    keys = ("dt", "rf", "grad", "b0_offset")
    ref = gradients(oracle, ph, seq, coils, None, keys)
    out = gradients(gpu_run(scanner=F.Scanner(field_strength=B0, concomitant="exact", z_asymmetry=a),
                            checkpoint_every=64), ph, seq, coils, None, keys)
    compare(ref, out, 1e-4)


def test_leading_fields_bound_pool_and_nonlinearity(gpu):
    """Leading-order concomitant fields, a Lorentzian pool and a gradient nonlinearity: the coefficients' gradient
    (through the encoding positions) and the true positions' (encoding and concomitant coefficients)."""
    ph = pool_phantom(8, 4, "lorentzian")
    seq = bp.gre_2d(8, ramps=True, spoiling="rf", slice_thickness=5e-3, rf_duration=1e-3, n_dummy=2, flip_deg=30.0,
                    system=bp.System(rf_raster=40e-6))
    coils = bp.birdcage_sensitivities(ph.pos, 2)
    gen = np.random.default_rng(5)
    alpha = np.zeros((3, 4, 4))
    alpha[:, 3, :] = gen.normal(size=(3, 4)) * 2e-2
    alpha[:, 2, :] = gen.normal(size=(3, 4)) * 1e-2
    beta = torch.zeros(3, 4, 4, dtype=f64)
    B0 = 0.55
    results = {}
    for name in ("oracle", "gpu"):
        a_leaf = torch.tensor(np.tril(alpha), dtype=f64, requires_grad=True)
        nl = F.GradientNonlinearity(a_leaf, beta, 0.25)
        p, s, c, m0, leaves = leaves_of(ph, seq, coils, None, ("rf",))
        if name == "oracle":
            enc = nl.encoding_positions(p.pos)
            sig, m = E.simulate_bound(replace(p, pos=enc), s, c, field=OF.Field(p.pos, F.GAMMA_BAR * B0, 0.5, 1))
        else:
            r = bp.simulate(p, s, c, engine="gpu", scanner=F.Scanner(field_strength=B0, nonlinearity=nl),
                            checkpoint_every=64)
            sig, m = r.signal, torch.cat([r.m, r.m_bound[:, None]], 1)
        keys = list(leaves) + ["alpha"]
        grads = torch.autograd.grad(weighted_loss(sig, m), list(leaves.values()) + [a_leaf], allow_unused=True)
        results[name] = {k: g for k, g in zip(keys, grads) if g is not None}
    compare(results["oracle"], results["gpu"], 1e-4)


def test_checkpoint_interval_does_not_change_the_gradients(gpu):
    ph = pool_phantom(6, 6, "lorentzian")
    seq = saturation_sequence()
    coils = bp.birdcage_sensitivities(ph.pos, 2)
    a = gradients(gpu_run(checkpoint_every=64), ph, seq, coils)
    b = gradients(gpu_run(checkpoint_every=256), ph, seq, coils)
    errors = {k: rel(a[k], b[k]) for k in b if b[k].abs().max() > 0}
    print({k: f"{v:.1e}" for k, v in errors.items()}, flush=True)
    assert max(errors.values()) < 1e-5, errors


def test_full_variant_gradient_refusals(gpu):
    ph = pool_phantom(4, 7, "lorentzian")
    seq = bp.gre_2d(4)
    t1 = ph.t1.clone().requires_grad_()
    for kw in (dict(pulses="reuse"), dict(readouts="nufft")):
        with pytest.raises(NotImplementedError):
            bp.simulate(replace(ph, t1=t1), seq, engine="gpu", **kw)
    with pytest.raises(ValueError):
        bp.simulate(replace(ph, t1=t1), seq, engine="gpu", checkpoint_every=100)
