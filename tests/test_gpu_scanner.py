"""The scanner field model on the GPU against the FP64 oracle: concomitant fields (leading order and exact, through
the full-variant kernel), with a bound pool too, bound-only RF without a bound pool, and a gradient nonlinearity
(encoding positions) forward and with its coefficients' gradient through the adjoint."""

from dataclasses import replace

import numpy as np
import pytest
import torch

import blochparty_amd as bp
from blochparty_amd import fields as F
from cases import rel
from oracle import core
from oracle import exchange as E
from oracle import fields as OF

pytestmark = pytest.mark.gpu
f64 = torch.float64


def field_case(n=12, seed=0):
    """A slice-selective spoiled GRE over spins spread +-11 cm in plane and +-4 cm in z (large concomitant fields at
    0.35 T), with B1 and off-resonance."""
    ph = bp.shepp_logan(n)
    g = torch.Generator().manual_seed(seed)
    N = len(ph)
    pos = ph.pos.clone()
    pos[:, 2] = (torch.rand(N, generator=g, dtype=f64) - 0.5) * 0.08
    ph = replace(ph, pos=pos, b1=torch.linspace(0.9, 1.1, N, dtype=f64),
                 db0=(torch.rand(N, generator=g, dtype=f64) - 0.5) * 60)
    seq = bp.gre_2d(n, ramps=True, spoiling="rf", slice_thickness=0.1, rf_duration=1e-3, n_dummy=2, flip_deg=30.0,
                    system=bp.System(rf_raster=20e-6))
    return ph, seq, bp.birdcage_sensitivities(ph.pos, 3)


@pytest.mark.parametrize("mode", ["leading", "exact"])
def test_concomitant_fields_against_the_oracle(gpu, mode):
    ph, seq, coils = field_case()
    scanner = F.Scanner(field_strength=0.35, concomitant=mode, z_asymmetry=0.45)
    out = bp.simulate(ph, seq, coils, engine="gpu", scanner=scanner)
    field = OF.Field(ph.pos.double(), F.GAMMA_BAR * 0.35, 0.45, 1 if mode == "leading" else 2)
    ref_s, ref_m = core.simulate(ph, seq, coils, field=field)
    plain_s, _ = core.simulate(ph, seq, coils)
    e_s, e_m = rel(out.signal.to(torch.complex128), ref_s), rel(out.m.double(), ref_m)
    effect = rel(plain_s, ref_s)
    print(f"concomitant ({mode}) vs oracle: signal {e_s:.1e}, m {e_m:.1e}; the field's effect {effect:.2e}", flush=True)
    assert effect > 1e-2  # the field model matters here
    assert torch.isfinite(out.signal).all()
    assert e_s < 1e-5 and e_m < 3e-5, (e_s, e_m)  # m: this per-segment float32 kernel's floor (1e-5 - 2e-5)


def test_concomitant_fields_with_a_bound_pool(gpu):
    ph, seq, coils = field_case(seed=1)
    N = len(ph)
    g = torch.Generator().manual_seed(2)
    pool = bp.BoundPool(f=0.05 + 0.1 * torch.rand(N, generator=g, dtype=f64), k=20.0, t1=1.0, t2=1e-5, dw=-150.0)
    ph = replace(ph, bound=pool)
    out = bp.simulate(ph, seq, coils, engine="gpu", scanner=F.Scanner(field_strength=0.35))
    ref_s, ref_m = E.simulate_bound(ph, seq, coils, field=OF.Field(ph.pos.double(), F.GAMMA_BAR * 0.35, 0.5, 1))
    assert out.m.shape == (N, 3) and out.m_bound.shape == (N,)
    m4 = torch.cat([out.m, out.m_bound[:, None]], 1).double()
    e_s, e_m = rel(out.signal.to(torch.complex128), ref_s), rel(m4, ref_m)
    print(f"concomitant + bound pool vs oracle: signal {e_s:.1e}, m {e_m:.1e}", flush=True)
    assert e_s < 1e-5 and e_m < 1e-5, (e_s, e_m)


def test_bound_only_rf_without_a_bound_pool(gpu):
    """RF that only a bound pool feels does nothing to the water, with or without a bound pool: the same as the
    sequence with that RF switched off."""
    ph, seq, coils = field_case(seed=3)
    only = [bool(r) and i % 7 == 3 for i, r in enumerate(seq.is_rf)]
    seq_b = replace(seq, bound_only=only)
    rf_off = seq.rf.clone()
    rf_off[torch.tensor(only)] = 0.0
    out = bp.simulate(ph, seq_b, coils, engine="gpu")
    ref_s, ref_m = core.simulate(ph, replace(seq, rf=rf_off), coils)
    e_s = rel(out.signal.to(torch.complex128), ref_s)
    print(f"bound-only RF without a pool vs oracle: {e_s:.1e}", flush=True)
    assert rel(core.simulate(ph, seq, coils)[0], ref_s) > 1e-3  # the switched-off RF mattered
    assert e_s < 1e-5


def nonlinearity(seed=0, requires_grad=False):
    gen = np.random.default_rng(seed)
    alpha, beta = np.zeros((3, 4, 4)), np.zeros((3, 4, 4))
    alpha[:, 3, :] = gen.normal(size=(3, 4)) * 2e-2
    alpha[:, 2, :] = gen.normal(size=(3, 4)) * 1e-2
    beta[:, 3, 1:] = gen.normal(size=(3, 3)) * 2e-2
    a = torch.tensor(np.tril(alpha), dtype=f64, requires_grad=requires_grad)
    return F.GradientNonlinearity(a, torch.tensor(np.tril(beta), dtype=f64), 0.25)


def test_gradient_nonlinearity_forward_and_coefficient_gradient(gpu):
    ph, seq, coils = field_case(seed=4)
    nl = nonlinearity()
    out = bp.simulate(ph, seq, coils, engine="gpu", scanner=F.Scanner(nonlinearity=nl))
    enc = OF.encoding(ph.pos.double(), nl.alpha.detach().numpy(), nl.beta.numpy(), 0.25)
    ref_s, _ = core.simulate(ph, seq, coils, enc_pos=enc)
    e_s = rel(out.signal.to(torch.complex128), ref_s)
    effect = rel(core.simulate(ph, seq, coils)[0], ref_s)
    print(f"nonlinearity vs oracle: {e_s:.1e} (its effect {effect:.1e})", flush=True)
    assert effect > 1e-2 and e_s < 1e-5
    # dL/dalpha (L = sum |s|^2) through the adjoint against the oracle's autograd (both through the package's solid
    # harmonics, pinned against SciPy on the CPU)
    grads = {}
    for name in ("gpu", "oracle"):
        nl_g = nonlinearity(requires_grad=True)
        if name == "gpu":
            s = bp.simulate(ph, seq, coils, engine="gpu", scanner=F.Scanner(nonlinearity=nl_g),
                            checkpoint_every=64).signal
        else:
            s = core.simulate(ph, seq, coils, enc_pos=nl_g.encoding_positions(ph.pos.double()))[0]
        (grads[name],) = torch.autograd.grad((s.abs() ** 2).sum(), [nl_g.alpha])
    e_g = rel(grads["gpu"].double(), grads["oracle"])
    print(f"dL/dalpha vs oracle autograd: {e_g:.1e}", flush=True)
    assert e_g < 1e-4


def test_field_model_refusals():
    """Concomitant fields run on the plain stepped path: reused pulses and NUFFT readouts raise NotImplementedError,
    forward and with gradients."""
    ph, seq, coils = field_case(n=4)
    scanner = F.Scanner(field_strength=1.5)
    t1 = ph.t1.clone().requires_grad_()
    with pytest.raises(NotImplementedError):
        bp.simulate(replace(ph, t1=t1), seq, coils, engine="gpu", scanner=scanner, pulses="surrogate")
    with pytest.raises(NotImplementedError):
        bp.simulate(ph, seq, coils, engine="gpu", scanner=scanner, pulses="reuse")
