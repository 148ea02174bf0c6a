"""The reference's input objects cross the boundary by ``from_reference``: tensors shared, autograd graph kept (CPU).

The stand-in dataclasses below carry the reference's field sets (src/blochparty/phantom.py:17-31,
src/blochparty/sequence.py:28-51) without importing it."""

import dataclasses
from dataclasses import dataclass, field

import pytest
import torch

import blochparty_amd as bp
from blochparty_amd.interop import from_reference

f64 = torch.float64


@dataclass
class RefPhantom:
    pos: torch.Tensor
    pd: torch.Tensor
    t1: torch.Tensor
    t2: torch.Tensor
    db0: torch.Tensor
    b1: torch.Tensor | None = None
    bound: object | None = None
    pools: tuple | None = None
    diffusion: torch.Tensor | None = None
    restriction: object | None = None
    motion: object | None = None


@dataclass
class RefSequence:
    dt: torch.Tensor
    rf: torch.Tensor
    grad: torch.Tensor
    adc_phase: torch.Tensor
    is_rf: list
    adc: list
    spoil: list
    meta: dict = field(default_factory=dict)
    b0_offset: torch.Tensor | None = None
    rf_freq: torch.Tensor | None = None
    bound_only: list | None = None


def ref_phantom(t1_param):
    n = t1_param.shape[0]
    g = torch.Generator().manual_seed(0)
    return RefPhantom(pos=torch.rand(n, 3, generator=g, dtype=f64), pd=torch.ones(n, dtype=f64),
                      t1=2.0 * t1_param, t2=torch.full((n,), 0.08, dtype=f64), db0=torch.zeros(n, dtype=f64))


def test_phantom_tensors_are_shared_and_the_graph_is_kept():
    t1_param = torch.linspace(0.5, 1.0, 5, dtype=f64, requires_grad=True)
    ref = ref_phantom(t1_param)
    assert not ref.t1.is_leaf and ref.t1.requires_grad
    with pytest.raises(RuntimeError):  # what the old recipe, Phantom(**dataclasses.asdict(p)), runs into
        bp.Phantom(**dataclasses.asdict(ref))
    ph = from_reference(ref)
    assert isinstance(ph, bp.Phantom)
    for f in dataclasses.fields(ref):
        assert getattr(ph, f.name) is getattr(ref, f.name), f.name
    (ph.t1**2).sum().backward()  # the gradient reaches the parameter the map was computed from
    assert torch.allclose(t1_param.grad, 8.0 * t1_param.detach())


def test_sequence_fields_and_identity_of_our_objects():
    seq = bp.gre_2d(4, n_dummy=1)
    ref = RefSequence(**{f.name: getattr(seq, f.name) for f in dataclasses.fields(seq)})
    out = from_reference(ref)
    assert isinstance(out, bp.Sequence)
    assert out.dt is seq.dt and out.rf is seq.rf and out.is_rf is seq.is_rf
    assert from_reference(seq) is seq


def test_unknown_fields_are_refused():
    @dataclass
    class Other:
        pos: torch.Tensor
        pd: torch.Tensor
        t1: torch.Tensor
        t2: torch.Tensor
        db0: torch.Tensor
        temperature: float | None = None

    z = torch.zeros(2, dtype=f64)
    assert isinstance(from_reference(Other(z.view(2, 1).expand(2, 3), z, z, z, z)), bp.Phantom)
    with pytest.raises(TypeError, match="temperature"):
        from_reference(Other(z.view(2, 1).expand(2, 3), z, z, z, z, temperature=310.0))
    with pytest.raises(TypeError):
        from_reference({"pos": z})
