"""``simulate(..., engine="gpu")``: the drop-in entry point (reference: ``src/blochparty/engines/__init__.py:22-144``
and ``engines/gpu/__init__.py:54-96``).

The signature, the option names, the engine resolution and the error types are the reference's. This package holds
one engine, ``"hip"`` (hand-written HIP kernels for gfx950); ``"gpu"`` resolves to it when a ROCm GPU is visible
and raises ``RuntimeError`` otherwise, as the reference does without a GPU. The reference's CPU engines (``"core"``,
``"chunked"``) and its Metal / CUDA backends are not part of this package: naming them raises ``ValueError``.
Results come back float32 / complex64 on the CPU; with inputs that require grad, gradients flow through the
hand-written adjoint (``engines/adjoint.py``).
"""

from __future__ import annotations

import dataclasses

import torch
from torch import Tensor

from ..interop import from_reference, scanner_from_reference
from ..phantom import Phantom
from ..result import SimResult
from ..sequence import Sequence
from . import hip
from .capabilities import check, features

ENGINES = ("gpu", "hip")
# the reference's GPU options (engines/__init__.py:23), then rf_grad: how rf is differentiated through reused pulses
GPU_OPTIONS = ("chunk", "group", "mma", "readouts", "nufft_width", "nufft_terms", "pulses", "rf_grad")
_OTHER_ENGINES = ("core", "chunked", "metal", "cuda")


def available(engine: str = "gpu") -> bool:
    """Whether ``engine`` runs here."""
    if engine in ("gpu", "hip"):
        return hip.available()
    if engine in _OTHER_ENGINES:
        return False
    raise ValueError(f"engine must be one of {ENGINES}, not {engine!r}")


def resolve(engine: str) -> str:
    """``"gpu"`` is ``"hip"`` where a ROCm GPU and the built library are (RuntimeError when not)."""
    if engine in _OTHER_ENGINES:
        raise ValueError(f"engine={engine!r} is not part of blochparty_amd (its engines: {ENGINES})")
    if engine not in ENGINES:
        raise ValueError(f"engine must be one of {ENGINES}, not {engine!r}")
    if not hip.available():
        raise RuntimeError("engine='gpu': no GPU engine here (the HIP engine needs a ROCm GPU and the built "
                           "library, python -m blochparty_amd.build)")
    return "hip"


def requires_grad(*objs) -> bool:
    """Whether any tensor in ``objs`` (tensors, dataclasses such as Phantom and Sequence, tuples) requires grad
    (reference ``engines/gpu/__init__.py:39-51``)."""
    for o in objs:
        if isinstance(o, Tensor):
            if o.requires_grad:
                return True
        elif dataclasses.is_dataclass(o) and not isinstance(o, type):
            if requires_grad(*(getattr(o, f.name) for f in dataclasses.fields(o))):
                return True
        elif isinstance(o, tuple) and requires_grad(*o):
            return True
    return False


def simulate(
    phantom: Phantom,
    seq: Sequence,
    coils: Tensor | None = None,
    *,
    m_init: Tensor | None = None,
    checkpoint_every: int | None = None,
    merge_free_runs: bool = True,
    scanner=None,
    seed: int = 0,
    maps=None,
    engine: str = "gpu",
    **options,
) -> SimResult:
    """Run ``seq`` on ``phantom`` on the GPU.

    coils: (N, C) complex receive sensitivities (default one uniform coil). m_init: (N, 3) starting magnetization
    (default equilibrium). checkpoint_every: the adjoint stores the state every this many segments (a multiple of 64;
    default about sqrt(segments)). merge_free_runs must stay True (the core's option). scanner: the field model
    (``fields.Scanner``, or the reference's): a gradient nonlinearity encodes every spin at f(r) (the coefficients'
    gradients through autograd); concomitant fields take the full-variant kernels (as a bound pool and bound-only RF
    do), with pulses="step" and readouts="direct", forward and with gradients. maps: refused (NotImplementedError) unless None. seed: only
    meaningful for diffusion, which is refused.
    options: ``chunk`` (segments per dispatch, default 32768), ``group`` (256), ``mma`` (the matrix cores always
    contract the coils here; any value is accepted), ``readouts`` ("direct", or "nufft": readouts of >= 16 equal
    samples as a type-1 NUFFT with ``nufft_width`` kernel cells (6) and ``nufft_terms`` decay terms (3); under
    autograd the reverse sweep takes each run as the transpose of its NUFFT), ``pulses`` ("step"; "reuse": a pulse
    that recurs up to a global RF phase is composed once per spin into its affine map, in float64, and applied at
    every occurrence; "surrogate": those maps from
    the learned RF surrogate, trained on the GPU on first use, with the exact composition as the fallback outside its
    trained box and as the judge of its error on every run, reported in ``SimResult.surrogate``; under autograd both
    run as maps in the adjoint too, with gradients for the spin inputs, coils, m_init and each occurrence's RF
    amplitude and phase, and ValueError when dt, grad or b0_offset require a gradient), ``rf_grad`` (None, or
    "scale": with reused pulses under autograd, rf's gradient is each occurrence's complex-amplitude component, exact
    for flip angles and RF phases and not for a pulse's shape; with rf requiring grad, pulses="reuse" / "surrogate"
    raise ValueError unless rf_grad="scale" accepts that).
    Other keywords raise TypeError. ``phantom`` and ``seq`` may be the reference package's objects (or any dataclass
    with their fields): they are taken over field by field, tensors shared (``interop.from_reference``).
    """
    name = resolve(engine)
    phantom, seq = from_reference(phantom), from_reference(seq)
    scanner = scanner_from_reference(scanner)
    if unknown := sorted(set(options) - set(GPU_OPTIONS)):
        raise TypeError(f"simulate(engine={engine!r}) got unexpected keyword argument(s) {', '.join(unknown)}")
    if not merge_free_runs:
        raise ValueError(f"merge_free_runs=False is the core's segment-by-segment stepping, not engine={engine!r}'s")
    chunk = int(options.get("chunk", 32768))
    group = int(options.get("group", 256))
    readouts = options.get("readouts", "direct")
    pulses = options.get("pulses", "step")
    if group != 256:
        raise ValueError("the HIP kernels run workgroups of 256")
    if readouts not in ("direct", "nufft"):
        raise ValueError(f"readouts must be 'direct' or 'nufft', not {readouts!r}")
    nufft = None
    if readouts == "nufft":
        nufft = (int(options.get("nufft_width", 6)), int(options.get("nufft_terms", 3)))
    if pulses not in ("step", "reuse", "surrogate"):
        raise ValueError(f"pulses must be 'step', 'reuse' or 'surrogate', not {pulses!r}")
    if chunk < 1:
        raise ValueError("chunk must be positive")
    rf_grad = options.get("rf_grad")
    if rf_grad not in (None, "scale"):
        raise ValueError(f"rf_grad must be None or 'scale', not {rf_grad!r}")
    check(name, features(phantom, seq, m_init, scanner, maps), grad=False)
    conc = None
    if scanner is not None:  # the spins at their encoding positions; concomitant coefficients at the true ones
        from ..fields import spin_fields

        fields = spin_fields(scanner, phantom.pos)
        phantom = dataclasses.replace(phantom, pos=fields.pos)
        conc = (fields.conc, fields.mode) if fields.mode else None
    grad = torch.is_grad_enabled() and requires_grad(phantom, seq, coils, m_init)
    check(name, features(phantom, seq, m_init, scanner, maps), grad=grad)
    if grad:
        from . import mt

        if mt.needed(phantom, seq, conc):
            if pulses != "step" or nufft is not None:
                raise NotImplementedError("hip engine: a bound pool, concomitant fields or bound-only RF run with "
                                          "pulses='step' and readouts='direct'")
            from .full_adjoint import simulate_grad as simulate_grad_full

            return simulate_grad_full(phantom, seq, coils, m_init, checkpoint_every, conc)
        from .adjoint import simulate_grad

        return simulate_grad(phantom, seq, coils, m_init, checkpoint_every, pulses, nufft, rf_grad)
    from .driver import forward

    with torch.no_grad():
        return forward(phantom, seq, coils, m_init, chunk=chunk, reuse=pulses == "reuse",
                       surrogate=pulses == "surrogate", nufft=nufft, conc=conc)


__all__ = ["ENGINES", "GPU_OPTIONS", "available", "resolve", "simulate"]
