"""Conversion of the reference package's input objects into this package's, sharing every tensor.

blochparty's ``Phantom`` (``src/blochparty/phantom.py:17-31``) and ``Sequence`` (``src/blochparty/sequence.py:28-51``)
are dataclasses whose field names this package's ``Phantom`` and ``Sequence`` repeat. ``from_reference`` builds this
package's object from the other's fields by attribute access: no copy, no ``dataclasses.asdict`` (which deep-copies
every tensor and fails on a non-leaf tensor that requires grad), so a T1 map computed from a parameter that requires
grad stays in the autograd graph and its gradient reaches that parameter.
"""

from __future__ import annotations

import dataclasses

from .phantom import Phantom
from .sequence import Sequence

_TARGETS = (Phantom, Sequence)


def _target_of(obj) -> type:
    names = {f.name for f in dataclasses.fields(obj)}
    for cls in _TARGETS:
        required = {f.name for f in dataclasses.fields(cls)
                    if f.default is dataclasses.MISSING and f.default_factory is dataclasses.MISSING}
        if required <= names:
            return cls
    raise TypeError(f"{type(obj).__name__} has neither Phantom's nor Sequence's fields")


def from_reference(obj):
    """This package's ``Phantom`` / ``Sequence`` with the same field values as ``obj`` (the reference's object, or
    any dataclass with those field names). Tensors are the same objects, not copies. A field this package's class
    does not have raises ``TypeError`` unless its value is None. Objects of this package are returned as they are.
    """
    if isinstance(obj, _TARGETS):
        return obj
    if not dataclasses.is_dataclass(obj) or isinstance(obj, type):
        raise TypeError(f"from_reference takes a Phantom or Sequence dataclass, not {type(obj).__name__}")
    cls = _target_of(obj)
    ours = {f.name for f in dataclasses.fields(cls)}
    values = {f.name: getattr(obj, f.name) for f in dataclasses.fields(obj)}
    extra = sorted(k for k, v in values.items() if k not in ours and v is not None)
    if extra:
        raise TypeError(f"{cls.__name__} has no field(s) {', '.join(extra)}")
    return cls(**{k: v for k, v in values.items() if k in ours})


def scanner_from_reference(obj):
    """This package's ``fields.Scanner`` from the reference's (``src/blochparty/fields.py:108-135``) or any object
    with its attributes (field_strength, concomitant, z_asymmetry, nonlinearity with alpha, beta, r0); tensors shared.
    None stays None."""
    from .fields import GradientNonlinearity, Scanner

    if obj is None or isinstance(obj, Scanner):
        return obj
    nl = getattr(obj, "nonlinearity", None)
    if nl is not None and not isinstance(nl, GradientNonlinearity):
        nl = GradientNonlinearity(nl.alpha, nl.beta, nl.r0)
    return Scanner(field_strength=getattr(obj, "field_strength", None), concomitant=getattr(obj, "concomitant", None),
                   z_asymmetry=getattr(obj, "z_asymmetry", 0.5), nonlinearity=nl)


__all__ = ["from_reference", "scanner_from_reference"]
