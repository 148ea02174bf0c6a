"""What the GPU engine accepts, forward and with gradients: one table, one check, one wording.

Reference: ``src/blochparty/engines/capabilities.py:12-91``. The feature names and the message are the reference's;
the ``hip`` row is this engine's: the plain Bloch path with B1+, an initial magnetization, a gradient nonlinearity
(encoding positions), a rate-model bound pool, bound-only RF segments and a scanner's concomitant fields, forward and
with gradients. Everything else (the generalized bound pool, free pools, diffusion, motion,
field maps) is refused with ``NotImplementedError``.
"""

from __future__ import annotations

FEATURES = {
    "b1": "B1+ (Phantom.b1)",
    "maps": "field maps (FieldMaps)",
    "bound": "a bound pool (Phantom.bound)",
    "generalized": "the generalized Bloch bound pool",
    "pools": "free exchanging pools (Phantom.pools)",
    "diffusion": "diffusion (Phantom.diffusion)",
    "motion": "motion (Phantom.motion)",
    "scanner": "the scanner's concomitant fields (scanner= with a field strength)",
    "nonlinearity": "the gradient nonlinearity (Scanner.nonlinearity)",
    "bound_only": "bound-pool-only RF segments (Sequence.bound_only)",
    "m_init": "an initial magnetization (m_init)",
}
SUPPORTED = {
    ("hip", "forward"): frozenset({"b1", "m_init", "bound", "bound_only", "scanner", "nonlinearity"}),
    ("hip", "gradients"): frozenset({"b1", "m_init", "bound", "bound_only", "scanner", "nonlinearity"}),
}


def features(phantom, seq, m_init=None, scanner=None, maps=None) -> frozenset:
    """The ``FEATURES`` a case uses."""
    bound_only = getattr(seq, "bound_only", None)
    bound = getattr(phantom, "bound", None)
    present = {
        "b1": getattr(phantom, "b1", None) is not None,
        "maps": maps is not None,
        "bound": bound is not None,
        "generalized": bound is not None and getattr(bound, "model", None) == "generalized",
        "pools": bool(getattr(phantom, "pools", None)),
        "diffusion": getattr(phantom, "diffusion", None) is not None,
        "motion": getattr(phantom, "motion", None) is not None,
        "scanner": scanner is not None and getattr(scanner, "concomitant", None) is not None
        and getattr(scanner, "field_strength", None) is not None,
        "nonlinearity": getattr(scanner, "nonlinearity", None) is not None,
        "bound_only": bound_only is not None and any(bound_only),
        "m_init": m_init is not None,
    }
    return frozenset(k for k, v in present.items() if v)


def unsupported(engine: str, used: frozenset, grad: bool = False) -> list[str]:
    missing = set(used - SUPPORTED[(engine, "forward")])
    if grad:
        missing |= used - SUPPORTED[(engine, "gradients")]
    return sorted(missing)


def check(engine: str, used: frozenset, grad: bool = False) -> None:
    """``NotImplementedError`` when ``engine`` does not run a case with features ``used``."""
    bad = unsupported(engine, used, grad)
    if bad:
        what = "; ".join(FEATURES[k] for k in bad)
        kind = "gradients through" if grad else "simulating with"
        raise NotImplementedError(f"{engine} engine: {kind} {what} is not implemented; use blochparty's core engine")
