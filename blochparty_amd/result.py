"""What a simulation returns (reference: ``src/blochparty/result.py:10-15``)."""

from __future__ import annotations

from dataclasses import dataclass

from torch import Tensor


@dataclass
class SimResult:
    """``signal``: (C, n_adc) complex k-space in ADC order; ``m``: (N, 3) magnetization after the last segment.

    ``m_bound``: (N,) the bound pool's Mz after the last segment (with ``Phantom.bound``); ``m_pools`` keeps the
    reference's field name (free exchanging pools are refused by the capability check). ``surrogate`` (not in the reference): with
    ``pulses="surrogate"``, the judge's report of the run (largest map difference against the exact composition on a
    sample of spins, spins outside the trained box, whether the exact maps replaced the network)."""

    signal: Tensor
    m: Tensor
    m_bound: Tensor | None = None
    m_pools: Tensor | None = None
    surrogate: dict | None = None
