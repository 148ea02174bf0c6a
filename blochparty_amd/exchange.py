"""Magnetization transfer: a bound (semi-solid) pool exchanging longitudinal magnetization with the free water, the
rate model (Henkelman 1993; Graham's saturation rate), for the GPU engine's ``Phantom.bound``.

Reference interface: ``src/blochparty/exchange.py`` — ``BoundPool`` (:62-101), the lineshapes (:141-160),
``rf_frequency`` (:175-203), ``z_exchange`` (:209-247), ``saturation_rate`` (:250-256). Per spin the free pool keeps
its Bloch dynamics and the bound pool only Mzb (equilibrium f PD):

    dMz/dt  = R1 (PD - Mz)     - k f Mz + k Mzb          (+ the RF rotation of the free pool)
    dMzb/dt = R1b (f PD - Mzb) + k f Mz - k Mzb - W Mzb,

with W = w1^2 G(2 pi Delta) [1/s] while RF plays, w1 = 2 pi b1 |RF| [rad/s] and Delta the RF's frequency minus the
spin's off-resonance minus the pool's shift. Lineshapes G [s]: Lorentzian T2b / (1 + x^2), x = 2 pi Delta T2b; the
super-Lorentzian T2b F(x), F(x) = sqrt(2 pi) int_0^1 exp(-2 x^2 / (3u^2 - 1)^2) / |3u^2 - 1| du, held at its value at
``cutoff`` [Hz] closer to resonance, from a table of ln F on a log grid (``superlorentzian_table``). The
generalized Bloch model and the cut-free lineshape are not part of this package.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, replace

import numpy as np
import torch
from torch import Tensor

SL_CUTOFF = 1e3  # [Hz] the reference's default cut
LINESHAPES = ("superlorentzian", "lorentzian")
SL_X_MIN, SL_X_MAX, SL_POINTS = 1e-4, 30.0, 4097  # the table's range of x = 2 pi Delta T2b and its size


@dataclass
class BoundPool:
    """Bound-pool parameters per spin (tensors (N,), or scalars broadcast to N)."""

    f: Tensor | float  # pool size relative to the free pool, M0b / M0
    k: Tensor | float  # exchange rate from the bound to the free pool [Hz]
    t1: Tensor | float  # [s]
    t2: Tensor | float  # [s], the lineshape's width (~10 us in tissue)
    dw: Tensor | float = 0.0  # resonance shift from water [Hz]
    lineshape: str = "superlorentzian"
    cutoff: float | None = SL_CUTOFF  # [Hz]; None (no cut) is refused by this package
    model: str = "rate"  # "generalized" is refused by this package
    memory: float = 300.0  # (the generalized model's; kept for the interface)

    def __post_init__(self):
        if self.lineshape not in LINESHAPES:
            raise ValueError(f"lineshape must be one of {LINESHAPES}, got {self.lineshape!r}")
        if self.model not in ("rate", "generalized"):
            raise ValueError(f"model must be 'rate' or 'generalized', got {self.model!r}")
        if self.cutoff is not None and not self.cutoff > 0:
            raise ValueError(f"cutoff must be positive or None, got {self.cutoff!r}")

    def expand(self, n: int, like: Tensor) -> "BoundPool":
        conv = lambda v: torch.as_tensor(v, dtype=like.dtype, device=like.device).expand(n)
        return replace(self, f=conv(self.f), k=conv(self.k), t1=conv(self.t1), t2=conv(self.t2), dw=conv(self.dw))

    def to(self, dtype=None, device=None) -> "BoundPool":
        conv = lambda v: v.to(dtype=dtype, device=device) if isinstance(v, Tensor) else v
        return replace(self, f=conv(self.f), k=conv(self.k), t1=conv(self.t1), t2=conv(self.t2), dw=conv(self.dw))

    def __getitem__(self, idx) -> "BoundPool":
        pick = lambda v: v[idx] if isinstance(v, Tensor) and v.dim() > 0 else v
        return replace(self, f=pick(self.f), k=pick(self.k), t1=pick(self.t1), t2=pick(self.t2), dw=pick(self.dw))


def superlorentzian_integral(x: np.ndarray, nodes: int = 256) -> np.ndarray:
    """F(x) for x > 0. With u0 = 1/sqrt(3) (so 3 u0^2 = 1), u = u0 sin(phi) below the magic angle makes
    3u^2 - 1 = -cos^2(phi) and u = u0 cosh(psi) above it makes it sinh^2(psi), so

        F = sqrt(2 pi) u0 [ int_0^{pi/2} exp(-2 x^2 / cos^4 phi) / cos phi dphi
                          + int_0^{acosh sqrt 3} exp(-2 x^2 / sinh^4 psi) / sinh psi dpsi ],

    both integrands smooth and vanishing (faster than any power) at the magic angle, where Gauss-Legendre nodes
    cluster; ``nodes`` per integral."""
    x = np.atleast_1d(np.asarray(x, dtype=np.float64))[:, None]
    t, w = np.polynomial.legendre.leggauss(nodes)
    u0 = 1.0 / math.sqrt(3.0)
    phi = 0.25 * math.pi * (t + 1)
    c = np.cos(phi)
    below = 0.25 * math.pi * np.sum(w * np.exp(-2 * x * x / c**4) / c, axis=1)
    top = math.acosh(math.sqrt(3.0))
    psi = 0.5 * top * (t + 1)
    sh = np.sinh(psi)
    above = 0.5 * top * np.sum(w * np.exp(-2 * x * x / sh**4) / sh, axis=1)
    return math.sqrt(2 * math.pi) * u0 * (below + above)


_TABLE: dict = {}


def superlorentzian_table() -> tuple[float, float, np.ndarray]:
    """(ln x_min, step in ln x, ln F at SL_POINTS points from x_min to x_max), float64, computed once."""
    if "sl" not in _TABLE:
        lx = np.linspace(math.log(SL_X_MIN), math.log(SL_X_MAX), SL_POINTS)
        _TABLE["sl"] = (float(lx[0]), float(lx[1] - lx[0]),
                        np.log(np.maximum(superlorentzian_integral(np.exp(lx)), 1e-300)))
    return _TABLE["sl"]


def lineshape(pool: BoundPool, delta: Tensor) -> Tensor:
    """G(2 pi delta) [s] for offsets ``delta`` [Hz] (the pool's T2b broadcast)."""
    t2 = torch.as_tensor(pool.t2, dtype=delta.dtype, device=delta.device)
    if pool.lineshape == "lorentzian":
        x = 2 * math.pi * delta * t2
        return t2 / (1 + x * x)
    x = 2 * math.pi * delta.abs().clamp_min(pool.cutoff) * t2
    lx0, step, logf = superlorentzian_table()
    table = torch.as_tensor(logf, dtype=delta.dtype, device=delta.device)
    pos = ((torch.log(x.clamp_min(SL_X_MIN)) - lx0) / step).clamp(0, SL_POINTS - 1)
    i = pos.detach().floor().long().clamp(max=SL_POINTS - 2)
    a = pos - i
    val = torch.exp(table[i] * (1 - a) + table[i + 1] * a)
    return t2 * torch.where(x > SL_X_MAX, torch.zeros_like(val), val)


def rf_frequency(seq) -> Tensor:
    """(S,) each segment's RF frequency [Hz] in this package's convention (a pulse turning as exp(-i 2 pi f t) is
    resonant with spins at offset f): minus the phase step per unit time to the RF neighbours, the mean of the two
    steps where both neighbours carry RF, 0 with none. ``seq.rf_freq`` when it is set."""
    if seq.rf_freq is not None:
        if seq.rf_freq.shape != seq.dt.shape:
            raise ValueError(f"Sequence.rf_freq has shape {tuple(seq.rf_freq.shape)}, expected {tuple(seq.dt.shape)}")
        return seq.rf_freq
    rf, dt = seq.rf, seq.dt
    live = torch.tensor(seq.is_rf, device=dt.device) & ((rf[:, 0] ** 2 + rf[:, 1] ** 2) > 0)
    both = live[1:] & live[:-1]
    z0 = torch.complex(rf[:-1, 0], rf[:-1, 1])
    z1 = torch.complex(rf[1:, 0], rf[1:, 1])
    prod = torch.where(both, z1 * z0.conj(), torch.ones_like(z1))
    step = torch.where(both, -torch.angle(prod) / (math.pi * (dt[1:] + dt[:-1])), torch.zeros_like(dt[1:]))
    pad = dt.new_zeros(1)
    n = torch.cat([pad, both.to(dt.dtype)]) + torch.cat([both.to(dt.dtype), pad])
    return (torch.cat([pad, step]) + torch.cat([step, pad])) / n.clamp_min(1)


__all__ = ["BoundPool", "LINESHAPES", "SL_CUTOFF", "lineshape", "rf_frequency", "superlorentzian_integral",
           "superlorentzian_table"]
