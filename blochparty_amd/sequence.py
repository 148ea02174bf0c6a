"""The sequence as the engine sees it: a flat list of piecewise-constant segments.

Reference: ``src/blochparty/sequence.py:28-59`` (fields and their meaning), ``:92-142`` (``merge_free``). Continuous
quantities are tensors (they may carry gradients); the structural flags are Python booleans that steer control flow.
Units: RF in Hz (gamma_bar B1), gradients in Hz/m, so a gradient's time integral is a k-space position in cycles/m.
"""

from __future__ import annotations

import contextlib
import contextvars
from dataclasses import dataclass, field, replace

import numpy as np
import torch
from torch import Tensor

T_TOL = 1e-11  # [s] two event times closer than this are one breakpoint
AXES = {"x": 0, "y": 1, "z": 2}


@dataclass
class Sequence:
    dt: Tensor  # (S,) durations [s]
    rf: Tensor  # (S, 2) RF field (x, y) [Hz]
    grad: Tensor  # (S, 3) gradient [Hz/m]
    adc_phase: Tensor  # (S,) receiver phase [rad] where adc is set
    is_rf: list[bool]
    adc: list[bool]  # a sample at the end of the segment
    spoil: list[bool]  # ideal crusher: transverse magnetization zeroed at the end of the segment
    meta: dict = field(default_factory=dict)
    b0_offset: Tensor | None = None  # (S,) uniform field added to every spin's off-resonance [Hz]; None = zeros
    rf_freq: Tensor | None = None  # (S,) RF frequency [Hz] of the bound pool's saturation; None: exchange.rf_frequency
    bound_only: list[bool] | None = None  # RF felt by the bound pool only (the water ignores it)

    def __len__(self) -> int:
        return int(self.dt.shape[0])

    @property
    def n_adc(self) -> int:
        return int(sum(self.adc))

    @property
    def duration(self) -> Tensor:
        return self.dt.sum()

    def to(self, dtype=None, device=None) -> Sequence:
        conv = lambda t: None if t is None else t.to(dtype=dtype, device=device)
        return replace(
            self,
            dt=conv(self.dt),
            rf=conv(self.rf),
            grad=conv(self.grad),
            adc_phase=conv(self.adc_phase),
            b0_offset=conv(self.b0_offset),
            rf_freq=conv(self.rf_freq),
        )

    def merge_free(self) -> Sequence:
        """Every stretch of segments with neither RF nor ADC (a spoil may only end it) as one segment carrying the
        stretch's duration, gradient moment and b0_offset integral; RF and ADC segments are kept as they are. For
        static spins the state after such a stretch depends on nothing else (``sequence.py:92-142``)."""
        is_rf = np.asarray(self.is_rf, dtype=bool)
        adc = np.asarray(self.adc, dtype=bool)
        spoil = np.asarray(self.spoil, dtype=bool)
        n = len(is_rf)
        if n == 0:
            return self
        quiet = ~is_rf & ~adc
        new_group = np.ones(n, dtype=bool)
        new_group[1:] = ~(quiet[1:] & quiet[:-1] & ~spoil[:-1])
        gid = np.cumsum(new_group) - 1
        first = np.flatnonzero(new_group)
        last = np.append(first[1:] - 1, n - 1)
        if np.all(last == first):
            return self
        dev = self.dt.device
        g = torch.as_tensor(gid, device=dev)
        n_g = len(first)

        def gsum(x: Tensor) -> Tensor:
            return torch.zeros((n_g, *x.shape[1:]), dtype=x.dtype, device=dev).index_add(0, g, x)

        many = torch.as_tensor(last > first, device=dev)
        take = torch.as_tensor(first, device=dev)
        dt_sum = gsum(self.dt)
        dt = torch.where(many, dt_sum, self.dt[take])
        grad = torch.where(many[:, None], gsum(self.grad * self.dt[:, None]) / dt_sum[:, None], self.grad[take])
        b0 = None
        if self.b0_offset is not None:
            b0 = torch.where(many, gsum(self.b0_offset * self.dt) / dt_sum, self.b0_offset[take])
        rf_freq = None
        if self.rf_freq is not None:
            rf_freq = torch.where(many, torch.zeros_like(dt), self.rf_freq[take])
        meta = {k: v for k, v in self.meta.items() if k not in ("blocks", "block_segments", "segment_block")}
        meta["merged_free"] = True
        return Sequence(
            dt,
            self.rf[take],
            grad,
            self.adc_phase[take],
            is_rf[first].tolist(),
            adc[first].tolist(),
            spoil[last].tolist(),
            meta,
            b0,
            rf_freq,
            None if self.bound_only is None else np.asarray(self.bound_only, dtype=bool)[first].tolist(),
        )


# Flag lists (is_rf, adc, spoil, bound_only) as bool arrays. Converting a Python list of a million bools takes tens of
# milliseconds, and one case's preparation reads each list in several places; inside ``flag_memo()`` each list object
# is converted once (the memo holds the list itself, so an id is never reused while it lives). Outside it, every call
# converts, so lists changed in place between calls are always seen.
_FLAG_MEMO: contextvars.ContextVar[dict | None] = contextvars.ContextVar("blochparty_amd_flag_memo", default=None)


def flag_array(flags) -> np.ndarray:
    """``flags`` as a read-only bool array."""
    memo = _FLAG_MEMO.get()
    if memo is not None:
        hit = memo.get(id(flags))
        if hit is not None and hit[0] is flags:
            return hit[1]
    arr = np.asarray(flags, dtype=bool)
    if arr is flags:  # already a bool array: a read-only view, not the caller's array itself
        arr = arr.view()
    arr.flags.writeable = False
    if memo is not None:
        memo[id(flags)] = (flags, arr)
    return arr


@contextlib.contextmanager
def flag_memo():
    """Within the block, ``flag_array`` converts each list object once."""
    token = _FLAG_MEMO.set({})
    try:
        yield
    finally:
        _FLAG_MEMO.reset(token)

