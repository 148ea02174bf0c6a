"""Sequence building blocks and the block splitter that turns them into segments.

Reference: ``src/blochparty/sequences/builder.py`` — ``System`` (:23-43), ``Trapezoid`` (:61-86), ``RFPulse``
(:110-123), ``ADC`` (:126-142), ``trapezoid`` (:145-212), ``block_pulse`` (:267-283), ``sinc_pulse`` (:286-329),
``block_duration`` (:332-341) and ``SequenceBuilder`` (:366-574). The segment tables these produce are the GPU
engine's wire format, so the splitting rules are kept exactly: a block is cut at every gradient corner, at every
gradient-raster point of a sloped piece, at every RF raster step and at every ADC sample (which ends its segment);
each segment carries the exact time average of the piecewise-linear gradient (its midpoint value).
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch
from torch import Tensor

from .sequence import AXES, T_TOL, Sequence

PULSEQ_GAMMA_BAR = 42.576e6  # [Hz/T] pypulseq's rounded value, only to convert mT/m and T/m/s limits
MAX_RF_STEP = 20e-6  # [s] hard pulses are cut into segments no longer than this


@dataclass(frozen=True)
class System:
    """Hardware limits and rasters (3 T defaults: 40 mT/m, 170 T/m/s)."""

    max_grad: float = 40e-3 * PULSEQ_GAMMA_BAR  # [Hz/m]
    max_slew: float = 170.0 * PULSEQ_GAMMA_BAR  # [Hz/m/s]
    grad_raster: float = 10e-6
    rf_raster: float = 1e-6
    adc_raster: float = 100e-9
    rf_dead_time: float = 100e-6
    rf_ringdown_time: float = 30e-6

    def ramp_time(self, amplitude) -> float:
        """Shortest ramp on the gradient raster from 0 to ``amplitude`` [Hz/m]."""
        return ceil_to(abs(float(amplitude)) / self.max_slew, self.grad_raster)

    def check_raster(self, t: float, raster: float, what: str) -> None:
        ratio = t / raster
        if abs(ratio - round(ratio)) > 1e-6:
            raise ValueError(f"{what} = {t:.9g} s is not a multiple of the raster {raster:.9g} s")


def ceil_to(t: float, raster: float) -> float:
    """``t`` rounded up to the raster (with a 1e-9 raster-unit allowance for float noise)."""
    return math.ceil(t / raster - 1e-9) * raster


def axis_vector(value, axis: str | None, dtype) -> Tensor:
    """A (3,) tensor: ``value`` on ``axis``, or ``value`` itself when it is a 3-vector (axis None)."""
    if axis is None:
        if isinstance(value, Tensor):
            return value.to(dtype)
        return torch.stack([torch.as_tensor(v, dtype=dtype) for v in value])
    e = torch.zeros(3, dtype=dtype)
    e[AXES[axis]] = 1.0
    return torch.as_tensor(value, dtype=dtype) * e


@dataclass
class Trapezoid:
    """Trapezoid of (3,) ``amplitude`` [Hz/m]: ``rise``, ``flat``, ``fall`` [s] after ``delay``."""

    amplitude: Tensor
    rise: float
    flat: float
    fall: float
    delay: float = 0.0

    @property
    def duration(self) -> float:
        return self.delay + self.rise + self.flat + self.fall

    @property
    def area(self) -> Tensor:
        return self.amplitude * (self.flat + 0.5 * (self.rise + self.fall))

    def corners(self) -> tuple[np.ndarray, Tensor]:
        times = self.delay + np.cumsum([0.0, self.rise, self.flat, self.fall])
        level = torch.tensor([0.0, 1.0, 1.0, 0.0], dtype=self.amplitude.dtype)
        return times, level[:, None] * self.amplitude[None, :]


@dataclass
class RFPulse:
    """``signal`` (k, 2) [Hz] held for ``raster`` [s] each, starting ``delay`` into the block; ``center`` is the
    time from the pulse start to its effective centre."""

    signal: Tensor
    raster: float
    delay: float
    center: float

    @property
    def duration(self) -> float:
        return self.signal.shape[0] * self.raster


@dataclass
class ADC:
    """``n`` samples at ``delay + (i + 1/2) dwell`` into the block; receiver ``phase`` [rad]."""

    n: int
    dwell: float
    delay: float = 0.0
    phase: float | Tensor = 0.0

    @property
    def times(self) -> np.ndarray:
        return self.delay + (np.arange(self.n) + 0.5) * self.dwell

    @property
    def duration(self) -> float:
        return self.delay + self.n * self.dwell


def check_limits(g, sysm: System) -> None:
    """Raise when a gradient event exceeds the amplitude or slew limit on any axis, or jumps."""
    t, v = g.corners()
    v = v.detach()
    peak = float(v.abs().max()) if v.numel() else 0.0
    if peak > sysm.max_grad * (1 + 1e-9):
        raise ValueError(f"gradient {peak:.6g} Hz/m exceeds max_grad {sysm.max_grad:.6g} Hz/m")
    steps = torch.as_tensor(np.diff(t), dtype=v.dtype)
    change = (v[1:] - v[:-1]).abs().max(dim=1).values
    if bool(((steps <= T_TOL) & (change > 0)).any()):
        raise ValueError("gradient jumps instantaneously (infinite slew rate)")
    slew = float((change / steps.clamp_min(T_TOL)).max()) if steps.numel() else 0.0
    if slew > sysm.max_slew * (1 + 1e-9):
        raise ValueError(f"slew rate {slew:.6g} Hz/m/s exceeds max_slew {sysm.max_slew:.6g} Hz/m/s")


def trapezoid(
    axis: str | None = None,
    *,
    area=None,
    amplitude=None,
    flat_area=None,
    flat_time=None,
    duration=None,
    rise_time=None,
    delay: float = 0.0,
    system: System | None = None,
    dtype=torch.float64,
) -> Trapezoid:
    """A trapezoid within ``system``'s limits; the modes of ``builder.py:145-212``: ``area`` alone (shortest),
    ``area`` + ``duration`` (lowest amplitude), ``flat_area`` + ``flat_time``, ``amplitude`` + ``flat_time`` or
    ``duration``. Timings are fixed from values; areas / amplitudes may carry gradients."""
    sysm = system or System()
    vec = lambda x: axis_vector(x, axis, dtype)
    if amplitude is not None:
        if flat_time is None and duration is None:
            raise ValueError("amplitude needs flat_time or duration")
        amp = vec(amplitude)
        ramp = rise_time if rise_time is not None else sysm.ramp_time(amp.abs().max())
        flat = flat_time if flat_time is not None else duration - 2 * ramp
    elif flat_area is not None:
        if flat_time is None:
            raise ValueError("flat_area needs flat_time")
        amp = vec(flat_area) / flat_time
        ramp = rise_time if rise_time is not None else sysm.ramp_time(amp.abs().max())
        flat = flat_time
    elif area is not None:
        a_vec = vec(area)
        a = float(a_vec.abs().max())
        if duration is not None:
            if rise_time is not None:
                ramp = rise_time
            else:
                disc = duration**2 - 4 * a / sysm.max_slew
                if disc < 0:
                    raise ValueError(f"area {a:.6g} /m does not fit in {duration * 1e3:.4g} ms at the slew limit")
                ramp = ceil_to(0.5 * (duration - math.sqrt(disc)), sysm.grad_raster) if a > 0 else 0.0
            flat = duration - 2 * ramp
        elif a <= sysm.max_grad**2 / sysm.max_slew:  # a triangle reaches the area below max_grad
            ramp = rise_time if rise_time is not None else ceil_to(math.sqrt(a / sysm.max_slew), sysm.grad_raster)
            flat = 0.0
        else:
            ramp = rise_time if rise_time is not None else sysm.ramp_time(sysm.max_grad)
            flat = ceil_to(a / sysm.max_grad - ramp, sysm.grad_raster)
        if flat < -T_TOL:
            raise ValueError(f"ramps of {ramp * 1e6:.4g} us do not fit in {duration * 1e3:.4g} ms")
        flat = max(flat, 0.0)
        amp = a_vec / (flat + ramp) if flat + ramp > 0 else a_vec * 0.0
    else:
        raise ValueError("give area, flat_area or amplitude")
    if flat < -T_TOL:
        raise ValueError(f"negative flat time {flat:.3g} s")
    g = Trapezoid(amp, float(ramp), max(float(flat), 0.0), float(ramp), float(delay))
    check_limits(g, sysm)
    return g


def rf_steps(duration) -> int:
    """Segments of a hard pulse: at most ``MAX_RF_STEP`` each, from the duration alone."""
    return max(1, math.ceil(float(torch.as_tensor(duration).detach()) / MAX_RF_STEP - 1e-6))


def block_pulse(flip_deg, duration: float, *, phase=0.0, system: System | None = None, dtype=torch.float64,
                n_steps: int | None = None) -> RFPulse:
    """Hard pulse after the RF dead time (``builder.py:267-283``)."""
    sysm = system or System()
    n = rf_steps(duration) if n_steps is None else n_steps
    amp = torch.as_tensor(flip_deg, dtype=dtype) / 360.0 / duration
    ph = torch.as_tensor(phase, dtype=dtype)
    signal = torch.stack([amp * torch.cos(ph), amp * torch.sin(ph)])[None].expand(n, 2)
    return RFPulse(signal, float(duration) / n, sysm.rf_dead_time, 0.5 * duration)


def sinc_pulse(flip_deg, duration: float, *, slice_thickness: float, time_bw_product: float = 4.0,
               apodization: float = 0.5, slice_offset: float = 0.0, phase=0.0, system: System | None = None,
               dtype=torch.float64) -> tuple[RFPulse, Trapezoid, Tensor]:
    """Windowed-sinc slice-selective pulse (``builder.py:286-329``): envelope ((1-a) + a cos(2 pi t/T)) sinc(BW t)
    on the RF raster, scaled to ``flip_deg`` at the slice centre; slice-select gradient BW / thickness on z with the
    pulse centred on its flat top; returns (rf, gradient, rephase area from the centre to the ramp end)."""
    sysm = system or System()
    sysm.check_raster(duration, sysm.rf_raster, "RF duration")
    n = round(duration / sysm.rf_raster)
    t = (np.arange(n) + 0.5) * sysm.rf_raster - 0.5 * duration
    bw = time_bw_product / duration
    env = ((1 - apodization) + apodization * np.cos(2 * np.pi * t / duration)) * np.sinc(bw * t)
    env = torch.as_tensor(env / (env.sum() * sysm.rf_raster), dtype=dtype)  # unit area = one cycle
    amp = torch.as_tensor(flip_deg, dtype=dtype) / 360.0 * env
    g_amp = bw / slice_thickness
    ph = torch.as_tensor(phase, dtype=dtype) - torch.as_tensor(2 * np.pi * g_amp * slice_offset * t, dtype=dtype)
    signal = torch.stack([amp * torch.cos(ph), amp * torch.sin(ph)], dim=-1)
    ramp = sysm.ramp_time(g_amp)
    flat = ceil_to(duration, sysm.grad_raster)
    margin = 0.5 * (flat - duration)
    g_delay = max(0.0, sysm.rf_dead_time - ramp - margin)
    gz = trapezoid("z", amplitude=g_amp, flat_time=flat, rise_time=ramp, delay=g_delay, system=sysm, dtype=dtype)
    rf = RFPulse(signal, sysm.rf_raster, g_delay + ramp + margin, 0.5 * duration)
    rephase = -gz.amplitude * (0.5 * flat + 0.5 * ramp)
    return rf, gz, rephase


def block_duration(*grads, rf: RFPulse | None = None, adc: ADC | list[ADC] | None = None,
                   system: System | None = None) -> float:
    """Until the last gradient or ADC sample ends, or the RF pulse plus the ringdown time."""
    adcs = [] if adc is None else ([adc] if isinstance(adc, ADC) else list(adc))
    ends = [0.0] + [g.duration for g in grads] + [a.duration for a in adcs]
    if rf is not None:
        ends.append(rf.delay + rf.duration + (system or System()).rf_ringdown_time)
    return max(ends)


def _piecewise_linear_at(mid: np.ndarray, t: np.ndarray, v: Tensor) -> Tensor:
    """The piecewise-linear waveform through (t, v) at points ``mid`` that are no corners; zero outside."""
    j = np.searchsorted(t, mid, side="right") - 1
    inside = (j >= 0) & (j < len(t) - 1)
    j = np.clip(j, 0, len(t) - 2)
    frac = (mid - t[j]) / np.where(inside, t[j + 1] - t[j], 1.0)
    w = torch.as_tensor(np.where(inside, frac, 0.0), dtype=v.dtype, device=v.device)[:, None]
    on = torch.as_tensor(inside, dtype=v.dtype, device=v.device)[:, None]
    jt = torch.as_tensor(j, device=v.device)
    return (v[jt] * (1 - w) + v[jt + 1] * w) * on


def _raster_cuts(t: np.ndarray, sloped: np.ndarray, step: float) -> np.ndarray:
    """Points cutting every sloped interval of ``t`` into equal pieces no longer than ``step``."""
    pts = [np.zeros(0)]
    for a, b in zip(t[:-1][sloped], t[1:][sloped]):
        k = math.ceil((b - a) / step - 1e-9)
        if k > 1:
            pts.append(a + (b - a) * np.arange(1, k) / k)
    return np.concatenate(pts)


class SequenceBuilder:
    """Appends segments block by block (``builder.py:366-574``); numeric arguments may require grad."""

    def __init__(self, dtype=torch.float64, device=None, system: System | None = None):
        self.dtype, self.device = dtype, device
        self.system = system or System()
        self.time = 0.0
        self._dt, self._rf, self._grad, self._phase = [], [], [], []
        self._is_rf, self._adc, self._spoil = [], [], []

    def _t(self, x) -> Tensor:
        return torch.as_tensor(x, dtype=self.dtype, device=self.device)

    def _grad3(self, grad, n: int) -> Tensor:
        if grad is None:
            g = torch.zeros(3, dtype=self.dtype, device=self.device)
        elif isinstance(grad, Tensor):
            g = grad.to(dtype=self.dtype, device=self.device)
        else:
            g = torch.stack([self._t(x) for x in grad])
        return g if g.ndim == 2 else g.expand(n, 3)

    def _push(self, dt: Tensor, rf: Tensor, grad: Tensor, phase: Tensor, is_rf, adc) -> None:
        n = dt.shape[0]
        self._dt.append(dt)
        self._rf.append(rf)
        self._grad.append(grad)
        self._phase.append(phase)
        self._is_rf += is_rf if isinstance(is_rf, list) else [is_rf] * n
        self._adc += adc if isinstance(adc, list) else [adc] * n
        self._spoil += [False] * n
        self.time += float(dt.detach().sum())

    def rf(self, duration, *, flip_deg=None, amplitude=None, phase=0.0, grad=None, n_steps=None):
        """Hard pulse of ``flip_deg`` (or ``amplitude`` [Hz]) in ``n_steps`` segments."""
        dur = self._t(duration)
        amp = self._t(flip_deg) / 360.0 / dur if flip_deg is not None else self._t(amplitude)
        n = rf_steps(dur) if n_steps is None else n_steps
        ph = self._t(phase)
        rf = torch.stack([amp * torch.cos(ph), amp * torch.sin(ph)]).expand(n, 2)
        dt = (dur / n).expand(n)
        self._push(dt, rf, self._grad3(grad, n), torch.zeros_like(dt), True, False)
        return self

    def rf_waveform(self, samples, dt, grad=None):
        """Arbitrary RF: (k, 2) samples [Hz], one per step of ``dt``."""
        rf = self._t(samples)
        n = rf.shape[0]
        dts = self._t(dt).expand(n)
        self._push(dts, rf, self._grad3(grad, n), torch.zeros_like(dts), True, False)
        return self

    def delay(self, duration, grad=None):
        dt = self._t(duration).reshape(1)
        zeros = torch.zeros(1, 2, dtype=self.dtype, device=self.device)
        self._push(dt, zeros, self._grad3(grad, 1), torch.zeros_like(dt), False, False)
        return self

    def readout(self, n: int, dwell, grad=None, phase=0.0):
        """``n`` samples at (i + 1/2) dwell: segments of dwell/2, dwell x (n - 1), dwell/2, the first n sampled."""
        dw = self._t(dwell)
        dt = torch.cat([(0.5 * dw).reshape(1), dw.expand(n - 1), (0.5 * dw).reshape(1)])
        rf = torch.zeros(n + 1, 2, dtype=self.dtype, device=self.device)
        ph = self._t(phase).expand(n + 1)
        self._push(dt, rf, self._grad3(grad, n + 1), ph, False, [True] * n + [False])
        return self

    def block(self, *grads, rf: RFPulse | None = None, adc: ADC | list[ADC] | None = None,
              duration: float | None = None):
        """Events played together, split into segments at every event boundary (module docstring)."""
        sysm = self.system
        adcs = [] if adc is None else ([adc] if isinstance(adc, ADC) else list(adc))
        if rf is not None and rf.delay < sysm.rf_dead_time - T_TOL:
            raise ValueError(f"RF starts {rf.delay * 1e6:.4g} us into the block, before the dead time")
        end = block_duration(*grads, rf=rf, adc=adcs, system=sysm)
        if duration is not None:
            if duration < end - T_TOL:
                raise ValueError(f"block events last {end:.9g} s, longer than duration {duration:.9g} s")
            end = float(duration)
        if end <= T_TOL:
            return self
        cuts, waves = [np.array([0.0, end])], []
        for g in grads:
            tc, vc = g.corners()
            waves.append((tc, vc.to(dtype=self.dtype, device=self.device)))
            vd = vc.detach()
            sloped = (vd[1:] != vd[:-1]).any(dim=1).numpy() & (np.diff(tc) > T_TOL)
            cuts += [tc, _raster_cuts(tc, sloped, sysm.grad_raster)]
        if rf is not None:
            cuts.append(rf.delay + rf.raster * np.arange(rf.signal.shape[0] + 1))
        for a in adcs:
            cuts.append(a.times)
        t = np.sort(np.concatenate(cuts))
        t = t[np.concatenate([[True], np.diff(t) > T_TOL])]
        if t[-1] > end + T_TOL or t[0] < -T_TOL:
            raise ValueError(f"events extend outside the block [0, {end:.9g}] s")
        t = np.append(t[t < end - T_TOL], end)
        mid, n = 0.5 * (t[1:] + t[:-1]), len(t) - 1
        g = torch.zeros(n, 3, dtype=self.dtype, device=self.device)
        for tc, vc in waves:
            g = g + _piecewise_linear_at(mid, tc, vc)
        rf_seg = torch.zeros(n, 2, dtype=self.dtype, device=self.device)
        on_rf = np.zeros(n, dtype=bool)
        if rf is not None:
            on_rf = (mid > rf.delay) & (mid < rf.delay + rf.duration)
            k = np.clip(np.floor((mid - rf.delay) / rf.raster).astype(int), 0, rf.signal.shape[0] - 1)
            gate = torch.as_tensor(on_rf, dtype=self.dtype, device=self.device)[:, None]
            rf_seg = rf.signal.to(dtype=self.dtype, device=self.device)[torch.as_tensor(k, device=self.device)] * gate
        on_adc = np.zeros(n, dtype=bool)
        phase = torch.zeros(n, dtype=self.dtype, device=self.device)
        for a in adcs:
            ta = a.times
            seg = np.searchsorted(t, ta - T_TOL) - 1  # every sample ends a segment
            if np.any(seg < 0) or np.any(np.abs(t[seg + 1] - ta) > T_TOL) or np.any(on_adc[seg]):
                raise ValueError("ADC samples must lie inside the block, after its start, without overlap")
            on_adc[seg] = True
            ph = torch.as_tensor(a.phase, dtype=self.dtype, device=self.device).expand(a.n)
            phase = phase.index_put((torch.as_tensor(seg, device=self.device),), ph)
        self._push(self._t(np.diff(t)), rf_seg, g, phase, on_rf.tolist(), on_adc.tolist())
        return self

    def trapezoid(self, axis: str | None = None, **kwargs):
        return self.block(trapezoid(axis, system=self.system, dtype=self.dtype, **kwargs))

    def sinc_pulse(self, flip_deg, duration: float, *, slice_thickness: float, rephase: bool = True, **kwargs):
        rf, gz, reph = sinc_pulse(flip_deg, duration, slice_thickness=slice_thickness, system=self.system,
                                  dtype=self.dtype, **kwargs)
        self.block(gz, rf=rf)
        return self.block(trapezoid(area=reph, system=self.system, dtype=self.dtype)) if rephase else self

    def spoil(self):
        """Ideal spoiler at the end of the last segment."""
        self._spoil[-1] = True
        return self

    def build(self, **meta) -> Sequence:
        meta = {"system": self.system, **meta}
        return Sequence(
            dt=torch.cat(self._dt),
            rf=torch.cat(self._rf),
            grad=torch.cat(self._grad),
            adc_phase=torch.cat(self._phase),
            is_rf=list(self._is_rf),
            adc=list(self._adc),
            spoil=list(self._spoil),
            meta=meta,
        )
