"""Sequence building blocks, and a batched compiler that turns blocks into the segment table.

Reference interface: ``src/blochparty/sequences/builder.py`` — ``System`` (:23-43), ``Trapezoid`` (:61-86),
``RFPulse`` (:110-123), ``ADC`` (:126-142), ``trapezoid`` (:145-212), ``block_pulse`` (:267-283), ``sinc_pulse``
(:286-329), ``block_duration`` (:332-341), ``SequenceBuilder`` (:366-574). Names, fields, argument meanings and error
types follow that interface; the segment table it produces is the GPU engine's input format, pinned against the
reference's tables (tests/golden/sequences.npz).

Segment rules (what a block becomes):
- a block's knots are its start and end, every gradient lobe's corners, the gradient-raster points that cut each
  sloped piece of a lobe into equal parts, every RF raster edge and every ADC sample time; a knot within ``T_TOL`` of
  the previous one is the same knot, and knots at or past the block's end (less ``T_TOL``) give way to the end;
- one segment spans two consecutive knots; its gradient is the sum over the block's lobes of the lobe's level at the
  segment's midpoint (0 before and after, linear on the ramps, 1 on the flat top) times its (3,) amplitude, which is
  the segment's exact time average since no corner lies inside a segment;
- its RF is the raster sample covering the midpoint (zero outside the pulse); an ADC sample ends the segment whose
  end knot it is and gives that segment the ADC's receiver phase.

Design: builder calls do not split anything. ``block`` and the other appenders record rows in small
struct-of-arrays tables (blocks, lobes, RF pulses, ADC events; or ready segments for the direct appenders), and
``build`` splits every recorded block in one vectorised pass: one sort of all knots, one gather of RF samples and
one ``index_add`` of lobe contributions. Amplitudes and RF samples stay torch tensors through that gather and sum, so
flip angles and areas that require grad keep their graph; timings are Python floats. The sequence library
(``library.py``) records all TRs of a sequence as one batch, so a 64^3 GRE (1.1 M segments) builds in well under a
second.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import torch
from torch import Tensor

from .sequence import AXES, T_TOL, Sequence

PULSEQ_GAMMA_BAR = 42.576e6  # [Hz/T] pypulseq's rounded value, used only to express mT/m and T/m/s limits in Hz
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
        steps = t / raster
        if abs(steps - round(steps)) > 1e-6:
            raise ValueError(f"{what} = {t:.9g} s is not a multiple of the raster {raster:.9g} s")


def ceil_to(t, raster: float):
    """``t`` rounded up to a multiple of ``raster``, with 1e-9 of a raster step forgiven for float noise (scalars
    and arrays alike)."""
    if isinstance(t, np.ndarray):
        return np.ceil(t / raster - 1e-9) * raster
    return math.ceil(t / raster - 1e-9) * raster


def axis_vector(value, axis: str | None, dtype) -> Tensor:
    """A (3,) tensor: ``value`` along ``axis`` ("x", "y", "z"), or ``value`` itself as a 3-vector (axis None)."""
    if axis is not None:
        out = torch.zeros(3, dtype=dtype)
        out[AXES[axis]] = 1.0
        return out * torch.as_tensor(value, dtype=dtype)
    if isinstance(value, Tensor):
        return value.to(dtype)
    return torch.stack([torch.as_tensor(v, dtype=dtype) for v in value])


@dataclass
class Trapezoid:
    """A gradient lobe of (3,) ``amplitude`` [Hz/m]: ``rise``, ``flat`` and ``fall`` times [s] after ``delay``."""

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

    def corner_times(self) -> np.ndarray:
        """(4,) start of the rise, top of the rise, end of the flat top, end of the fall."""
        return self.delay + np.cumsum([0.0, self.rise, self.flat, self.fall])

    def corners(self) -> tuple[np.ndarray, Tensor]:
        """The lobe as a piecewise-linear waveform: corner times and (4, 3) values."""
        on = torch.tensor([0.0, 1.0, 1.0, 0.0], dtype=self.amplitude.dtype)
        return self.corner_times(), on[:, None] * self.amplitude[None, :]


@dataclass
class RFPulse:
    """``signal`` (k, 2) [Hz], each sample held for ``raster`` [s], starting ``delay`` after the block start;
    ``center``: from the pulse start to its effective centre."""

    signal: Tensor
    raster: float
    delay: float
    center: float

    @property
    def duration(self) -> float:
        return self.signal.shape[0] * self.raster


@dataclass
class ADC:
    """``n`` samples at ``delay + (i + 1/2) dwell`` after the block start, with receiver ``phase`` [rad]."""

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


def check_limits(g: Trapezoid, sysm: System) -> None:
    """ValueError when a lobe exceeds the amplitude limit, jumps (a ramp of zero time to a nonzero level), or
    exceeds the slew limit on any axis."""
    peak = float(g.amplitude.detach().abs().max()) if g.amplitude.numel() else 0.0
    if peak > sysm.max_grad * (1 + 1e-9):
        raise ValueError(f"gradient {peak:.6g} Hz/m exceeds max_grad {sysm.max_grad:.6g} Hz/m")
    if peak == 0.0:
        return
    for ramp in (g.rise, g.fall):
        if ramp <= T_TOL:
            raise ValueError("gradient jumps instantaneously (infinite slew rate)")
        if peak / ramp > sysm.max_slew * (1 + 1e-9):
            raise ValueError(f"slew rate {peak / ramp:.6g} Hz/m/s exceeds max_slew {sysm.max_slew:.6g} Hz/m/s")


# ---- lobe timing: one solver per way of asking for a lobe; the array forms serve the library's batched TRs


def shortest_ramp_flat(area, sysm: System, ramp=None):
    """Ramp and flat time of the shortest lobe of |``area``| (scalar or array): a triangle while that stays under
    max_grad, else ramps to max_grad and the flat top that completes the area, both on the gradient raster.
    ``ramp`` fixes the ramp time instead."""
    a = np.abs(np.asarray(area, dtype=np.float64))
    triangle = a <= sysm.max_grad**2 / sysm.max_slew
    ramp_tri = ceil_to(np.sqrt(a / sysm.max_slew), sysm.grad_raster) if ramp is None else np.full_like(a, ramp)
    ramp_max = sysm.ramp_time(sysm.max_grad) if ramp is None else ramp
    flat_max = ceil_to(a / sysm.max_grad - ramp_max, sysm.grad_raster)
    ramp = np.where(triangle, ramp_tri, ramp_max)
    flat = np.where(triangle, 0.0, flat_max)
    if ramp.ndim == 0:
        return float(ramp), float(flat)
    return ramp, flat


def ramp_for_duration(area, duration: float, sysm: System):
    """The ramp of the lowest lobe of |``area``| (scalar or array) lasting ``duration``: the smaller root of
    r (duration - r) max_slew = area, rounded up to the gradient raster (0 for zero area)."""
    a = np.abs(np.asarray(area, dtype=np.float64))
    disc = duration**2 - 4 * a / sysm.max_slew
    if np.any(disc < 0):
        worst = float(a.max())
        raise ValueError(f"area {worst:.6g} /m does not fit in {duration * 1e3:.4g} ms at the slew limit")
    ramp = np.where(a > 0, ceil_to(0.5 * (duration - np.sqrt(np.maximum(disc, 0.0))), sysm.grad_raster), 0.0)
    return float(ramp) if ramp.ndim == 0 else ramp


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
    """A lobe within ``system``'s limits, asked for one of these ways: ``area`` (the shortest lobe), ``area`` and
    ``duration`` (the lowest lobe of that length), ``flat_area`` and ``flat_time``, ``amplitude`` and ``flat_time`` or
    ``duration``. ``rise_time`` overrides the ramp. Timings come from values; areas and amplitudes may require grad.
    ``axis`` puts scalar values on "x", "y" or "z"; without it they are 3-vectors (the largest component sets the
    timing)."""
    sysm = system or System()
    as_vec = lambda v: axis_vector(v, axis, dtype)
    if amplitude is not None:
        if flat_time is None and duration is None:
            raise ValueError("amplitude needs flat_time or duration")
        amp = as_vec(amplitude)
        ramp = sysm.ramp_time(amp.detach().abs().max()) if rise_time is None else rise_time
        flat = flat_time if flat_time is not None else duration - 2 * ramp
    elif flat_area is not None:
        if flat_time is None:
            raise ValueError("flat_area needs flat_time")
        amp = as_vec(flat_area) / flat_time
        ramp = sysm.ramp_time(amp.detach().abs().max()) if rise_time is None else rise_time
        flat = flat_time
    elif area is not None:
        target = as_vec(area)
        peak = float(target.detach().abs().max())
        if duration is not None:
            ramp = ramp_for_duration(peak, duration, sysm) if rise_time is None else rise_time
            flat = duration - 2 * ramp
            if flat < -T_TOL:
                raise ValueError(f"ramps of {ramp * 1e6:.4g} us do not fit in {duration * 1e3:.4g} ms")
        else:
            ramp, flat = shortest_ramp_flat(peak, sysm, rise_time)
        flat = max(flat, 0.0)
        amp = target / (flat + ramp) if flat + ramp > 0 else target * 0.0
    else:
        raise ValueError("give area, flat_area or amplitude")
    if flat < -T_TOL:
        raise ValueError(f"negative flat time {flat:.3g} s")
    lobe = Trapezoid(amp, float(ramp), max(float(flat), 0.0), float(ramp), float(delay))
    check_limits(lobe, sysm)
    return lobe


def rf_steps(duration) -> int:
    """Segments of a hard pulse: as few as keep each at most ``MAX_RF_STEP``."""
    return max(1, math.ceil(float(torch.as_tensor(duration).detach()) / MAX_RF_STEP - 1e-6))


def _polar_rows(amp: Tensor, phase: Tensor) -> Tensor:
    """(..., 2) RF samples amp (cos phase, sin phase)."""
    return torch.stack([amp * torch.cos(phase), amp * torch.sin(phase)], dim=-1)


def block_pulse(flip_deg, duration: float, *, phase=0.0, system: System | None = None, dtype=torch.float64,
                n_steps: int | None = None) -> RFPulse:
    """A hard pulse of ``flip_deg`` after the RF dead time, in ``n_steps`` raster steps (default ``rf_steps``)."""
    sysm = system or System()
    k = rf_steps(duration) if n_steps is None else n_steps
    amp = torch.as_tensor(flip_deg, dtype=dtype) / (360.0 * duration)
    row = _polar_rows(amp, torch.as_tensor(phase, dtype=dtype))
    return RFPulse(row[None].expand(k, 2), float(duration) / k, sysm.rf_dead_time, 0.5 * duration)


def sinc_envelope(n: int, raster: float, duration: float, time_bw_product: float, apodization: float) -> np.ndarray:
    """A windowed sinc of ``n`` raster samples, normalised to one cycle of area (1 Hz x 1 s): at the sample centres
    t (from the pulse centre), ((1 - a) + a cos(2 pi t / T)) sinc(t TBW / T) for a pulse of length T."""
    t = (np.arange(n) + 0.5) * raster - 0.5 * duration
    hann = (1 - apodization) + apodization * np.cos(2 * np.pi * t / duration)
    env = hann * np.sinc(time_bw_product / duration * t)
    return env / (env.sum() * raster)


def sinc_pulse(flip_deg, duration: float, *, slice_thickness: float, time_bw_product: float = 4.0,
               apodization: float = 0.5, slice_offset: float = 0.0, phase=0.0, system: System | None = None,
               dtype=torch.float64) -> tuple[RFPulse, Trapezoid, Tensor]:
    """Slice-selective windowed sinc: the envelope of ``sinc_envelope`` on the RF raster, scaled to ``flip_deg`` at
    the slice centre and frequency-shifted to ``slice_offset``; a slice-select lobe on z of amplitude
    bandwidth / thickness whose flat top (on the gradient raster) centres the pulse. Returns (pulse, lobe, the z
    area from the pulse centre to the end of the lobe's fall, negated: the rephasing moment)."""
    sysm = system or System()
    sysm.check_raster(duration, sysm.rf_raster, "RF duration")
    n = round(duration / sysm.rf_raster)
    bandwidth = time_bw_product / duration
    g_amp = bandwidth / slice_thickness
    env = torch.as_tensor(sinc_envelope(n, sysm.rf_raster, duration, time_bw_product, apodization), dtype=dtype)
    t = (np.arange(n) + 0.5) * sysm.rf_raster - 0.5 * duration
    offset_phase = torch.as_tensor(2 * np.pi * g_amp * slice_offset * t, dtype=dtype)
    signal = _polar_rows(torch.as_tensor(flip_deg, dtype=dtype) / 360.0 * env,
                         torch.as_tensor(phase, dtype=dtype) - offset_phase)
    ramp = sysm.ramp_time(g_amp)
    flat = ceil_to(duration, sysm.grad_raster)
    lead = 0.5 * (flat - duration)  # the pulse sits centred on the flat top
    g_delay = max(0.0, sysm.rf_dead_time - ramp - lead)
    gz = trapezoid("z", amplitude=g_amp, flat_time=flat, rise_time=ramp, delay=g_delay, system=sysm, dtype=dtype)
    pulse = RFPulse(signal, sysm.rf_raster, g_delay + ramp + lead, 0.5 * duration)
    return pulse, gz, -gz.amplitude * 0.5 * (flat + ramp)


def block_duration(*grads, rf: RFPulse | None = None, adc: ADC | list[ADC] | None = None,
                   system: System | None = None) -> float:
    """When the last lobe or ADC sample ends, or the RF pulse plus the ringdown time, whichever is latest."""
    adcs = [] if adc is None else ([adc] if isinstance(adc, ADC) else list(adc))
    end = max([0.0, *(g.duration for g in grads), *(a.duration for a in adcs)])
    if rf is not None:
        end = max(end, rf.delay + rf.duration + (system or System()).rf_ringdown_time)
    return end


# ---- recorded rows and the batched compiler


@dataclass
class _Blocks:
    """Blocks recorded for the batched split, as struct-of-arrays rows (block indices are this table's)."""

    end: list = field(default_factory=list)  # block lengths [s]
    spoil: list = field(default_factory=list)  # block index whose last segment is spoiled
    fill_dt: dict = field(default_factory=dict)  # block -> tensor length of an event-free block (keeps its graph)
    # lobes: block, corner times (4,), (3,) amplitude, sloped (nonzero amplitude)
    lobe_block: list = field(default_factory=list)
    lobe_t: list = field(default_factory=list)
    lobe_amp: list = field(default_factory=list)
    # RF pulses: block, delay, raster, sample count, (k, 2) samples
    rf_block: list = field(default_factory=list)
    rf_delay: list = field(default_factory=list)
    rf_raster: list = field(default_factory=list)
    rf_signal: list = field(default_factory=list)
    # ADC events: block, sample times (n,), phase
    adc_block: list = field(default_factory=list)
    adc_t: list = field(default_factory=list)
    adc_phase: list = field(default_factory=list)

    grad_raster: float = 10e-6  # ramps are cut at this raster

    def __len__(self) -> int:
        return len(self.end)


def _concat_rows(parts, dtype) -> Tensor:
    return torch.cat([torch.as_tensor(p, dtype=dtype).reshape(-1, *np.shape(p)[1:]) if not isinstance(p, Tensor)
                      else p.to(dtype) for p in parts])


def compile_blocks(blocks: _Blocks, dtype, device=None) -> dict:
    """Split every recorded block into segments (the module's rules) in one pass. Returns the segment table of all
    blocks in block order: dt, rf, grad, adc_phase (tensors), is_rf, adc, spoil (bool arrays) and ``count`` (segments
    per block)."""
    B = len(blocks)
    end = np.asarray(blocks.end, dtype=np.float64)
    lb = np.asarray(blocks.lobe_block, dtype=np.int64)
    lt = np.asarray(blocks.lobe_t, dtype=np.float64).reshape(-1, 4)
    rb = np.asarray(blocks.rf_block, dtype=np.int64)
    rd = np.asarray(blocks.rf_delay, dtype=np.float64)
    rr = np.asarray(blocks.rf_raster, dtype=np.float64)
    rk = np.asarray([s.shape[0] for s in blocks.rf_signal], dtype=np.int64)
    ab = np.asarray(blocks.adc_block, dtype=np.int64)
    an = np.asarray([len(t) for t in blocks.adc_t], dtype=np.int64)
    amp = _concat_rows(blocks.lobe_amp, dtype).reshape(-1, 3) if len(lb) else torch.zeros(0, 3, dtype=dtype)
    sloped = (amp.detach() != 0).any(dim=1).numpy() if len(lb) else np.zeros(0, bool)

    # -- knots: (block, time)
    kb = [np.arange(B), np.arange(B)]
    kt = [np.zeros(B), end]
    if len(lb):
        kb.append(np.repeat(lb, 4))
        kt.append(lt.reshape(-1))
        for lo, hi in ((0, 1), (2, 3)):  # ramps: raster points cutting each into equal parts
            span = lt[:, hi] - lt[:, lo]
            parts = np.where(sloped & (span > T_TOL), np.ceil(span / blocks.grad_raster - 1e-9), 1).astype(np.int64)
            cut = parts - 1
            if cut.sum():
                owner = np.repeat(np.arange(len(lb)), cut)
                j = np.arange(cut.sum()) - np.repeat(np.cumsum(cut) - cut, cut) + 1
                kb.append(lb[owner])
                kt.append(lt[owner, lo] + span[owner] * j / parts[owner])
    if len(rb):
        owner = np.repeat(np.arange(len(rb)), rk + 1)
        j = np.arange(owner.size) - np.repeat(np.cumsum(rk + 1) - (rk + 1), rk + 1)
        kb.append(rb[owner])
        kt.append(rd[owner] + rr[owner] * j)
    if len(ab):
        kb.append(np.repeat(ab, an))
        kt.append(np.concatenate(blocks.adc_t))
    kb, kt = np.concatenate(kb), np.concatenate(kt)
    if np.any(kt < -T_TOL) or np.any(kt > end[kb] + T_TOL):
        bad = int(kb[np.flatnonzero((kt < -T_TOL) | (kt > end[kb] + T_TOL))[0]])
        raise ValueError(f"events extend outside the block [0, {end[bad]:.9g}] s")
    order = np.lexsort((kt, kb))
    kb, kt = kb[order], kt[order]
    first = np.ones(kb.size, dtype=bool)
    first[1:] = (kb[1:] != kb[:-1]) | (np.diff(kt) > T_TOL)
    keep = first & (kt < end[kb] - T_TOL)  # the block's end knot is put back below
    kb, kt = kb[keep], kt[keep]
    count = np.bincount(kb, minlength=B)  # segments per block (each block keeps its start knot: end > T_TOL)
    seg_block = kb
    t0 = kt
    nxt = np.empty_like(kt)
    nxt[:-1] = kt[1:]
    last = np.ones(kb.size, dtype=bool)
    last[:-1] = kb[1:] != kb[:-1]
    t1 = np.where(last, end[kb], nxt)
    S = kt.size
    seg0 = np.cumsum(count) - count  # first segment of each block
    mid = 0.5 * (t1 + t0)
    dt_np = t1 - t0

    # -- gradients: every lobe's level at the midpoints of its block's segments, times its amplitude
    grad = torch.zeros(S, 3, dtype=dtype)
    if len(lb):
        n_of = count[lb]
        pair_lobe = np.repeat(np.arange(len(lb)), n_of)
        pair_seg = seg0[lb][pair_lobe] + np.arange(n_of.sum()) - np.repeat(np.cumsum(n_of) - n_of, n_of)
        c = lt[pair_lobe]
        m = mid[pair_seg]
        with np.errstate(divide="ignore", invalid="ignore"):
            up = (m - c[:, 0]) / (c[:, 1] - c[:, 0])
            down = (c[:, 3] - m) / (c[:, 3] - c[:, 2])
        level = np.where(m < c[:, 1], up, np.where(m <= c[:, 2], 1.0, down))
        level = np.where((m < c[:, 0]) | (m >= c[:, 3]), 0.0, level)
        live = level != 0.0
        w = torch.as_tensor(level[live], dtype=dtype)
        grad = grad.index_add(0, torch.as_tensor(pair_seg[live]), w[:, None] * amp[torch.as_tensor(pair_lobe[live])])

    # -- RF: the raster sample under each midpoint inside the pulse
    rf = torch.zeros(S, 2, dtype=dtype)
    is_rf = np.zeros(S, dtype=bool)
    if len(rb):
        n_of = count[rb]
        pair_rf = np.repeat(np.arange(len(rb)), n_of)
        pair_seg = seg0[rb][pair_rf] + np.arange(n_of.sum()) - np.repeat(np.cumsum(n_of) - n_of, n_of)
        m = mid[pair_seg]
        d, r, k = rd[pair_rf], rr[pair_rf], rk[pair_rf]
        inside = (m > d) & (m < d + k * r)
        idx = np.clip(np.floor((m - d) / r).astype(np.int64), 0, k - 1) + (np.cumsum(rk) - rk)[pair_rf]
        signal = _concat_rows(blocks.rf_signal, dtype)
        rf = rf.index_put((torch.as_tensor(pair_seg[inside]),), signal[torch.as_tensor(idx[inside])])
        is_rf[pair_seg[inside]] = True

    # -- ADC: each sample is the end knot of one segment of its block
    adc = np.zeros(S, dtype=bool)
    phase = torch.zeros(S, dtype=dtype)
    if len(ab):
        sb = np.repeat(ab, an)
        st = np.concatenate(blocks.adc_t)
        # the first knot of the sample's block at or after (t - T_TOL): knots and queries sorted together
        qt = st - T_TOL
        allb = np.concatenate([kb, sb])
        allt = np.concatenate([kt, qt])
        kind = np.concatenate([np.ones(kb.size, np.int8), np.zeros(sb.size, np.int8)])  # queries first on ties
        o = np.lexsort((kind, allt, allb))
        is_knot = kind[o] == 1
        knots_before = np.cumsum(is_knot) - is_knot  # knots strictly before each entry
        pos = np.empty(sb.size, dtype=np.int64)
        pos[o[~is_knot] - kb.size] = knots_before[~is_knot]
        local = pos - seg0[sb]  # the knot's index in its block; the block's end knot is local index count
        seg = pos - 1
        knot_t = np.where(local < count[sb], kt[np.minimum(pos, S - 1)], end[sb])
        if np.any(local < 1) or np.any(local > count[sb]) or np.any(np.abs(knot_t - st) > T_TOL):
            raise ValueError("ADC samples must lie inside the block, after its start, without overlap")
        if np.unique(seg).size != seg.size:
            raise ValueError("ADC samples must lie inside the block, after its start, without overlap")
        adc[seg] = True
        if any(isinstance(p, Tensor) for p in blocks.adc_phase):
            ph = [torch.as_tensor(p, dtype=dtype).reshape(()) for p in blocks.adc_phase]
            per_sample = torch.cat([p.to(dtype).expand(int(n)) for p, n in zip(ph, an)])
        else:
            per_sample = torch.as_tensor(np.repeat(np.asarray(blocks.adc_phase, dtype=np.float64), an), dtype=dtype)
        phase = phase.index_put((torch.as_tensor(seg),), per_sample)

    dt = torch.as_tensor(dt_np, dtype=dtype)
    if blocks.fill_dt:
        fb = np.asarray(list(blocks.fill_dt), dtype=np.int64)
        dt = dt.index_put((torch.as_tensor(seg0[fb]),), torch.stack([v.to(dtype).reshape(())
                                                                       for v in blocks.fill_dt.values()]))
    spoil = np.zeros(S, dtype=bool)
    if blocks.spoil:
        sp = np.asarray(blocks.spoil, dtype=np.int64)
        spoil[seg0[sp] + count[sp] - 1] = True
    dev = lambda t: t if device is None else t.to(device)
    return {"dt": dev(dt), "rf": dev(rf), "grad": dev(grad), "adc_phase": dev(phase), "is_rf": is_rf, "adc": adc,
            "spoil": spoil, "count": count, "seg_block": seg_block}


class SequenceBuilder:
    """Appends blocks and plain segments in time order; ``build()`` gives the ``Sequence``. Numeric arguments may
    require grad (areas, amplitudes, flip angles, RF samples; the lengths of ``delay``, ``rf``, ``rf_waveform`` and
    ``readout``)."""

    def __init__(self, dtype=torch.float64, device=None, system: System | None = None):
        self.dtype, self.device = dtype, device
        self.system = system or System()
        self.time = 0.0
        self._blocks = _Blocks(grad_raster=self.system.grad_raster)
        # the timeline: ("B", block index) or ("S", index into _plain), in order
        self._items: list[tuple[str, int]] = []
        self._plain: list[dict] = []

    # ---- plain segments (no splitting)

    def _t(self, x) -> Tensor:
        return torch.as_tensor(x, dtype=self.dtype, device=self.device)

    def _grad_rows(self, grad, n: int) -> Tensor:
        if grad is None:
            return torch.zeros(n, 3, dtype=self.dtype, device=self.device)
        g = grad.to(dtype=self.dtype, device=self.device) if isinstance(grad, Tensor) else torch.stack(
            [self._t(x) for x in grad])
        return g if g.ndim == 2 else g.expand(n, 3)

    def _plain_segments(self, dt: Tensor, rf: Tensor, grad: Tensor, phase: Tensor, is_rf, adc) -> None:
        n = dt.shape[0]
        flags = lambda v: np.asarray(v, dtype=bool) if isinstance(v, list) else np.full(n, v, dtype=bool)
        self._plain.append({"dt": dt, "rf": rf, "grad": grad, "adc_phase": phase, "is_rf": flags(is_rf),
                            "adc": flags(adc), "spoil": np.zeros(n, dtype=bool)})
        self._items.append(("S", len(self._plain) - 1))
        self.time += float(dt.detach().sum())

    def rf(self, duration, *, flip_deg=None, amplitude=None, phase=0.0, grad=None, n_steps=None):
        """A hard pulse of ``flip_deg`` (or ``amplitude`` [Hz]) in ``n_steps`` equal segments (default
        ``rf_steps``), under a constant ``grad``."""
        length = self._t(duration)
        amp = self._t(flip_deg) / (360.0 * length) if flip_deg is not None else self._t(amplitude)
        k = rf_steps(length) if n_steps is None else n_steps
        dt = (length / k).expand(k)
        self._plain_segments(dt, _polar_rows(amp, self._t(phase)).expand(k, 2), self._grad_rows(grad, k),
                             torch.zeros_like(dt), True, False)
        return self

    def rf_waveform(self, samples, dt, grad=None):
        """Arbitrary RF: (k, 2) samples [Hz], one segment of ``dt`` each."""
        rf = self._t(samples)
        k = rf.shape[0]
        dts = self._t(dt).expand(k)
        self._plain_segments(dts, rf, self._grad_rows(grad, k), torch.zeros_like(dts), True, False)
        return self

    def delay(self, duration, grad=None):
        """One segment of ``duration`` under a constant ``grad``."""
        dt = self._t(duration).reshape(1)
        self._plain_segments(dt, torch.zeros(1, 2, dtype=self.dtype, device=self.device), self._grad_rows(grad, 1),
                             torch.zeros_like(dt), False, False)
        return self

    def readout(self, n: int, dwell, grad=None, phase=0.0):
        """``n`` samples at (i + 1/2) dwell under a constant ``grad``: segments dwell/2, (n - 1) x dwell, dwell/2,
        each of the first n ending on a sample."""
        dw = self._t(dwell)
        half = (0.5 * dw).reshape(1)
        dt = torch.cat([half, dw.expand(n - 1), half])
        self._plain_segments(dt, torch.zeros(n + 1, 2, dtype=self.dtype, device=self.device),
                             self._grad_rows(grad, n + 1), self._t(phase).expand(n + 1), False,
                             [True] * n + [False])
        return self

    # ---- blocks (split at build)

    def _new_block(self, length: float) -> int:
        rows = self._blocks
        rows.end.append(float(length))
        b = len(rows.end) - 1
        self._items.append(("B", b))
        self.time += float(length)
        return b

    def _lobe(self, b: int, g: Trapezoid) -> None:
        rows = self._blocks
        rows.lobe_block.append(b)
        rows.lobe_t.append(g.corner_times())
        rows.lobe_amp.append(g.amplitude.to(self.dtype).reshape(1, 3))

    def block(self, *grads, rf: RFPulse | None = None, adc: ADC | list[ADC] | None = None,
              duration: float | None = None):
        """Lobes, at most one RF pulse and ADC events played together, split at build by the module's rules.
        ``duration`` (at least the events' own) pads the block; a block of no length adds nothing."""
        sysm = self.system
        adcs = [] if adc is None else ([adc] if isinstance(adc, ADC) else list(adc))
        if rf is not None and rf.delay < sysm.rf_dead_time - T_TOL:
            raise ValueError(f"RF starts {rf.delay * 1e6:.4g} us into the block, before the dead time")
        length = block_duration(*grads, rf=rf, adc=adcs, system=sysm)
        if duration is not None:
            if duration < length - T_TOL:
                raise ValueError(f"block events last {length:.9g} s, longer than duration {duration:.9g} s")
            length = float(duration)
        if length <= T_TOL:
            return self
        b = self._new_block(length)
        rows = self._blocks
        for g in grads:
            self._lobe(b, g)
        if rf is not None:
            rows.rf_block.append(b)
            rows.rf_delay.append(float(rf.delay))
            rows.rf_raster.append(float(rf.raster))
            rows.rf_signal.append(rf.signal)
        for a in adcs:
            rows.adc_block.append(b)
            rows.adc_t.append(a.times)
            rows.adc_phase.append(a.phase)
        return self

    def trapezoid(self, axis: str | None = None, **kwargs):
        return self.block(trapezoid(axis, system=self.system, dtype=self.dtype, **kwargs))

    def sinc_pulse(self, flip_deg, duration: float, *, slice_thickness: float, rephase: bool = True, **kwargs):
        pulse, gz, moment = sinc_pulse(flip_deg, duration, slice_thickness=slice_thickness, system=self.system,
                                       dtype=self.dtype, **kwargs)
        self.block(gz, rf=pulse)
        if rephase:
            self.block(trapezoid(area=moment, system=self.system, dtype=self.dtype))
        return self

    def spoil(self):
        """An ideal spoiler at the end of the last segment so far."""
        kind, i = self._items[-1]
        if kind == "B":
            self._blocks.spoil.append(i)
        else:
            self._plain[i]["spoil"][-1] = True
        return self

    # ---- the library's batched recording (whole TR trains at once)

    def _record_train(self, train: "Train") -> None:
        """Append ``train``'s blocks (already in time order) to the timeline."""
        rows = self._blocks
        b0 = len(rows.end)
        rows.end.extend(train.end.tolist())
        self._items.extend(("B", b0 + i) for i in range(len(train.end)))
        self.time += float(train.end.sum())
        rows.spoil.extend((b0 + np.asarray(train.spoil, dtype=np.int64)).tolist())
        for i, v in train.fill_dt.items():
            rows.fill_dt[b0 + i] = v
        if len(train.lobe_block):
            rows.lobe_block.extend((b0 + train.lobe_block).tolist())
            rows.lobe_t.extend(list(train.lobe_t))
            rows.lobe_amp.append(train.lobe_amp.to(self.dtype))
        for blk, delay, raster, signal in train.rf:
            rows.rf_block.extend((b0 + blk).tolist())
            rows.rf_delay.extend([float(delay)] * len(blk))
            rows.rf_raster.extend([float(raster)] * len(blk))
            rows.rf_signal.extend(list(signal.to(self.dtype).unbind(0)) if len(blk) else [])
        for blk, times, phase in train.adc:
            rows.adc_block.extend((b0 + blk).tolist())
            rows.adc_t.extend([times] * len(blk))
            rows.adc_phase.extend(list(phase))

    def build(self, **meta) -> Sequence:
        meta = {"system": self.system, **meta}
        parts = {}
        if self._blocks.end:
            parts["B"] = compile_blocks(self._blocks, self.dtype, self.device)
        # segment ranges of every item, then one gather in timeline order
        keys = ("dt", "rf", "grad", "adc_phase")
        blk = parts.get("B")
        n_blk = 0 if blk is None else int(blk["count"].sum())
        seg0 = None if blk is None else np.cumsum(blk["count"]) - blk["count"]
        plain_len = [p["dt"].shape[0] for p in self._plain]
        plain0 = n_blk + np.concatenate([[0], np.cumsum(plain_len)]).astype(np.int64)
        starts, lens = [], []
        for kind, i in self._items:
            if kind == "B":
                starts.append(seg0[i])
                lens.append(blk["count"][i])
            else:
                starts.append(plain0[i])
                lens.append(plain_len[i])
        starts, lens = np.asarray(starts, dtype=np.int64), np.asarray(lens, dtype=np.int64)
        take = np.repeat(starts - (np.cumsum(lens) - lens), lens) + np.arange(lens.sum())
        src = {k: ([blk[k]] if blk is not None else []) + [p[k] for p in self._plain] for k in
               keys + ("is_rf", "adc", "spoil")}
        cat = {k: torch.cat([t.to(dtype=self.dtype, device=self.device) for t in src[k]]) for k in keys}
        flags = {k: np.concatenate(src[k]) for k in ("is_rf", "adc", "spoil")}
        if np.array_equal(take, np.arange(take.size)):
            out = cat
        else:
            idx = torch.as_tensor(take, device=self.device)
            out = {k: v[idx] for k, v in cat.items()}
            flags = {k: v[take] for k, v in flags.items()}
        return Sequence(dt=out["dt"], rf=out["rf"], grad=out["grad"], adc_phase=out["adc_phase"],
                        is_rf=flags["is_rf"].tolist(), adc=flags["adc"].tolist(), spoil=flags["spoil"].tolist(),
                        meta=meta)


@dataclass
class Train:
    """Blocks of a whole sequence recorded at once (``SequenceBuilder._record_train``), block indices local:
    ``end`` (B,) lengths; ``spoil`` blocks whose last segment is spoiled; ``fill_dt`` event-free blocks whose
    length is a tensor; lobes as arrays (L,), (L, 4) corner times, (L, 3) amplitudes; ``rf`` a list of (blocks
    (R,), delay, raster, (R, k, 2) samples); ``adc`` a list of (blocks (A,), sample times (n,), phases (A,))."""

    end: np.ndarray
    lobe_block: np.ndarray
    lobe_t: np.ndarray
    lobe_amp: Tensor
    rf: list
    adc: list
    spoil: list
    fill_dt: dict = field(default_factory=dict)
