"""Ready-made sequences for the GPU path: spoiled Cartesian GRE (2D / 3D, ideal or realistic) and 2D TSE.

Reference interface: ``src/blochparty/sequences/library.py`` — ``gre_cartesian`` (:43-138), ``gre_2d`` / ``gre_3d``
(:143-150), ``tse_2d`` (:392-640). Argument names and meanings, the timing of every event and the k-space order are
the reference's (its segment tables are the golden fixtures these are checked against, tests/golden/sequences.npz);
multi-slice cross-talk (``slice_order``, ``other_rf_raster``, ``bound_only_beyond``) and partial acquisitions
(``acquired_lines``) are not part of this package and raise ``NotImplementedError``.

Design: a sequence is a train of identical repetitions that differ only in amplitudes (phase encodes, RF phase, flip
angles) and in whether they sample. Each builder first fixes the repetition's timing once (``_Excitation``,
``_Readout``, the lobe durations, the padding), then lays every block of every repetition out as rows of a
``Train`` in one pass of array arithmetic: the per-repetition lobe ramps (which depend on each phase encode's area)
come from the vectorised timing solvers of ``builder.py``, and the RF samples of all repetitions are one (R, k, 2)
tensor. ``SequenceBuilder`` splits the whole train in one batched pass.
"""

from __future__ import annotations

import math
from typing import NamedTuple

import numpy as np
import torch
from torch import Tensor

from .builder import (
    RFPulse,
    SequenceBuilder,
    System,
    Train,
    Trapezoid,
    axis_vector,
    block_duration,
    block_pulse,
    ceil_to,
    check_limits,
    ramp_for_duration,
    rf_steps as hard_pulse_steps,
    shortest_ramp_flat,
    sinc_pulse,
    trapezoid,
)
from .sequence import T_TOL, Sequence

SPOILING = ("ideal", "gradient", "rf")


# ---- the pieces of a repetition


class _Excitation(NamedTuple):
    """A pulse block: its lobes (none, or the slice select), the pulse, the z moment that rephases it, the block's
    length, and the pulse centre's distance from the block's start (``lead``) and end (``tail``)."""

    lobes: tuple
    pulse: RFPulse
    rephase_z: float
    length: float
    lead: float
    tail: float


def _excitation(flip_deg, rf_duration, slice_thickness, slice_offset, time_bw_product, apodization, sysm, dtype,
                n_steps=None) -> _Excitation:
    if slice_thickness is None:
        pulse = block_pulse(flip_deg, 20e-6 if rf_duration is None else rf_duration, system=sysm, dtype=dtype,
                            n_steps=n_steps)
        lobes, rephase_z = (), 0.0
    else:
        pulse, gz, moment = sinc_pulse(flip_deg, 2e-3 if rf_duration is None else rf_duration,
                                       slice_thickness=slice_thickness, time_bw_product=time_bw_product,
                                       apodization=apodization, slice_offset=slice_offset, system=sysm, dtype=dtype)
        lobes, rephase_z = (gz,), float(moment[2])
    length = block_duration(*lobes, rf=pulse, system=sysm)
    lead = pulse.delay + pulse.center
    return _Excitation(lobes, pulse, rephase_z, length, lead, length - lead)


class _Readout(NamedTuple):
    """The readout lobe on x with its ADC on the flat top: ``adc_delay`` from the block start to the ADC,
    ``to_centre`` from the block start to the k = 0 sample, ``moment_centre`` the x moment up to that sample,
    ``moment`` the lobe's whole area, ``length`` the block's."""

    lobe: Trapezoid
    adc_delay: float
    to_centre: float
    moment_centre: float
    moment: float
    length: float


def _readout(n: int, dwell: float, fov: float, ramps: bool, sysm: System, dtype, echo_centred: bool) -> _Readout:
    """Samples i = 0..n-1 at kx = (i - n/2) / fov. ``echo_centred`` (spin echoes) starts the lobe so the ADC sits
    half a dwell into the flat top's first raster and puts k = 0 at the lobe's centre."""
    sysm.check_raster(dwell, sysm.adc_raster, "dwell")
    level, flat = (1.0 / fov) / dwell, n * dwell
    ramp = sysm.ramp_time(level) if ramps else 0.0
    start = max(0.0, 0.5 * dwell - ramp) if echo_centred else 0.0
    if ramps:
        lobe = trapezoid("x", amplitude=level, flat_time=flat, system=sysm, dtype=dtype)
        lobe = Trapezoid(lobe.amplitude, lobe.rise, lobe.flat, lobe.fall, start)
    else:
        lobe = Trapezoid(axis_vector(level, "x", dtype), 0.0, flat, 0.0, start)
    if echo_centred:
        adc_delay = start + ramp - 0.5 * dwell
        to_centre = start + ramp + 0.5 * flat
        moment_centre = level * (0.5 * ramp + 0.5 * flat)
    else:
        adc_delay = ramp
        to_centre = ramp + (n / 2 + 0.5) * dwell
        moment_centre = level * 0.5 * ramp + (n / 2 + 0.5) / fov
    return _Readout(lobe, adc_delay, to_centre, moment_centre, level * (ramp + flat), lobe.duration)


def _shortest_length(area: float, sysm: System) -> float:
    ramp, flat = shortest_ramp_flat(area, sysm)
    return 0.0 + ramp + flat + ramp


def _lobes_in(areas: np.ndarray, length: float, ramps: bool, sysm: System) -> tuple[np.ndarray, np.ndarray]:
    """Lobes of (L, 3) ``areas`` that each last ``length``: with ramps the lowest trapezoid (its ramp from the
    largest component), else a rectangle. Returns (L, 4) corner times and (L, 3) amplitudes; the lobes are checked
    against the system's limits."""
    L = areas.shape[0]
    if not ramps:
        corners = np.tile(np.cumsum([0.0, 0.0, length, 0.0]), (L, 1))
        return corners, areas / length if length > 0 else areas * 0.0
    peak = np.abs(areas).max(axis=1)
    ramp = ramp_for_duration(peak, length, sysm)
    flat = length - 2 * ramp
    if np.any(flat < -T_TOL):
        raise ValueError(f"ramps of {float(ramp.max()) * 1e6:.4g} us do not fit in {length * 1e3:.4g} ms")
    flat = np.maximum(flat, 0.0)
    span = flat + ramp
    amps = np.where(span[:, None] > 0, areas / np.where(span > 0, span, 1.0)[:, None], 0.0)
    top = np.abs(amps).max(axis=1)
    if np.any(top > sysm.max_grad * (1 + 1e-9)):
        raise ValueError(f"gradient {float(top.max()):.6g} Hz/m exceeds max_grad {sysm.max_grad:.6g} Hz/m")
    live = top > 0
    if np.any(live & (ramp <= T_TOL)):
        raise ValueError("gradient jumps instantaneously (infinite slew rate)")
    if np.any(live & (top > sysm.max_slew * (1 + 1e-9) * np.where(live, ramp, 1.0))):
        raise ValueError(f"slew rate exceeds max_slew {sysm.max_slew:.6g} Hz/m/s")
    corners = np.cumsum(np.stack([np.zeros(L), ramp, flat, ramp], axis=1), axis=1)
    return corners, amps


def _shortest_lobe(area: np.ndarray, ramps: bool, sysm: System) -> tuple[np.ndarray, np.ndarray]:
    """The shortest lobe of (3,) ``area``: with ramps the shortest trapezoid of its largest component, else a
    rectangle as long as that trapezoid. Returns (4,) corner times and (3,) amplitudes."""
    peak = float(np.abs(area).max())
    if not ramps:
        return _lobes_in(area[None], _shortest_length(peak, sysm), False, sysm)
    ramp, flat = shortest_ramp_flat(peak, sysm)
    flat = max(flat, 0.0)
    lobe = Trapezoid(torch.as_tensor(area / (flat + ramp) if flat + ramp > 0 else area * 0.0), ramp, flat, ramp)
    check_limits(lobe, sysm)
    return lobe.corner_times(), lobe.amplitude.numpy()


def _turned(signal: Tensor, phase: Tensor) -> Tensor:
    """(R, k, 2): ``signal`` (k, 2) turned by each of the (R,) phases [rad]."""
    c, s = torch.cos(phase)[:, None], torch.sin(phase)[:, None]
    x, y = signal[:, 0][None], signal[:, 1][None]
    return torch.stack([c * x - s * y, s * x + c * y], dim=-1)


def _rf_spoil_phase(j: np.ndarray, increment_deg: float) -> np.ndarray:
    """Quadratic RF spoiling: the j-th repetition's phase is increment j (j + 1) / 2 degrees, wrapped to 360."""
    return np.radians(np.mod(increment_deg * j * (j + 1) / 2, 360.0))


class _Layout:
    """Rows of a ``Train`` laid out as repetitions x slots: block (rep, slot) is rep * slots + slot."""

    def __init__(self, reps: int, lengths: list):
        self.reps, self.slots = reps, len(lengths)
        self.end = np.tile(np.asarray(lengths, dtype=np.float64), reps)
        self.lobe_block, self.lobe_t, self.lobe_amp = [], [], []
        self.rf, self.adc, self.spoil = [], [], []

    def blocks(self, slot: int, reps=None) -> np.ndarray:
        r = np.arange(self.reps) if reps is None else np.asarray(reps, dtype=np.int64)
        return r * self.slots + slot

    def lobes(self, slot: int, corners: np.ndarray, amps, reps=None) -> None:
        """One lobe per repetition in ``slot``: (R, 4) or (4,) corner times, (R, 3) or (3,) amplitudes."""
        blk = self.blocks(slot, reps)
        self.lobe_block.append(blk)
        self.lobe_t.append(np.broadcast_to(np.asarray(corners, dtype=np.float64), (blk.size, 4)))
        a = amps if isinstance(amps, Tensor) else torch.as_tensor(np.asarray(amps, dtype=np.float64))
        self.lobe_amp.append(a.reshape(-1, 3).expand(blk.size, 3) if a.numel() == 3 else a.reshape(-1, 3))

    def train(self, dtype) -> Train:
        L = sum(b.size for b in self.lobe_block)
        return Train(end=self.end,
                     lobe_block=np.concatenate(self.lobe_block) if L else np.zeros(0, np.int64),
                     lobe_t=np.concatenate(self.lobe_t) if L else np.zeros((0, 4)),
                     lobe_amp=torch.cat([a.to(dtype) for a in self.lobe_amp]) if L else torch.zeros(0, 3, dtype=dtype),
                     rf=self.rf, adc=self.adc, spoil=np.concatenate(self.spoil).tolist() if self.spoil else [])


def _lobe_of(g: Trapezoid) -> tuple[np.ndarray, Tensor]:
    return g.corner_times(), g.amplitude


def _phase_encodes(pe_shape, fov: float) -> tuple[np.ndarray, np.ndarray]:
    """(ky, kz) of every line, the last phase-encode axis fastest; sample i of an axis of m at (i - m//2) / fov."""
    grids = np.meshgrid(*(np.arange(m) for m in pe_shape), indexing="ij")
    k = [((g - m // 2) / fov).reshape(-1) for g, m in zip(grids, pe_shape)]
    ky = k[-1]
    kz = k[0] if len(k) == 2 else np.zeros_like(ky)
    return ky, kz


# ---- GRE


def gre_cartesian(
    shape: tuple[int, ...],
    fov: float = 0.22,
    flip_deg=15.0,
    tr=10e-3,
    te: float | None = None,
    dwell: float = 10e-6,
    rf_duration: float | None = None,
    rf_phase: float = 0.0,
    n_dummy: int = 0,
    dtype=torch.float64,
    *,
    spoiling: str = "ideal",
    slice_thickness: float | None = None,
    slice_offset: float = 0.0,
    ramps: bool = False,
    system: System | None = None,
    spoiler_cycles: float = 4.0,
    rf_spoil_increment: float = 117.0,
    time_bw_product: float = 4.0,
    apodization: float = 0.5,
    rf_steps: int | None = None,
) -> Sequence:
    """Cartesian spoiled GRE over ``(ny, nx)`` or ``(nz, ny, nx)``: one x line per TR, the last phase-encode axis
    fastest, so the ADC samples reshape to ``shape``; sample i of an axis of n sits at k = (i - n/2) / fov. With the
    defaults it is ideal (hard pulse, instantaneous gradients, ideal spoiling); ``spoiling`` ("gradient", "rf"),
    ``slice_thickness`` (a windowed-sinc slab) or ``ramps`` make it realistic. ``n_dummy`` TRs without sampling
    come first. ``flip_deg`` and ``tr`` may require grad."""
    if spoiling not in SPOILING:
        raise ValueError(f"spoiling must be one of {SPOILING}, not {spoiling!r}")
    if spoiling == "ideal" and slice_thickness is None and not ramps:
        return _gre_ideal(shape, fov, flip_deg, tr, te, dwell, 20e-6 if rf_duration is None else rf_duration,
                          rf_phase, n_dummy, dtype, rf_steps)
    return _gre_realistic(shape, fov, flip_deg, tr, te, dwell, rf_duration, rf_phase, n_dummy, dtype, spoiling,
                          slice_thickness, slice_offset, ramps, system or System(), spoiler_cycles,
                          rf_spoil_increment, time_bw_product, apodization, rf_steps)


def gre_2d(n: int = 64, fov: float = 0.22, **kwargs) -> Sequence:
    return gre_cartesian((n, n), fov, **kwargs)


def gre_3d(n: int = 32, fov: float = 0.22, **kwargs) -> Sequence:
    return gre_cartesian((n, n, n), fov, **kwargs)


def _gre_ideal(shape, fov, flip_deg, tr, te, dwell, rf_len, rf_phase, n_dummy, dtype, n_steps) -> Sequence:
    """The ideal GRE as plain segments, every TR at once: the hard pulse in equal steps, one prephasing segment,
    the readout (half a dwell, n - 1 dwells, half a dwell; the first n segments end on samples), one spoiled fill
    segment. A dummy TR is the pulse and one spoiled segment."""
    *pe_shape, nx = shape
    t_read = nx * dwell
    t_centre = (nx / 2 + 0.5) * dwell  # readout start -> the k = 0 sample
    t_pre = 1e-3 if te is None else te - rf_len / 2 - t_centre
    if t_pre <= 0:
        raise ValueError(f"TE too short: prephaser would last {t_pre * 1e3:.3f} ms")
    t_fill = tr - rf_len - t_pre - t_read
    if float(t_fill) <= 0:
        raise ValueError(f"TR too short by {-float(t_fill) * 1e3:.3f} ms")
    t = lambda x: torch.as_tensor(x, dtype=dtype)
    k_rf = hard_pulse_steps(t(rf_len)) if n_steps is None else n_steps
    ky, kz = _phase_encodes(pe_shape, fov)
    lines = ky.size
    rf_row = torch.stack([torch.cos(t(rf_phase)), torch.sin(t(rf_phase))]) * (t(flip_deg) / 360.0 / t(rf_len))
    step = t(rf_len) / k_rf
    half = t(0.5 * dwell).reshape(1)
    # one imaging TR: k_rf pulse steps, prephaser, nx + 1 readout segments, fill
    per_tr = k_rf + 1 + (nx + 1) + 1
    dt_tr = torch.cat([step.expand(k_rf), t(t_pre).reshape(1), half, t(dwell).expand(nx - 1), half,
                       torch.as_tensor(t_fill, dtype=dtype).reshape(1)])
    rf_tr = torch.cat([rf_row[None].expand(k_rf, 2), torch.zeros(per_tr - k_rf, 2, dtype=dtype)])
    grad = torch.zeros(lines, per_tr, 3, dtype=dtype)
    grad[:, k_rf, 0] = -(nx / 2 + 0.5) / fov / t_pre
    grad[:, k_rf, 1] = torch.as_tensor(ky / t_pre, dtype=dtype)
    grad[:, k_rf, 2] = torch.as_tensor(kz / t_pre, dtype=dtype)
    grad[:, k_rf + 1 : k_rf + 2 + nx, 0] = (1.0 / fov) / dwell
    flags = np.zeros((3, per_tr), dtype=bool)  # is_rf, adc, spoil
    flags[0, :k_rf] = True
    flags[1, k_rf + 1 : k_rf + 1 + nx] = True
    flags[2, -1] = True
    # a dummy TR: the pulse and one spoiled segment of the rest of the TR
    dt_d = torch.cat([step.expand(k_rf), (tr - t(rf_len)).reshape(1)])
    flags_d = np.zeros((3, k_rf + 1), dtype=bool)
    flags_d[0, :k_rf] = True
    flags_d[2, -1] = True
    dt = torch.cat([dt_d.repeat(n_dummy), dt_tr.repeat(lines)])
    rf = torch.cat([rf_tr[: k_rf + 1].repeat(n_dummy, 1), rf_tr.repeat(lines, 1)])
    grad = torch.cat([torch.zeros(n_dummy * (k_rf + 1), 3, dtype=dtype), grad.reshape(-1, 3)])
    fl = np.concatenate([np.tile(flags_d, n_dummy), np.tile(flags, lines)], axis=1)
    return Sequence(dt=dt, rf=rf, grad=grad, adc_phase=torch.zeros(dt.shape[0], dtype=dtype),
                    is_rf=fl[0].tolist(), adc=fl[1].tolist(), spoil=fl[2].tolist(),
                    meta={"system": System(), "kspace_shape": tuple(shape), "fov": fov,
                          "te": rf_len / 2 + t_pre + t_centre, "tr": tr, "dwell": dwell})


def _gre_realistic(shape, fov, flip_deg, tr, te, dwell, rf_duration, rf_phase, n_dummy, dtype, spoiling,
                   slice_thickness, slice_offset, ramps, sysm, spoiler_cycles, rf_spoil_increment, time_bw_product,
                   apodization, n_steps=None) -> Sequence:
    """Blocks per TR: the excitation (its RF turned by the TR's spoiling phase), the prephaser (readout prephase,
    phase encodes and the slice rephase, in one lobe), the readout, the spoiler lobe (unless spoiling is ideal: it
    rewinds the phase encodes and adds ``spoiler_cycles`` across a pixel on x), the fill to TR."""
    *pe_shape, nx = shape
    exc = _excitation(flip_deg, rf_duration, slice_thickness, slice_offset, time_bw_product, apodization, sysm,
                      dtype, n_steps)
    ro = _readout(nx, dwell, fov, ramps, sysm, dtype, echo_centred=False)
    ky, kz = _phase_encodes(pe_shape, fov)
    k_max = float(np.maximum(np.abs(ky), np.abs(kz)).max())
    pre_x = -ro.moment_centre
    if te is None:
        t_pre = _shortest_length(max(abs(pre_x), k_max + abs(exc.rephase_z)), sysm) if ramps else 1e-3
    else:
        t_pre = te - exc.tail - ro.to_centre
        if t_pre <= 0:
            raise ValueError(f"TE too short: prephaser would last {t_pre * 1e3:.3f} ms")
    spoiled = spoiling != "ideal"
    spoil_x = spoiler_cycles * nx / fov - (pre_x + ro.moment) if spoiled else 0.0
    t_spoil = _shortest_length(max(abs(spoil_x), k_max), sysm) if spoiled else 0.0
    t_fill = tr - (exc.length + t_pre + ro.length + t_spoil)
    if float(t_fill) < -T_TOL:
        raise ValueError(f"TR too short by {-float(t_fill) * 1e3:.3f} ms")
    fill = isinstance(t_fill, Tensor) or t_fill > T_TOL

    reps = n_dummy + ky.size
    ky_all = np.concatenate([np.zeros(n_dummy), ky])
    kz_all = np.concatenate([np.zeros(n_dummy), kz])
    j = np.arange(reps)
    psi = _rf_spoil_phase(j, rf_spoil_increment) if spoiling == "rf" else np.zeros(reps)
    lengths = [exc.length, t_pre, ro.length] + ([t_spoil] if spoiled else []) + ([float(t_fill)] if fill else [])
    lay = _Layout(reps, lengths)
    for g in exc.lobes:
        lay.lobes(0, *_lobe_of(g))
    phases = torch.as_tensor(psi, dtype=dtype) + torch.as_tensor(rf_phase, dtype=dtype)
    lay.rf.append((lay.blocks(0), exc.pulse.delay, exc.pulse.raster, _turned(exc.pulse.signal.to(dtype), phases)))
    pre_area = np.stack([np.full(reps, pre_x), ky_all, kz_all + exc.rephase_z], axis=1)
    lay.lobes(1, *_lobes_in(pre_area, t_pre, ramps, sysm))
    lay.lobes(2, *_lobe_of(ro.lobe))
    sampled = np.arange(n_dummy, reps)
    lay.adc.append((lay.blocks(2, sampled), ro.adc_delay + (np.arange(nx) + 0.5) * dwell, psi[sampled]))
    if spoiled:
        lay.lobes(3, *_lobes_in(np.stack([np.full(reps, spoil_x), -ky_all, -kz_all], axis=1), t_spoil, ramps,
                                sysm))
    train = lay.train(dtype)
    if fill and isinstance(t_fill, Tensor):
        train.fill_dt = {int(b): t_fill for b in lay.blocks(lay.slots - 1)}
    if spoiling == "ideal":
        train.spoil = lay.blocks(lay.slots - 1).tolist()
    b = SequenceBuilder(dtype=dtype, system=sysm)
    b._record_train(train)
    return b.build(kspace_shape=tuple(shape), fov=fov, te=exc.tail + t_pre + ro.to_centre, tr=tr, dwell=dwell,
                   spoiling=spoiling, slice_thickness=slice_thickness, slice_offset=slice_offset, ramps=ramps,
                   rf_phases=psi.tolist(), system=sysm)


# ---- TSE


def _line_order(n: int, ordering: str) -> list[int]:
    """k-space lines in acquisition order: ``linear`` 0..n-1; ``centric`` from the centre outwards (at equal
    distance the line above the centre first); ``reverse_centric`` from the edges inwards."""
    centre = n // 2
    if ordering == "linear":
        return list(range(n))
    if ordering == "centric":
        return sorted(range(n), key=lambda i: (abs(i - centre), i < centre))
    if ordering == "reverse_centric":
        return sorted(range(n), key=lambda i: (-abs(i - centre), i < centre))
    raise ValueError(f"ordering must be 'linear', 'centric' or 'reverse_centric', not {ordering!r}")


def tse_2d(
    n: int = 64,
    fov: float = 0.22,
    etl: int = 8,
    esp: float | None = None,
    tr: float = 3.0,
    excite_flip_deg=90.0,
    refocus_flip_deg=180.0,
    dwell: float = 10e-6,
    rf_duration: float | None = None,
    ordering: str = "linear",
    n_dummy: int = 0,
    dtype=torch.float64,
    *,
    slice_thickness: float | None = None,
    slice_offset: float = 0.0,
    crusher_cycles: float = 2.0,
    spoiling: str = "ideal",
    spoiler_cycles: float = 4.0,
    ramps: bool = True,
    system: System | None = None,
    time_bw_product: float = 4.0,
    apodization: float = 0.5,
    rf_steps: int | None = None,
    slice_order=None,
    other_rf_raster: float | None = None,
    bound_only_beyond: float | None = None,
    readout_oversampling: int = 1,
    acquired_lines: int | None = None,
) -> Sequence:
    """2D turbo spin echo (CPMG), one slice: a 90 along x, then ``etl`` refocusing pulses along y every ``esp``
    (default: the shortest on twice the gradient raster), crushers on z around each, one phase-encode line per echo
    rewound after it. Shot s acquires lines order[e * shots + s] at its echoes e. ``refocus_flip_deg`` is a scalar
    or ``etl`` values and may require grad, as may ``excite_flip_deg``."""
    if slice_order is not None or bound_only_beyond is not None or other_rf_raster is not None:
        raise NotImplementedError("multi-slice cross-talk (slice_order, bound_only_beyond) is not in this package")
    if acquired_lines is not None:
        raise NotImplementedError("acquired_lines (a partial acquisition) is not in this package")
    if n % etl:
        raise ValueError(f"n = {n} is not a multiple of etl = {etl}")
    if spoiling not in ("ideal", "gradient"):
        raise ValueError(f"spoiling must be 'ideal' or 'gradient', not {spoiling!r}")
    over = int(readout_oversampling)
    if over < 1:
        raise ValueError("readout_oversampling must be a positive integer")
    sysm = system or System()
    shots = n // etl
    if isinstance(refocus_flip_deg, Tensor) and refocus_flip_deg.ndim > 0:
        flips = refocus_flip_deg
    elif isinstance(refocus_flip_deg, (list, tuple)):
        flips = list(refocus_flip_deg)
    else:
        flips = [refocus_flip_deg] * etl
    if len(flips) != etl:
        raise ValueError(f"{len(flips)} refocusing flips for etl = {etl}")
    order = _line_order(n, ordering)
    line = np.asarray([[order[e * shots + s] for e in range(etl)] for s in range(shots)], dtype=np.int64)

    pulse_args = (rf_duration, slice_thickness, slice_offset, time_bw_product, apodization, sysm, dtype, rf_steps)
    exc = _excitation(excite_flip_deg, *pulse_args)
    refs = [_excitation(f, *pulse_args) for f in flips]
    ref = refs[0]
    ro = _readout(n * over, dwell / over, fov * over, ramps, sysm, dtype, echo_centred=True)
    crush = crusher_cycles / (slice_thickness if slice_thickness is not None else fov / n)
    t_pre = _shortest_length(max(ro.moment_centre, abs(crush + exc.rephase_z)), sysm)
    t_pe = _shortest_length(max(n // 2 / fov, crush), sysm)
    # half echo spacings: excitation centre -> first refocusing centre, refocusing centre -> echo, echo -> next
    half = (exc.tail + t_pre + ref.lead, ref.tail + t_pe + ro.to_centre, ro.length - ro.to_centre + t_pe + ref.lead)
    if esp is None:
        esp = ceil_to(2 * max(half), 2 * sysm.grad_raster)
    elif esp < 2 * max(half) - T_TOL:
        raise ValueError(f"echo spacing too short: minimum {2 * max(half) * 1e3:.3f} ms")
    pads = [esp / 2 - h for h in half]

    # slots of one shot
    lengths, slot = [], {}

    def add(name, length, keep=True):
        if keep:
            slot[name] = len(lengths)
            lengths.append(length)

    add("exc", exc.length)
    add("pre", t_pre)
    add("pad0", pads[0], pads[0] > T_TOL)
    for e in range(etl):
        add(("ref", e), refs[e].length)
        add(("pe", e), t_pe)
        add(("pad1", e), pads[1], pads[1] > T_TOL)
        add(("ro", e), ro.length)
        add(("pad2", e), pads[2], pads[2] > T_TOL)
        add(("rw", e), t_pe)
    spoiler = None
    if spoiling == "gradient":
        spoiler = _shortest_lobe(np.array([0.0, 0.0, spoiler_cycles * crush / crusher_cycles]), ramps, sysm)
        add("spoiler", float(spoiler[0][-1]))
    shot_time = 0.0
    for length in lengths:
        shot_time += length
    t_fill = tr - shot_time
    if t_fill < -T_TOL:
        raise ValueError(f"TR too short by {-t_fill * 1e3:.3f} ms")
    add("fill", t_fill, t_fill > T_TOL)

    reps = n_dummy + shots
    lay = _Layout(reps, lengths)
    for g in exc.lobes:
        lay.lobes(slot["exc"], *_lobe_of(g))
    lay.rf.append((lay.blocks(slot["exc"]), exc.pulse.delay, exc.pulse.raster, exc.pulse.signal[None].expand(
        reps, *exc.pulse.signal.shape)))
    lay.lobes(slot["pre"], *_lobes_in(np.array([[ro.moment_centre, 0.0, crush + exc.rephase_z]]), t_pre, ramps,
                                      sysm))
    acquired = np.zeros((reps, etl), dtype=bool)
    acquired[n_dummy:] = True
    lines_all = np.full((reps, etl), n // 2, dtype=np.int64)
    lines_all[n_dummy:] = line
    ky = np.where(acquired, (lines_all - n // 2) / fov, 0.0)
    quarter = torch.as_tensor(math.pi / 2, dtype=dtype).reshape(1)
    for e in range(etl):
        r = refs[e]
        for g in r.lobes:
            lay.lobes(slot[("ref", e)], *_lobe_of(g))
        lay.rf.append((lay.blocks(slot[("ref", e)]), r.pulse.delay, r.pulse.raster,
                       _turned(r.pulse.signal.to(dtype), quarter).expand(reps, *r.pulse.signal.shape)))
        enc = np.stack([np.zeros(reps), ky[:, e], np.full(reps, crush)], axis=1)
        lay.lobes(slot[("pe", e)], *_lobes_in(enc, t_pe, ramps, sysm))
        lay.lobes(slot[("ro", e)], *_lobe_of(ro.lobe))
        sampled = np.flatnonzero(acquired[:, e])
        lay.adc.append((lay.blocks(slot[("ro", e)], sampled),
                        ro.adc_delay + (np.arange(n * over) + 0.5) * (dwell / over), np.zeros(sampled.size)))
        rewind = enc * np.array([1.0, -1.0, 1.0])
        lay.lobes(slot[("rw", e)], *_lobes_in(rewind, t_pe, ramps, sysm))
    if spoiler is not None:
        lay.lobes(slot["spoiler"], *spoiler)
    train = lay.train(dtype)
    if spoiling == "ideal":
        train.spoil = lay.blocks(lay.slots - 1).tolist()
    b = SequenceBuilder(dtype=dtype, system=sysm)
    b._record_train(train)

    kspace_index = [i for s in range(shots) for e in range(etl)
                    for i in range(line[s, e] * n * over, (line[s, e] + 1) * n * over)]
    echo_of = [0] * n
    for s in range(shots):
        for e in range(etl):
            echo_of[line[s, e]] = e
    extra = {} if over == 1 else {"readout_oversampling": over}
    return b.build(matrix=(n, n * over), kspace_index=torch.tensor(kspace_index), fov=fov, dwell=dwell / over, tr=tr,
                   esp=esp, etl=etl, te=(echo_of[n // 2] + 1) * esp, echo_times=[(e + 1) * esp for e in range(etl)],
                   line_echo=echo_of, ordering=ordering, slice_thickness=slice_thickness, ramps=ramps, system=sysm,
                   **extra)
