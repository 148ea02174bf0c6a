"""Ready-made sequences that feed the GPU path: spoiled Cartesian GRE (2D / 3D, ideal or realistic) and 2D TSE.

Reference: ``src/blochparty/sequences/library.py`` — ``gre_cartesian`` (:43-138), ``gre_2d`` / ``gre_3d``
(:143-150), the realistic blocks (:188-380) and ``tse_2d`` (:392-640). The timing rules are kept exactly (the
segment tables are compared with the reference's as golden fixtures); multi-slice cross-talk (``slice_order``,
``bound_only_beyond``) is not part of this package and raises ``NotImplementedError``.
"""

from __future__ import annotations

import itertools
import math
from dataclasses import replace
from typing import NamedTuple

import torch
from torch import Tensor

from .builder import (
    ADC,
    RFPulse,
    SequenceBuilder,
    System,
    Trapezoid,
    axis_vector,
    block_duration,
    block_pulse,
    ceil_to,
    sinc_pulse,
    trapezoid,
)
from .sequence import T_TOL, Sequence

SPOILING = ("ideal", "gradient", "rf")


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
    """Cartesian spoiled GRE, ``(ny, nx)`` or ``(nz, ny, nx)``: one x line per TR, ky fastest, so the ADC samples
    reshape to ``shape``; sample i along an axis of n sits at k = (i - n/2) / fov. Ideal by default (hard pulse,
    instantaneous gradients, ideal spoiling); ``spoiling``, ``slice_thickness`` or ``ramps`` give the realistic
    build. ``flip_deg`` / ``tr`` may require grad."""
    if spoiling not in SPOILING:
        raise ValueError(f"spoiling must be one of {SPOILING}, not {spoiling!r}")
    if spoiling != "ideal" or slice_thickness is not None or ramps:
        return _gre_realistic(shape, fov, flip_deg, tr, te, dwell, rf_duration, rf_phase, n_dummy, dtype, spoiling,
                              slice_thickness, slice_offset, ramps, system or System(), spoiler_cycles,
                              rf_spoil_increment, time_bw_product, apodization, rf_steps)
    rf_duration = 20e-6 if rf_duration is None else rf_duration
    *pe_shape, nx = shape
    dk = 1.0 / fov
    t_read = nx * dwell
    t_echo = (nx / 2 + 0.5) * dwell  # readout start -> the k = 0 sample
    t_pre = 1e-3 if te is None else te - rf_duration / 2 - t_echo
    if t_pre <= 0:
        raise ValueError(f"TE too short: prephaser would last {t_pre * 1e3:.3f} ms")
    t_fill = tr - rf_duration - t_pre - t_read
    if float(t_fill) <= 0:
        raise ValueError(f"TR too short by {-float(t_fill) * 1e3:.3f} ms")
    b = SequenceBuilder(dtype=dtype)
    for _ in range(n_dummy):
        b.rf(rf_duration, flip_deg=flip_deg, phase=rf_phase, n_steps=rf_steps)
        b.delay(tr - rf_duration).spoil()
    for idx in itertools.product(*(range(m) for m in pe_shape)):
        k_pe = [(i - m // 2) * dk for i, m in zip(idx, pe_shape)]
        ky, kz = k_pe[-1], (k_pe[0] if len(k_pe) == 2 else 0.0)
        b.rf(rf_duration, flip_deg=flip_deg, phase=rf_phase, n_steps=rf_steps)
        b.delay(t_pre, grad=(-(nx / 2 + 0.5) * dk / t_pre, ky / t_pre, kz / t_pre))
        b.readout(nx, dwell, grad=(dk / dwell, 0.0, 0.0))
        b.delay(t_fill).spoil()
    return b.build(kspace_shape=tuple(shape), fov=fov, te=rf_duration / 2 + t_pre + t_echo, tr=tr, dwell=dwell)


def gre_2d(n: int = 64, fov: float = 0.22, **kwargs) -> Sequence:
    return gre_cartesian((n, n), fov, **kwargs)


def gre_3d(n: int = 32, fov: float = 0.22, **kwargs) -> Sequence:
    return gre_cartesian((n, n, n), fov, **kwargs)


class _Excitation(NamedTuple):
    grads: tuple
    rf: RFPulse
    rephase: Tensor  # (3,) undoes the slice-select moment after the pulse centre
    duration: float
    pre: float  # block start -> pulse centre
    post: float  # pulse centre -> block end


def _excitation(flip_deg, rf_duration, slice_thickness, slice_offset, time_bw_product, apodization,
                sysm: System, dtype, n_steps=None) -> _Excitation:
    """Hard pulse, or a windowed sinc with its slice-select gradient when ``slice_thickness`` is set."""
    if slice_thickness is None:
        rf = block_pulse(flip_deg, 20e-6 if rf_duration is None else rf_duration, system=sysm, dtype=dtype,
                         n_steps=n_steps)
        grads, reph = (), torch.zeros(3, dtype=dtype)
    else:
        rf, gz, reph = sinc_pulse(flip_deg, 2e-3 if rf_duration is None else rf_duration,
                                  slice_thickness=slice_thickness, time_bw_product=time_bw_product,
                                  apodization=apodization, slice_offset=slice_offset, system=sysm, dtype=dtype)
        grads = (gz,)
    dur = block_duration(*grads, rf=rf, system=sysm)
    centre = rf.delay + rf.center
    return _Excitation(grads, rf, reph, dur, centre, dur - centre)


def _rotated(rf: RFPulse, phase) -> RFPulse:
    """``rf`` with every sample turned by ``phase`` [rad]."""
    ph = torch.as_tensor(phase, dtype=rf.signal.dtype)
    c, s = torch.cos(ph), torch.sin(ph)
    x, y = rf.signal.unbind(-1)
    return replace(rf, signal=torch.stack([c * x - s * y, s * x + c * y], dim=-1))


class _Readout(NamedTuple):
    grad: Trapezoid
    adc_delay: float
    t_k0: float  # block start -> the k = 0 sample
    area_k0: float  # x moment up to the k = 0 sample
    area: float
    duration: float


def _readout(n: int, dwell: float, fov: float, ramps: bool, sysm: System, dtype, centred: bool) -> _Readout:
    """Readout lobe on x, ADC on the flat top, sample i at kx = (i - n/2) / fov; ``centred`` puts the k = 0 sample
    in the middle of the lobe (spin echoes)."""
    sysm.check_raster(dwell, sysm.adc_raster, "dwell")
    dk = 1.0 / fov
    amp, flat = dk / dwell, n * dwell
    if ramps:
        g = trapezoid("x", amplitude=amp, flat_time=flat, system=sysm, dtype=dtype)
    else:
        g = Trapezoid(axis_vector(amp, "x", dtype), 0.0, flat, 0.0)
    r = g.rise
    g_delay = max(0.0, 0.5 * dwell - r) if centred else 0.0
    g = replace(g, delay=g_delay)
    if centred:
        adc_delay, t_k0, area_k0 = g_delay + r - 0.5 * dwell, g_delay + r + 0.5 * flat, amp * (0.5 * r + 0.5 * flat)
    else:
        adc_delay, t_k0, area_k0 = r, r + (n / 2 + 0.5) * dwell, amp * 0.5 * r + (n / 2 + 0.5) * dk
    return _Readout(g, adc_delay, t_k0, area_k0, amp * (r + flat), g.duration)


def _shortest(area: float, sysm: System) -> float:
    return trapezoid("x", area=area, system=sysm).duration


def _lobe(area, duration: float | None, ramps: bool, sysm: System, dtype) -> Trapezoid:
    """A lobe of (3,) ``area``: a trapezoid of ``duration`` (None: shortest), or without ramps a rectangle."""
    if ramps:
        return trapezoid(area=area, duration=duration, system=sysm, dtype=dtype)
    a = axis_vector(area, None, dtype)
    if duration is None:
        duration = _shortest(float(a.detach().abs().max()), sysm)
    return Trapezoid(a / duration if duration > 0 else a * 0.0, 0.0, duration, 0.0)


def _rf_spoil_phase(j: int, increment_deg: float) -> float:
    """Quadratic RF spoiling: increment j (j + 1) / 2 degrees, wrapped."""
    return math.radians((increment_deg * j * (j + 1) / 2) % 360.0)


def _pad(b: SequenceBuilder, t: float) -> None:
    if t > T_TOL:
        b.delay(t)


def _gre_realistic(shape, fov, flip_deg, tr, te, dwell, rf_duration, rf_phase, n_dummy, dtype, spoiling,
                   slice_thickness, slice_offset, ramps, sysm, spoiler_cycles, rf_spoil_increment, time_bw_product,
                   apodization, n_steps=None) -> Sequence:
    *pe_shape, nx = shape
    dk = 1.0 / fov
    exc = _excitation(flip_deg, rf_duration, slice_thickness, slice_offset, time_bw_product, apodization, sysm,
                      dtype, n_steps)
    ro = _readout(nx, dwell, fov, ramps, sysm, dtype, centred=False)
    lines = []
    for idx in itertools.product(*(range(m) for m in pe_shape)):
        k_pe = [(i - m // 2) * dk for i, m in zip(idx, pe_shape)]
        lines.append((k_pe[-1], k_pe[0] if len(k_pe) == 2 else 0.0))
    k_max = max(max(abs(ky), abs(kz)) for ky, kz in lines)
    pre_x, reph_z = -ro.area_k0, exc.rephase[2]
    if te is None:
        t_pre = _shortest(max(abs(pre_x), k_max + abs(float(reph_z))), sysm) if ramps else 1e-3
    else:
        t_pre = te - exc.post - ro.t_k0
        if t_pre <= 0:
            raise ValueError(f"TE too short: prephaser would last {t_pre * 1e3:.3f} ms")
    spoil_x, t_spoil = 0.0, 0.0
    if spoiling != "ideal":
        spoil_x = spoiler_cycles * nx / fov - (pre_x + ro.area)
        t_spoil = _shortest(max(abs(spoil_x), k_max), sysm)
    t_fill = tr - (exc.duration + t_pre + ro.duration + t_spoil)
    if float(t_fill) < -T_TOL:
        raise ValueError(f"TR too short by {-float(t_fill) * 1e3:.3f} ms")
    b = SequenceBuilder(dtype=dtype, system=sysm)
    rf_phases = []

    def one_tr(j, ky, kz, acquire):
        psi = _rf_spoil_phase(j, rf_spoil_increment) if spoiling == "rf" else 0.0
        rf_phases.append(psi)
        b.block(*exc.grads, rf=_rotated(exc.rf, rf_phase + psi))
        b.block(_lobe((pre_x, ky, kz + reph_z), t_pre, ramps, sysm, dtype))
        b.block(ro.grad, adc=ADC(nx, dwell, ro.adc_delay, psi) if acquire else None, duration=ro.duration)
        if spoiling != "ideal":
            b.block(_lobe((spoil_x, -ky, -kz), t_spoil, ramps, sysm, dtype))
        if isinstance(t_fill, Tensor) or t_fill > T_TOL:
            b.delay(t_fill)
        if spoiling == "ideal":
            b.spoil()

    for j in range(n_dummy):
        one_tr(j, 0.0, 0.0, acquire=False)
    for j, (ky, kz) in enumerate(lines):
        one_tr(n_dummy + j, ky, kz, acquire=True)
    return b.build(kspace_shape=tuple(shape), fov=fov, te=exc.post + t_pre + ro.t_k0, tr=tr, dwell=dwell,
                   spoiling=spoiling, slice_thickness=slice_thickness, slice_offset=slice_offset, ramps=ramps,
                   rf_phases=rf_phases, system=sysm)


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
    """2D turbo spin echo (CPMG), single slice (``library.py:392-640``): excitation along x, ``etl`` refocusing
    pulses along y every ``esp`` with crushers on z, one phase-encode line per echo (rewound after it).
    ``refocus_flip_deg`` is a scalar or ``etl`` values and may require grad."""
    if slice_order is not None or bound_only_beyond is not None or other_rf_raster is not None:
        raise NotImplementedError("multi-slice cross-talk (slice_order, bound_only_beyond) is not in this package")
    if acquired_lines is None and n % etl:
        raise ValueError(f"n = {n} is not a multiple of etl = {etl}")
    if acquired_lines is not None and (acquired_lines < n or ordering != "linear"):
        raise ValueError("acquired_lines needs ordering 'linear' and at least n lines")
    if spoiling not in ("ideal", "gradient"):
        raise ValueError(f"spoiling must be 'ideal' or 'gradient', not {spoiling!r}")
    sysm = system or System()
    dk, n_shots = 1.0 / fov, n // etl
    flips = refocus_flip_deg
    if not isinstance(flips, Tensor) or flips.ndim == 0:
        flips = [flips] * etl if not isinstance(flips, (list, tuple)) else list(flips)
    if len(flips) != etl:
        raise ValueError(f"{len(flips)} refocusing flips for etl = {etl}")
    order = list(range(n))
    if ordering == "centric":
        order = sorted(order, key=lambda i: (abs(i - n // 2), i < n // 2))
    elif ordering == "reverse_centric":
        order = sorted(order, key=lambda i: (-abs(i - n // 2), i < n // 2))
    elif ordering != "linear":
        raise ValueError(f"ordering must be 'linear', 'centric' or 'reverse_centric', not {ordering!r}")
    lines = [[order[e * n_shots + s] for e in range(etl)] for s in range(n_shots)]
    if acquired_lines is not None:
        band = [[] for _ in range(etl)]
        for i in range(n):
            band[min(etl - 1, (i - n // 2 + acquired_lines // 2) * etl // acquired_lines)].append(i)
        n_shots = max(len(x) for x in band)
        lines = [[x[s] if s < len(x) else None for x in band] for s in range(n_shots)]
    args = (rf_duration, slice_thickness, slice_offset, time_bw_product, apodization, sysm, dtype, rf_steps)
    exc = _excitation(excite_flip_deg, *args)
    refs = [_excitation(f, *args) for f in flips]
    ref = refs[0]
    os_ = int(readout_oversampling)
    if os_ < 1:
        raise ValueError("readout_oversampling must be a positive integer")
    ro = _readout(n * os_, dwell / os_, fov * os_, ramps, sysm, dtype, centred=True)
    crush = crusher_cycles / (slice_thickness if slice_thickness is not None else fov / n)
    t_p = _shortest(max(ro.area_k0, abs(crush + float(exc.rephase[2]))), sysm)
    t_c = _shortest(max(n // 2 * dk, crush), sysm)
    h0 = exc.post + t_p + ref.pre
    h1 = ref.post + t_c + ro.t_k0
    h2 = ro.duration - ro.t_k0 + t_c + ref.pre
    if esp is None:
        esp = ceil_to(2 * max(h0, h1, h2), 2 * sysm.grad_raster)
    elif esp < 2 * max(h0, h1, h2) - T_TOL:
        raise ValueError(f"echo spacing too short: minimum {2 * max(h0, h1, h2) * 1e3:.3f} ms")
    b = SequenceBuilder(dtype=dtype, system=sysm)
    kspace_index = []
    for shot in range(n_dummy + n_shots):
        acquire = shot >= n_dummy
        t_start = b.time
        b.block(*exc.grads, rf=exc.rf)
        b.block(_lobe((ro.area_k0, 0.0, crush + exc.rephase[2]), t_p, ramps, sysm, dtype))
        _pad(b, esp / 2 - h0)
        for e, r in enumerate(refs):
            line = lines[shot - n_dummy][e] if acquire else n // 2
            read = acquire and line is not None
            ky = (line - n // 2) * dk if read else 0.0
            b.block(*r.grads, rf=_rotated(r.rf, math.pi / 2))
            b.block(_lobe((0.0, ky, crush), t_c, ramps, sysm, dtype))
            _pad(b, esp / 2 - h1)
            b.block(ro.grad, adc=ADC(n * os_, dwell / os_, ro.adc_delay) if read else None, duration=ro.duration)
            _pad(b, esp / 2 - h2)
            b.block(_lobe((0.0, -ky, crush), t_c, ramps, sysm, dtype))
            if read:
                kspace_index += range(line * n * os_, (line + 1) * n * os_)
        if spoiling == "gradient":
            b.block(_lobe((0.0, 0.0, spoiler_cycles * crush / crusher_cycles), None, ramps, sysm, dtype))
        t_fill = tr - (b.time - t_start)
        if t_fill < -T_TOL:
            raise ValueError(f"TR too short by {-t_fill * 1e3:.3f} ms")
        _pad(b, t_fill)
        if spoiling == "ideal":
            b.spoil()
    line_echo = [0] * n
    for s in range(n_shots):
        for e in range(etl):
            if lines[s][e] is not None:
                line_echo[lines[s][e]] = e
    return b.build(matrix=(n, n * os_), kspace_index=torch.tensor(kspace_index), fov=fov, dwell=dwell / os_, tr=tr,
                   esp=esp, etl=etl, te=(line_echo[n // 2] + 1) * esp, echo_times=[(e + 1) * esp for e in range(etl)],
                   line_echo=line_echo, ordering=ordering, slice_thickness=slice_thickness, ramps=ramps, system=sysm,
                   **({} if os_ == 1 else {"readout_oversampling": os_}))
