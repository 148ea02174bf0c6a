"""Pulseq import: a ``.seq`` file (format 1.4 or 1.5) -> ``Sequence``, parsed here (no pypulseq).

Reference interface: ``src/blochparty/sequences/pulseq/read.py:53`` (``read_pulseq``: its arguments, conventions and
segment rules; the reference reads through pypulseq, which this package does not use) and the conventions of
``sequences/pulseq/__init__.py``. Format: the Pulseq specification, sections [VERSION], [DEFINITIONS], [BLOCKS],
[RF], [GRADIENTS], [TRAP], [ADC], [EXTENSIONS], [SHAPES].

Segments of a block: its knots are the block start and end, every gradient corner (trapezoid corners, arbitrary
waveform samples and their first / last points), the ``ramp_raster`` points cutting each sloped gradient piece into
equal parts, the RF raster edges of a pulse on the RF raster (else its time points, sloped pieces cut at
``rf_raster``, every piece at ``max_rf_step``), the RF centre, and every ADC sample time; knots within 1e-11 s of
the previous one are one. A segment's gradient is the piecewise-linear waveform at its midpoint (the exact time
average: every corner is a knot); its RF is the raster sample under its midpoint (or the interpolated waveform),
times exp(i (phase + 2 pi f t)) with t the sample's time from the pulse start; an ADC sample ends its segment.

Conventions (as the reference's): units are Pulseq's (Hz, Hz/m, s). ``convention="physical"`` (default): a frequency
offset f excites the spins at +f (slice position f / G), so the field is the complex conjugate of the Pulseq
waveform, rf_x + i rf_y = conj(signal exp(i phase(t))), and the receiver phase is psi = -(phase_offset + 2 pi f t_i
+ modulation_i), t_i = (i + 1/2) dwell: a real pulse tips +z to +y, Pulseq RF phase phi gives Mxy = i e^{-i phi}
sin(flip), and equal RF and ADC phases cancel. ``convention="mrzero"`` takes the phases as written (the mirror
image). Label, trigger and soft-delay extensions are ignored; any other extension raises ``NotImplementedError``.
"""

from __future__ import annotations

import math
import re
from pathlib import Path

import numpy as np
import torch

from .builder import MAX_RF_STEP
from .sequence import T_TOL, Sequence

IGNORED_EXTENSIONS = {"LABELSET", "LABELINC", "TRIGGERS", "DELAYS"}
GAMMA_DEFAULT = 42.576e6  # [Hz/T] for ppm offsets when the file names no B0 / gamma
B0_DEFAULT = 3.0


def decompress_shape(data: np.ndarray, num_samples: int) -> np.ndarray:
    """A [SHAPES] entry's samples: stored as they are when there are ``num_samples`` of them; otherwise the
    run-length-coded derivative (a value written twice is followed by how many more times it repeats), summed."""
    data = np.asarray(data, dtype=np.float64)
    if len(data) == num_samples:
        return data
    out = np.empty(num_samples)
    i = o = 0
    n = len(data)
    while i < n:
        v = data[i]
        if i + 2 < n and data[i + 1] == v:  # v, v, r: v repeated r + 2 times
            r = int(round(data[i + 2])) + 2
            out[o : o + r] = v
            o += r
            i += 3
        else:
            out[o] = v
            o += 1
            i += 1
    if o != num_samples:
        raise ValueError(f"shape decompresses to {o} samples, expected {num_samples}")
    return np.cumsum(out)


def _sections(text: str) -> dict[str, list[str]]:
    """Section name -> its non-comment, non-empty lines."""
    out: dict[str, list[str]] = {}
    cur = None
    for raw in text.splitlines():
        line = raw.split("#", 1)[0].strip()
        if not line:
            continue
        m = re.fullmatch(r"\[(\w+)\]", line)
        if m:
            cur = m.group(1).upper()
            out.setdefault(cur, [])
        elif cur is not None:
            out[cur].append(line)
    return out


def _numbers(line: str) -> list[float]:
    """The leading numeric fields of a line (format 1.5 RF rows end with a one-letter use code)."""
    out = []
    for x in line.split():
        try:
            out.append(float(x))
        except ValueError:
            break
    return out


class SeqFile:
    """The parsed events of a ``.seq`` file (times in s, RF in Hz, gradients in Hz/m)."""

    def __init__(self, text: str):
        sec = _sections(text)
        ver = {ln.split()[0].lower(): ln.split()[1] for ln in sec.get("VERSION", [])}
        self.version = (int(ver.get("major", 1)), int(ver.get("minor", 4)), int(ver.get("revision", 0)))
        if self.version < (1, 4, 0):
            raise NotImplementedError(f"Pulseq format {'.'.join(map(str, self.version))}: 1.4 or later is read")
        self.defs: dict[str, object] = {}
        for ln in sec.get("DEFINITIONS", []):
            key, *vals = ln.split()
            try:
                nums = [float(v) for v in vals]
                self.defs[key] = nums[0] if len(nums) == 1 else nums
            except ValueError:
                self.defs[key] = " ".join(vals)
        d = self.defs
        self.grad_raster = float(d.get("GradientRasterTime", 10e-6))
        self.rf_raster = float(d.get("RadiofrequencyRasterTime", 1e-6))
        self.adc_raster = float(d.get("AdcRasterTime", 100e-9))
        self.block_raster = float(d.get("BlockDurationRaster", 10e-6))
        self.f0 = float(d.get("Gamma", GAMMA_DEFAULT)) * float(d.get("B0", B0_DEFAULT))  # [Hz] for ppm offsets
        self.shapes = self._shapes(sec.get("SHAPES", []))
        v15 = self.version >= (1, 5, 0)
        self.rf = {}
        for ln in sec.get("RF", []):
            x = _numbers(ln)
            if v15:  # id amp mag phase time center delay freq_ppm phase_ppm freq phase [use]
                rid, amp, mag, ph, tid, centre, delay, fppm, pppm, freq, phase = x[:11]
            else:  # id amp mag phase time delay freq phase
                rid, amp, mag, ph, tid, delay, freq, phase = x[:8]
                centre, fppm, pppm = None, 0.0, 0.0
            self.rf[int(rid)] = dict(amp=amp, mag=int(mag), phase_id=int(ph), time=int(tid),
                                     centre=None if centre is None else centre * 1e-6, delay=delay * 1e-6,
                                     freq=freq + fppm * 1e-6 * self.f0, phase=phase + pppm * 1e-6 * self.f0)
        self.grad = {}
        for ln in sec.get("GRADIENTS", []):
            x = _numbers(ln)
            if len(x) >= 7:  # id amp first last amp_shape time_shape delay
                gid, amp, first, last, sid, tid, delay = x[:7]
            else:  # id amp amp_shape time_shape delay (no first / last: the waveform starts and ends at 0)
                gid, amp, sid, tid, delay = x[:5]
                first = last = 0.0
            self.grad[int(gid)] = dict(kind="arb", amp=amp, first=first, last=last, shape=int(sid), time=int(tid),
                                       delay=delay * 1e-6)
        for ln in sec.get("TRAP", []):
            gid, amp, rise, flat, fall, delay = _numbers(ln)[:6]
            self.grad[int(gid)] = dict(kind="trap", amp=amp, rise=rise * 1e-6, flat=flat * 1e-6, fall=fall * 1e-6,
                                       delay=delay * 1e-6)
        self.adc = {}
        for ln in sec.get("ADC", []):
            x = _numbers(ln)
            if v15 and len(x) >= 9:  # id num dwell delay freq_ppm phase_ppm freq phase phase_shape
                aid, num, dwell, delay, fppm, pppm, freq, phase, mod = x[:9]
            else:  # id num dwell delay freq phase
                aid, num, dwell, delay, freq, phase = x[:6]
                fppm = pppm = mod = 0.0
            self.adc[int(aid)] = dict(num=int(num), dwell=dwell * 1e-9, delay=delay * 1e-6,
                                      freq=freq + fppm * 1e-6 * self.f0, phase=phase + pppm * 1e-6 * self.f0,
                                      mod=int(mod))
        self.ext_types: dict[int, str] = {}
        self.ext_lists: dict[int, tuple[int, int, int]] = {}
        for ln in sec.get("EXTENSIONS", []):
            parts = ln.split()
            if parts[0].lower() == "extension":
                self.ext_types[int(parts[2])] = parts[1].upper()
            elif len(parts) >= 4 and parts[0].lstrip("-").isdigit():
                lid, typ, ref, nxt = (int(v) for v in parts[:4])
                self.ext_lists[lid] = (typ, ref, nxt)
        self.blocks = [tuple(int(v) for v in ln.split()[:8]) for ln in sec.get("BLOCKS", [])]

    @staticmethod
    def _shapes(lines: list[str]) -> dict[int, np.ndarray]:
        out, i = {}, 0
        while i < len(lines):
            head = lines[i].split()
            if head[0].lower() != "shape_id":
                raise ValueError(f"[SHAPES]: expected shape_id, got {lines[i]!r}")
            sid = int(head[1])
            num = int(lines[i + 1].split()[1])
            i += 2
            vals = []
            while i < len(lines) and lines[i].split()[0].lower() != "shape_id":
                vals.append(float(lines[i]))
                i += 1
            out[sid] = decompress_shape(np.array(vals), num)
        return out

    def extensions_of(self, lid: int) -> list[str]:
        names, seen = [], set()
        while lid and lid not in seen:
            seen.add(lid)
            typ, _, nxt = self.ext_lists[lid]
            names.append(self.ext_types.get(typ, f"#{typ}"))
            lid = nxt
        return names

    def grad_waveform(self, gid: int) -> tuple[np.ndarray, np.ndarray]:
        """(times from the block start, values) of gradient event ``gid`` as a piecewise-linear waveform."""
        g = self.grad[gid]
        if g["kind"] == "trap":
            t = np.cumsum([0.0, g["rise"], g["flat"], g["fall"]])
            v = g["amp"] * np.array([0.0, 1.0, 1.0, 0.0])
        else:
            w = g["amp"] * self.shapes[g["shape"]]
            if g["time"] == 0:  # samples at the raster cells' centres, first / last at the waveform's ends
                tc = (np.arange(len(w)) + 0.5) * self.grad_raster
                t = np.concatenate([[0.0], tc, [len(w) * self.grad_raster]])
                v = np.concatenate([[g["first"]], w, [g["last"]]])
            else:  # an explicit time shape (in raster units), e.g. an extended trapezoid
                t = self.shapes[g["time"]] * self.grad_raster
                v = w
        keep = np.concatenate([[True], np.diff(t) > T_TOL])
        return g["delay"] + t[keep], v[keep]

    def rf_waveform(self, rid: int) -> dict:
        """The pulse's complex samples and their times from the pulse start, whether they sit on the RF raster, its
        centre, its delay, frequency and phase."""
        r = self.rf[rid]
        mag = self.shapes[r["mag"]]
        phase = self.shapes[r["phase_id"]] if r["phase_id"] else np.zeros_like(mag)
        sig = r["amp"] * mag * np.exp(2j * math.pi * phase)
        if r["time"] == 0:
            t = (np.arange(len(sig)) + 0.5) * self.rf_raster
            on_raster = True
        else:
            t = self.shapes[r["time"]] * self.rf_raster
            on_raster = bool(np.allclose(t, (np.arange(len(t)) + 0.5) * self.rf_raster, atol=T_TOL, rtol=0))
        centre = r["centre"]
        if centre is None:  # format 1.4: the middle of the samples at the peak magnitude
            peak = np.abs(sig).max()
            at = t[np.abs(sig) >= peak * (1 - 1e-5)]
            lo = at.min() - (0.5 * self.rf_raster if on_raster else 0.0)
            hi = at.max() + (0.5 * self.rf_raster if on_raster else 0.0)
            centre = 0.5 * (lo + hi)
        return dict(sig=sig, t=t, on_raster=on_raster, centre=centre, delay=r["delay"], freq=r["freq"],
                    phase=r["phase"])


def _cuts(t: np.ndarray, sloped: np.ndarray, step: float) -> np.ndarray:
    """Points cutting every sloped interval of knots ``t`` into equal pieces no longer than ``step``."""
    a, b = t[:-1][sloped], t[1:][sloped]
    parts = np.ceil((b - a) / step - 1e-9).astype(np.int64)
    more = np.maximum(parts - 1, 0)
    if more.sum() == 0:
        return np.zeros(0)
    owner = np.repeat(np.arange(len(a)), more)
    j = np.arange(more.sum()) - np.repeat(np.cumsum(more) - more, more) + 1
    return a[owner] + (b - a)[owner] * j / parts[owner]


def read_pulseq(path_or_text, *, dtype=torch.float64, device=None, ramp_raster: float | None = None,
                rf_raster: float | None = None, max_rf_step: float | None = None, convention: str = "physical",
                **meta) -> Sequence:
    """A ``.seq`` file (a path, or its text) as a ``Sequence`` (the module's rules and conventions).

    ramp_raster: the longest piece a sloped gradient section is cut into (default the gradient raster).
    rf_raster: likewise for time-shaped RF that varies or carries a frequency offset (default the RF raster);
        pulses on the RF raster are stepped sample by sample.
    max_rf_step: the longest piece of a time-shaped pulse (default ``MAX_RF_STEP``).
    convention: "physical" (default) or "mrzero".
    meta: extra entries for ``Sequence.meta``.
    """
    if convention not in ("physical", "mrzero"):
        raise ValueError(f"convention must be 'physical' or 'mrzero', not {convention!r}")
    text = path_or_text
    if isinstance(path_or_text, Path) or (isinstance(path_or_text, str) and "\n" not in path_or_text):
        text = Path(path_or_text).read_text()
    f = SeqFile(text)
    sign = -1.0 if convention == "physical" else 1.0
    ramp_raster = ramp_raster or f.grad_raster
    rf_step = rf_raster or f.rf_raster
    max_step = max_rf_step or MAX_RF_STEP
    dts, grads, rfs, rf_on, adc_on, psis, counts, spans = [], [], [], [], [], [], [], []
    n_seg = 0
    for blk in f.blocks:
        bid, dur_units, rid, gx, gy, gz, aid = blk[:7]
        ext = blk[7] if len(blk) > 7 else 0
        if ext:
            unknown = [e for e in f.extensions_of(ext) if e not in IGNORED_EXTENSIONS]
            if unknown:
                raise NotImplementedError(f"block {bid}: unsupported Pulseq extension(s) {sorted(set(unknown))}")
        dur = dur_units * f.block_raster
        if dur <= T_TOL:
            spans.append((n_seg, n_seg))
            continue
        knots = [np.array([0.0, dur])]
        shapes = []
        for gid in (gx, gy, gz):
            if not gid:
                shapes.append(None)
                continue
            t, v = f.grad_waveform(gid)
            shapes.append((t, v))
            knots += [t, _cuts(t, np.diff(v) != 0, ramp_raster)]
        pulse = f.rf_waveform(rid) if rid else None
        if pulse is not None:
            tr = pulse["t"]
            if pulse["on_raster"]:
                edges = np.arange(len(tr) + 1) * f.rf_raster
            else:
                sloped = (np.diff(pulse["sig"]) != 0) | (pulse["freq"] != 0)
                edges = np.sort(np.concatenate([tr, _cuts(tr, sloped, rf_step)]))
                edges = np.concatenate([edges, _cuts(edges, np.ones(len(edges) - 1, dtype=bool), max_step)])
            knots += [pulse["delay"] + edges, np.array([pulse["delay"] + pulse["centre"]])]
            rf_lo, rf_hi = pulse["delay"] + edges.min(), pulse["delay"] + edges.max()
        adc = f.adc[aid] if aid else None
        if adc is not None:
            t_adc = adc["delay"] + (np.arange(adc["num"]) + 0.5) * adc["dwell"]
            knots.append(t_adc)
        t = np.sort(np.concatenate(knots))
        t = t[np.concatenate([[True], np.diff(t) > T_TOL])]
        if t[-1] > dur + T_TOL:
            raise ValueError(f"block {bid}: events last {t[-1]:.9g} s, longer than the block ({dur:.9g} s)")
        t = np.append(t[t < dur - T_TOL], dur)
        if t[0] > T_TOL:
            t = np.concatenate([[0.0], t])
        mid = 0.5 * (t[1:] + t[:-1])
        n = len(mid)
        g = np.zeros((n, 3))
        for axis, shape in enumerate(shapes):
            if shape is not None:
                g[:, axis] = np.interp(mid, shape[0], shape[1], left=0.0, right=0.0)
        b1 = np.zeros(n, dtype=complex)
        is_rf = np.zeros(n, dtype=bool)
        if pulse is not None:
            is_rf = (mid > rf_lo) & (mid < rf_hi)
            since = mid[is_rf] - pulse["delay"]
            sig, tr = pulse["sig"], pulse["t"]
            if pulse["on_raster"]:
                k = np.clip(np.floor(since / f.rf_raster).astype(np.int64), 0, len(sig) - 1)
                ts, vals = tr[k], sig[k]
            else:
                ts = since
                vals = np.interp(since, tr, sig.real) + 1j * np.interp(since, tr, sig.imag)
            b1[is_rf] = vals * np.exp(1j * (pulse["phase"] + 2 * math.pi * pulse["freq"] * ts))
        if sign < 0:
            b1 = b1.conj()
        is_adc = np.zeros(n, dtype=bool)
        psi = np.zeros(n)
        if adc is not None:
            at = np.searchsorted(t, t_adc - T_TOL)
            if np.any(at >= len(t)) or np.any(np.abs(t[np.minimum(at, len(t) - 1)] - t_adc) > T_TOL):
                raise ValueError(f"block {bid}: ADC samples off the block's knots")
            ph = adc["phase"] + 2 * math.pi * adc["freq"] * (t_adc - adc["delay"])
            if adc["mod"]:
                ph = ph + f.shapes[adc["mod"]]
            is_adc[at - 1] = True
            psi[at - 1] = sign * ph
            counts.append(adc["num"])
        dts.append(np.diff(t))
        grads.append(g)
        rfs.append(b1)
        rf_on.append(is_rf)
        adc_on.append(is_adc)
        psis.append(psi)
        spans.append((n_seg, n_seg + n))
        n_seg += n
    cat = lambda parts: torch.as_tensor(np.concatenate(parts), dtype=dtype, device=device)
    b1_all = np.concatenate(rfs) if rfs else np.zeros(0, dtype=complex)
    out = Sequence(
        dt=cat(dts), rf=torch.as_tensor(np.stack([b1_all.real, b1_all.imag], -1).reshape(-1, 2), dtype=dtype,
                                        device=device),
        grad=cat(grads).reshape(-1, 3), adc_phase=cat(psis), is_rf=np.concatenate(rf_on).tolist(),
        adc=np.concatenate(adc_on).tolist(), spoil=[False] * n_seg,
        meta={"adc_per_readout": counts, "block_segments": spans, "definitions": dict(f.defs),
              "convention": convention, "pulseq_version": f.version},
    )
    out.meta.update(meta)
    return out


__all__ = ["SeqFile", "decompress_shape", "read_pulseq"]
