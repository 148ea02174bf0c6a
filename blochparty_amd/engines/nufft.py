"""NUFFT readouts (``simulate(..., readouts="nufft")``): a long readout with a constant gradient is a type-1
nonuniform FFT of the spins' transverse magnetization, evaluated at exactly the readout's samples.

Reference: ``src/blochparty/engines/gpu/driver.py:65-106`` (which readouts), ``:236-267`` (spins sorted along the
readout's phase), ``:549-580`` (grid and decay terms), ``:703-725`` (FFT, kernel deconvolution, terms recombined),
``cuda/kernels.cu.in:90-140`` (the spreading) and ``docs/PERFORMANCE.md:100-148`` (the method). Forward only, as
there.

For a run of n samples with one dt, gradient and b0_offset, spin s contributes c_s e^{-2 pi i q_s (i + 1)}
e^{-R2_s (i + 1) dt} to sample i, where q_s is its phase per sample in cycles and c_s its coil-weighted Mxy at the
run's start. With j = i - n // 2 the sum over spins is a type-1 NUFFT of the weights c_s e^{-2 pi i q_s (n // 2 + 1)}
at the grid positions x_s = frac(q_s) ng of a 2x-oversampled grid (ng = 2 n):

- the decay is e^{-R2 t} ~ sum_m L_m(t) e^{-R2 tau_m} (Lagrange interpolation at Chebyshev nodes tau_m of the
  readout's span): one grid per term;
- every spin spreads onto the ``width`` cells around x_s with the exponential-of-semicircle kernel
  psi(d) = exp(beta (sqrt(1 - (2 d / width)^2) - 1)), beta = 2.30 width (Barnett, Magland, af Klinteberg 2019);
- the grid's DFT at the n modes j, divided by psi's Fourier transform, recombined over the terms and demodulated,
  is the readout.

On the GPU (``csrc/nufft.hip``, ``csrc/forward.hip``):

- The spins are sorted by frac(q) of the first run's kind, so each workgroup of 256 spins covers a narrow window of
  grid cells (its start wb and width W per workgroup and run kind, computed here).
- The forward kernel stages each spin's (cell, term) columns of its window and contracts them with the coil
  sensitivities on the matrix cores, as it does samples. It writes the workgroup's window of grid columns, then
  jumps over the run in one exact free step.
- ``hb_nufft_grid`` adds, for every grid cell, the columns of the workgroups whose windows hold it, in workgroup
  order (the lists are built here): no atomics.
- ``hb_nufft_signal`` evaluates the DFT at the n modes, deconvolves, recombines the terms and demodulates in float64.

A run kind whose windows exceed ``WMAX`` cells keeps the per-sample path.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import torch

from ..sequence import Sequence, flag_array

MIN_SAMPLES = 16  # readouts shorter than this keep the per-sample path
WMAX = 32  # widest window of grid cells one workgroup may span
MAX_TERMS = 8
BETA_PER_WIDTH = 2.30
TOL = 1e-9


def readout_runs(seq: Sequence) -> list[tuple[int, int]]:
    """(first segment, samples) of the readout runs a NUFFT takes: stretches of >= MIN_SAMPLES ADC segments without
    RF in which every segment has the previous one's dt, gradient and b0_offset (to TOL) and no segment but the last
    is spoiled; of those, only the kind (samples, dt) that occurs most often (ties: the earliest). A readout whose
    first sample ends a half-dwell segment starts its run at the second sample, where the dt becomes constant."""
    adc, rf, sp = (flag_array(f) for f in (seq.adc, seq.is_rf, seq.spoil))
    S = len(adc)
    if S == 0:
        return []
    dt = seq.dt.detach().double().cpu().numpy()
    g = seq.grad.detach().double().cpu().numpy()
    b0 = np.zeros_like(dt) if seq.b0_offset is None else seq.b0_offset.detach().double().cpu().numpy()
    sample = adc & ~rf
    g_scale = TOL * max(float(np.abs(g).max()), 1.0)
    # link[s]: segment s continues the stretch that segment s - 1 is in
    link = np.zeros(S, dtype=bool)
    link[1:] = (sample[1:] & sample[:-1] & ~sp[:-1] & (np.abs(np.diff(dt)) <= TOL * dt[1:])
                & (np.abs(np.diff(g, axis=0)) <= g_scale).all(1) & (np.abs(np.diff(b0)) <= TOL))
    first = np.flatnonzero(sample & ~link)
    stop = np.flatnonzero(~np.append(link[1:], False))  # the last segment of every stretch (and of non-samples)
    last = stop[np.searchsorted(stop, first)]
    length = last + 1 - first
    long_ = length >= MIN_SAMPLES
    first, length = first[long_], length[long_]
    if first.size == 0:
        return []
    kind = np.stack([length, np.round(dt[first] * 1e12)], axis=1)
    kinds, inverse, counts = np.unique(kind, axis=0, return_inverse=True, return_counts=True)
    inverse = inverse.reshape(-1)
    best = counts.max()
    chosen = inverse[np.flatnonzero(counts[inverse] == best)[0]]  # the earliest run of a most common kind
    keep = inverse == chosen
    return list(zip(first[keep].tolist(), length[keep].tolist()))


def es_kernel(z: np.ndarray, beta: float) -> np.ndarray:
    return np.where(np.abs(z) < 1, np.exp(beta * (np.sqrt(np.clip(1 - z * z, 0, None)) - 1)), 0.0)


def es_hat(kp: np.ndarray, ng: int, w: int, beta: float) -> np.ndarray:
    """The kernel's Fourier transform at modes kp of an ng-cell grid, by trapezoid quadrature on 400 w points."""
    t = np.linspace(-w / 2, w / 2, 400 * w + 1)
    return np.trapezoid(es_kernel(2 * t / w, beta)[None, :] * np.cos(2 * np.pi * np.outer(kp, t) / ng), t, axis=1)


def sample_cycles(pos64: np.ndarray, db0: np.ndarray, seq: Sequence, a: int) -> np.ndarray:
    """(N,) float64 phase per sample [cycles] of run kind starting at segment a, as the kernel forms it:
    pos . (g dt) + db0 dt + b0_offset dt."""
    dt = float(seq.dt[a])
    g = seq.grad[a].detach().double().cpu().numpy()
    b0 = 0.0 if seq.b0_offset is None else float(seq.b0_offset[a])
    return pos64 @ (g * dt) + db0 * dt + b0 * dt


@dataclass
class Plan:
    runs: list  # [(first segment, n)]
    width: int
    terms: int
    n: int
    ng: int
    beta: float
    shift: int
    nodes: np.ndarray  # (terms,) tau_m [s]
    lagrange: np.ndarray  # (terms, n)
    hat: np.ndarray  # (n,) at modes j = i - n // 2
    kind: np.ndarray  # (runs,) run kind (distinct dt, gradient, b0_offset)
    kinds: list = field(default_factory=list)  # first segment of each kind
    win: np.ndarray | None = None  # (kinds, n_blocks, 2) int32 window start wb, width W
    csr_off: np.ndarray | None = None  # (kinds, ng + 1) int32
    csr: np.ndarray | None = None  # (entries, 2) int32 (block, local cell), per kind concatenated
    csr_base: np.ndarray | None = None  # (kinds,) int32 first entry of each kind

    @property
    def cols(self) -> int:
        return WMAX * self.terms


def make_plan(seq: Sequence, width: int = 6, terms: int = 3) -> Plan | None:
    """The run selection and the grid parameters (no spin-dependent parts), or None without readout runs."""
    if not 2 <= int(width) <= 16:
        raise ValueError("nufft_width must be 2..16")
    if not 1 <= int(terms) <= MAX_TERMS:
        raise ValueError(f"nufft_terms must be 1..{MAX_TERMS}")
    runs = readout_runs(seq)
    if not runs:
        return None
    n, a0 = runs[0][1], runs[0][0]
    dt = float(seq.dt[a0])
    ng = 2 * n
    beta = BETA_PER_WIDTH * width
    nodes = 0.5 * (dt + n * dt) + 0.5 * (n * dt - dt) * np.cos(np.pi * (2 * np.arange(terms) + 1) / (2 * terms))
    t = (np.arange(n) + 1) * dt
    lag = np.ones((terms, n))
    for a in range(terms):
        for b in range(terms):
            if a != b:
                lag[a] *= (t - nodes[b]) / (nodes[a] - nodes[b])
    kp = np.arange(n) - n // 2
    hat = es_hat(kp, ng, width, beta)
    g = seq.grad.detach().double().cpu().numpy()
    b0 = (np.zeros(len(seq)) if seq.b0_offset is None else seq.b0_offset.detach().double().cpu().numpy())
    keys, kind, firsts = {}, [], []
    for a, _ in runs:
        key = (round(float(seq.dt[a]) * 1e12), *np.round(g[a] * 1e6).astype(np.int64).tolist(), round(b0[a] * 1e9))
        if key not in keys:
            keys[key] = len(firsts)
            firsts.append(a)
        kind.append(keys[key])
    return Plan(runs, int(width), int(terms), n, ng, beta, n // 2 + 1, nodes, lag, hat, np.asarray(kind, np.int32),
                firsts)


def sort_order(plan: Plan, pos64: np.ndarray, db0: np.ndarray, seq: Sequence) -> np.ndarray:
    """The permutation of the spins sorted by the first kind's frac(q) (a workgroup then spans few grid cells)."""
    q = sample_cycles(pos64, db0, seq, plan.kinds[0])
    return np.argsort(q - np.floor(q), kind="stable")


def windows(plan: Plan, pos64: np.ndarray, db0: np.ndarray, seq: Sequence, n: int, n_pad: int) -> bool:
    """Every workgroup's window (start, width incl. the kernel's cells) for every run kind, and the gather lists;
    False when a window exceeds WMAX (the caller keeps the per-sample path)."""
    n_blocks = n_pad // 256
    w = plan.width
    win = np.zeros((len(plan.kinds), n_blocks, 2), dtype=np.int32)
    offs, ents, bases = [], [], []
    total = 0
    for k, a in enumerate(plan.kinds):
        q = sample_cycles(pos64, db0, seq, a)
        x = (q - np.floor(q)) * plan.ng
        base = np.floor(x - 0.5 * w).astype(np.int64) + 1
        pad = np.full(n_pad - n, base[-1] if n else 0, dtype=np.int64)  # padding spins: no weight, any cell
        b = np.concatenate([base, pad]).reshape(n_blocks, 256)
        # no margin: the kernel takes every window cell with |z| < 1, and a support cell that float64 rounding
        # could put past the window's edge has |z| ~ 1, where the kernel is e^-beta (1e-6 at width 6)
        lo, hi = b.min(1), b.max(1)
        width = hi - lo + w
        if width.max() > WMAX:
            return False
        win[k, :, 0], win[k, :, 1] = lo, width
        # cell u (mod ng) <- (block, local cell) for every block whose window holds it, blocks in order
        lists = [[] for _ in range(plan.ng)]
        for blk in range(n_blocks):
            for lc in range(int(width[blk])):
                lists[(int(lo[blk]) + lc) % plan.ng].append((blk, lc))
        off = np.zeros(plan.ng + 1, dtype=np.int32)
        off[1:] = np.cumsum([len(x) for x in lists])
        offs.append(off)
        ents.append(np.asarray([e for x in lists for e in x], dtype=np.int32).reshape(-1, 2))
        bases.append(total)
        total += int(off[-1])
    plan.win, plan.csr_off = win, np.stack(offs)
    plan.csr, plan.csr_base = np.concatenate(ents), np.asarray(bases, dtype=np.int32)
    return True


def device_params(plan: Plan) -> np.ndarray:
    """The float32 parameter block the kernels read: ng, width, terms, beta, shift, cols, then tau_m (8), then
    2 / width and beta log2(e) (the kernel's z per cell and its exponent in base 2)."""
    p = np.zeros(8 + MAX_TERMS, dtype=np.float32)
    p[:6] = [plan.ng, plan.width, plan.terms, plan.beta, plan.shift, plan.cols]
    p[6 : 6 + plan.terms] = plan.nodes
    p[6 + MAX_TERMS] = 2.0 / plan.width
    p[7 + MAX_TERMS] = plan.beta / math.log(2.0)
    return p


def dft_tables(plan: Plan) -> tuple[np.ndarray, np.ndarray]:
    """(n, ng, 2) float64 twiddles e^{-2 pi i j u / ng} at the modes j = i - n // 2, and the (terms, n) float64
    Lagrange weights divided by the kernel's transform."""
    j = np.arange(plan.n) - plan.n // 2
    ph = -2 * np.pi * np.outer(j, np.arange(plan.ng)) / plan.ng
    tw = np.stack([np.cos(ph), np.sin(ph)], -1)
    return np.ascontiguousarray(tw), np.ascontiguousarray(plan.lagrange / plan.hat[None, :])


def to_device(plan: Plan, dev) -> dict:
    tw, lw = dft_tables(plan)
    return dict(
        nu=torch.from_numpy(device_params(plan)).to(dev),
        nu_win=torch.from_numpy(np.ascontiguousarray(plan.win)).to(dev),
        nu_csr_off=torch.from_numpy(np.ascontiguousarray(plan.csr_off)).to(dev),
        nu_csr=torch.from_numpy(np.ascontiguousarray(plan.csr)).to(dev),
        nu_csr_base=torch.from_numpy(plan.csr_base).to(dev),
        nu_tw=torch.from_numpy(tw).to(dev),
        nu_lw=torch.from_numpy(lw).to(dev),
    )


__all__ = ["Plan", "make_plan", "readout_runs", "sort_order", "windows", "to_device", "WMAX", "MIN_SAMPLES"]
