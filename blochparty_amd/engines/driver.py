"""The GPU engine's host side: phantom, coils and sequence into device buffers, dispatch in chunks of segments,
readback. The role of the reference's ``_forward`` (``src/blochparty/engines/gpu/driver.py:197-756``), rebuilt for
the HIP kernels' layouts:

- spins as float32 structure-of-arrays (8, n_pad) ``px py pz pd r1 r2 db0 b1`` (``driver.py:320-335``) plus the
  float64 positions' low parts (3, n_pad), for the float64 k.r phase of free segments;
- sensitivities in the matrix-core operand order (n_pad / 64, tiles, 16, 64) (see ``csrc/forward.hip``);
- segments as 128-byte records: float32 dt, rf, g, b0_offset, psi (wrapped mod 2 pi in float64 first,
  ``driver.py:357``), float64 dt, moment g dt and b0_offset dt, flags (RF 1, ADC 2, spoil 4, ``driver.py:375-382``,
  and cache hints), and the rotation-vector factors -2 pi dt (1, rf, g, b0_offset) formed in float64;
- the state as float32 hi + lo, split from float64 to ~48 bits (``driver.py:386-433``).

The signal is summed per workgroup into partial rows and reduced across workgroups in float64 (``hb_signal_reduce``),
so the result does not depend on scheduling. ``prepare_case`` / ``run_forward`` split the upload from the timed GPU
work (the benchmark times ``run_forward`` on resident buffers).
"""

from __future__ import annotations

import functools
import math
from dataclasses import dataclass, field

import numpy as np
import torch

from ..phantom import Phantom
from ..result import SimResult
from ..sequence import Sequence, flag_array, flag_memo
from . import hip, pulses

BLOCK = 256
RF, ADC, SPOIL, SAME_FREE, SAME_RF_DT, NEXT_FREE, NEXT_RF_DT, PULSE, PULSE_SPOIL, RUN, SAME_RUN, TRAIN, MERGED, \
    SPOILED_FREE, NUFFT, XY_DEAD, SAME_FREE_DT, PULSE_BODY = (1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096,
                                                              8192, 16384, 32768, 65536, 131072)
NUFFT_RUNS = 64  # NUFFT runs per dispatch at most (slots of the grid partials)
MIN_RUN = 4  # readout runs shorter than this are stepped segment by segment
MIN_TRAIN = 4  # likewise RF trains
REDUCE_CHUNK = 256  # blocks per first-pass chunk of hb_signal_reduce
SEG_DTYPE = np.dtype([("f", "<f4", (8,)), ("d", "<f8", (5,)), ("flags", "<i4"), ("pad", "<i4", (5,)),
                      ("w", "<f4", (8,))])
assert SEG_DTYPE.itemsize == 128
PARTIAL_BYTES = 1 << 30  # workgroup partial sums of one dispatch at most (one buffer, reused)
DEFAULT_CHUNK = 32768  # segments per dispatch at most (the reference's default ``chunk``)
SAME_TOL = 1e-12  # relative: segments this close share their precession / relaxation factors


def _close(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    return np.abs(a - b) <= SAME_TOL * np.maximum(np.abs(a), np.abs(b))


def _span_index(starts: np.ndarray, lengths: np.ndarray) -> np.ndarray:
    """The indices of the ranges [starts[i], starts[i] + lengths[i]), concatenated in order."""
    lengths = np.maximum(lengths, 0)
    total = int(lengths.sum())
    if total == 0:
        return np.zeros(0, dtype=np.int64)
    offsets = np.cumsum(lengths) - lengths  # where each range begins in the output
    return np.arange(total, dtype=np.int64) + np.repeat(starts - offsets, lengths)


def segment_table(seq: Sequence, starts: list[int], pulses=(), runs: bool = False, nufft=None) -> np.ndarray:
    """The (S,) records of ``seq``; ``starts`` are the dispatch starts (the factor caches restart there); ``pulses``:
    reused pulses (``pulses.reused_pulses``), whose occurrences are applied as maps and not stepped; ``runs``: mark
    readout runs (identical free ADC segments) for the forward kernel's per-run Mz step (not with checkpoints: the
    adjoint replays segment by segment)."""
    S = len(seq)
    dt = seq.dt.detach().double().cpu().numpy()
    rf = seq.rf.detach().double().cpu().numpy()
    g = seq.grad.detach().double().cpu().numpy()
    b0 = np.zeros(S) if seq.b0_offset is None else seq.b0_offset.detach().double().cpu().numpy()
    psi = np.remainder(seq.adc_phase.detach().double().cpu().numpy(), 2 * math.pi)
    is_rf, adc, spoil = flag_array(seq.is_rf), flag_array(seq.adc), flag_array(seq.spoil)
    rec = np.zeros(S, dtype=SEG_DTYPE)
    packed = np.empty((S, 8), dtype=np.float32)
    for c, col in enumerate((dt, rf[:, 0], rf[:, 1], g[:, 0], g[:, 1], g[:, 2], b0, psi)):
        packed[:, c] = col
    rec["f"] = packed
    moments = np.empty((S, 5))
    moments[:, 0] = dt
    moments[:, 1:4] = g * dt[:, None]
    moments[:, 4] = b0 * dt
    rec["d"] = moments
    flags = is_rf * RF | adc * ADC | spoil * SPOIL
    # cache hints: a free segment equal to the previous free one, an RF segment with the previous RF one's dt,
    # both within one dispatch
    first = np.zeros(S, dtype=bool)
    first[np.asarray(starts, dtype=np.int64)] = True
    chunk_id = np.cumsum(first) - 1
    stepped = np.ones(S, dtype=bool)  # segments the kernels step (not inside a reused pulse's occurrence)
    for pid, (_, k, occ) in enumerate(pulses):
        heads = np.fromiter((a for a, _ in occ), dtype=np.int64, count=len(occ))
        phase = np.fromiter((ph for _, ph in occ), dtype=np.float64, count=len(occ))
        stepped[_span_index(heads, np.full(len(occ), k))] = False
        flags[heads] |= PULSE
        flags[heads[spoil[heads + k - 1]]] |= PULSE_SPOIL
        rec["pad"][heads, 1], rec["pad"][heads, 2] = pid, k
        rec["f"][heads, 7] = np.remainder(phase, 2 * math.pi).astype(np.float32)
    if nufft is not None:  # NUFFT readout runs: slot in their dispatch, kind, length; their samples are not stepped
        heads = np.fromiter((a for a, _ in nufft.runs), dtype=np.int64, count=len(nufft.runs))
        lens = np.fromiter((k for _, k in nufft.runs), dtype=np.int64, count=len(nufft.runs))
        if len(heads):
            if np.any(np.diff(heads) <= 0):
                raise ValueError("NUFFT runs must be in segment order")
            c = chunk_id[heads]
            slot = np.arange(len(heads)) - np.searchsorted(c, c, side="left")  # the run's rank in its dispatch
            flags[heads] |= NUFFT
            flags[heads[spoil[heads + lens - 1]]] |= PULSE_SPOIL
            rec["pad"][heads, 1], rec["pad"][heads, 2], rec["pad"][heads, 3] = slot, np.asarray(nufft.kind), lens
            stepped[_span_index(heads, lens)] = False
    key = np.stack([dt, g[:, 0], g[:, 1], g[:, 2], b0], 1)
    zapped = np.zeros(S, dtype=bool)
    if runs:
        merge_free(flags, rec, is_rf, adc, spoil, chunk_id, stepped)
        zapped = zap_spoiled(flags, rec, is_rf, spoil, chunk_id, stepped)
        if nufft is not None:  # a run followed by a Mz-only record: its Mxy at the end is never read
            for a, k in nufft.runs:
                if a + k < S and zapped[a + k]:
                    flags[a] |= XY_DEAD
    # the free segments' cache key in moments (equal to (dt, g, b0_offset) for single segments; merged ones carry
    # their sums)
    fkey = rec["d"].copy()
    for kind, k, bit, nxt, cols in ((~is_rf & stepped & ~zapped, fkey, SAME_FREE, NEXT_FREE, slice(0, 5)),
                                    (is_rf & stepped, key, SAME_RF_DT, NEXT_RF_DT, slice(0, 1))):
        idx = np.flatnonzero(kind)
        if len(idx) > 1:
            prev, cur = idx[:-1], idx[1:]
            same = _close(k[cur, cols], k[prev, cols]).all(1) & (chunk_id[cur] == chunk_id[prev])
            flags[cur[same]] |= bit  # forward order: equal to the previous one of the kind
            flags[prev[same]] |= nxt  # reverse order: equal to the next one of the kind
    # free records with the previous free record's float32 duration (in the dispatch): their relaxation factors
    # are reused even where the phase differs
    idx = np.flatnonzero(~is_rf & stepped & ~zapped)
    if len(idx) > 1:
        prev, cur = idx[:-1], idx[1:]
        same_dt = (rec["f"][cur, 0] == rec["f"][prev, 0]) & (chunk_id[cur] == chunk_id[prev])
        flags[cur[same_dt]] |= SAME_FREE_DT
    if runs:
        mark_runs(flags, rec, key, is_rf, adc, spoil, chunk_id, stepped)
        # SAME_RUN: the previous readout run in the dispatch (per-sample or NUFFT: both relax Mz once over the run
        # and keep the factor) has the same float32 sample dt and length
        heads = np.flatnonzero(flags & (RUN | NUFFT))
        flags[heads] &= ~SAME_RUN
        prev, cur = heads[:-1], heads[1:]
        same = ((chunk_id[cur] == chunk_id[prev]) & (rec["pad"][cur, 3] == rec["pad"][prev, 3])
                & (rec["f"][cur, 0] == rec["f"][prev, 0]))
        flags[cur[same]] |= SAME_RUN
    mark_trains(flags, rec, is_rf, adc, spoil, chunk_id, stepped)
    # an occurrence's or a NUFFT run's other segments: the identity (the record is applied at its first segment)
    spans = [(a, k) for _, k, occ in pulses for a, _ in occ] + (list(nufft.runs) if nufft is not None else [])
    if spans:
        sa = np.fromiter((a for a, _ in spans), dtype=np.int64, count=len(spans))
        sk = np.fromiter((k for _, k in spans), dtype=np.int64, count=len(spans))
        body = _span_index(sa + 1, sk - 1)
        flags[body] = PULSE_BODY
        rec["pad"][body, 3] = np.repeat(sa, sk - 1)
    if nufft is not None:
        for r, (a, _) in enumerate(nufft.runs):
            rec["pad"][a, 4] = r  # the run's index (the adjoint's grid)
    rec["flags"] = flags
    w = -2 * math.pi * dt  # the rotation-vector factors, formed in float64
    for c, col in enumerate((w, w * rf[:, 0], w * rf[:, 1], w * g[:, 0], w * g[:, 1], w * g[:, 2], w * b0,
                             (w * rf[:, 0]) ** 2 + (w * rf[:, 1]) ** 2)):
        packed[:, c] = col
    rec["w"] = packed
    for _, _, occ in pulses:  # a reused pulse's head is never stepped: its w_rf holds cos, sin of the RF phase
        heads = np.fromiter((a for a, _ in occ), dtype=np.int64, count=len(occ))
        rec["w"][heads, 1] = np.fromiter((math.cos(ph) for _, ph in occ), dtype=np.float64, count=len(occ))
        rec["w"][heads, 2] = np.fromiter((math.sin(ph) for _, ph in occ), dtype=np.float64, count=len(occ))
    rec["pad"][:, 0] = np.where(adc, np.cumsum(adc) - 1, -1)  # ADC row
    ends_adc = np.flatnonzero((flags & MERGED != 0) & (flags & ADC != 0))  # merged runs sampled at their end
    rec["pad"][ends_adc, 0] = rec["pad"][ends_adc + rec["pad"][ends_adc, 3] - 1, 0]
    return rec


def _stretches(ok: np.ndarray, link: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """(first, end) of every maximal stretch of ``ok`` segments, where ``link[s]`` says segment s continues the
    stretch of segment s - 1 (it implies ok[s] and ok[s - 1])."""
    head = ok & ~link
    tail = ok.copy()
    tail[:-1] &= ~link[1:]
    return np.flatnonzero(head), np.flatnonzero(tail) + 1


def _chained(ok: np.ndarray, also: np.ndarray) -> np.ndarray:
    """link[s] = ok[s] and ok[s - 1] and ``also[s]`` (the condition on the pair (s - 1, s))."""
    link = np.zeros(ok.size, dtype=bool)
    link[1:] = ok[1:] & ok[:-1] & also[1:]
    return link


def mark_runs(flags, rec, key, is_rf, adc, spoil, chunk_id, stepped) -> None:
    """HB_RUN on the first of every maximal run (>= MIN_RUN) of consecutive free ADC segments with one key (dt, g,
    b0_offset), none spoiled, in one dispatch; its length in pad[3]. The run's first segment keeps its SAME_FREE
    hint (the kernel's factor cache); the rest of the run is skipped, so hints on it are cleared."""
    ok = ~is_rf & adc & ~spoil & stepped
    pair = np.zeros(ok.size, dtype=bool)
    pair[1:] = _close(key[1:], key[:-1]).all(1) & (chunk_id[1:] == chunk_id[:-1])
    first, end = _stretches(ok, _chained(ok, pair))
    keep = end - first >= MIN_RUN
    first, end = first[keep], end[keep]
    flags[first] |= RUN
    rec["pad"][first, 3] = end - first
    inner = np.zeros(ok.size + 1, dtype=np.int64)
    np.add.at(inner, first + 1, 1)
    np.add.at(inner, end, -1)
    flags[np.cumsum(inner[:-1]) > 0] &= ~(SAME_FREE | NEXT_FREE)


def merge_free(flags, rec, is_rf, adc, spoil, chunk_id, stepped) -> None:
    """Merge every maximal run (>= 2) of consecutive stepped free segments in one dispatch, none sampled but possibly
    the last, only the last spoiled, into its first record (MERGED, the count in pad[3]; ADC when the last one is
    sampled, the sample taken after the merged step): free precession and relaxation compose exactly, so the run is
    one free segment of the summed duration, gradient moment and b0_offset area. Only for forwards without
    checkpoints (the adjoint steps and differentiates every segment)."""
    ok = stepped & ~is_rf & ~adc
    S = ok.size
    same_chunk = np.zeros(S, dtype=bool)
    same_chunk[1:] = chunk_id[1:] == chunk_id[:-1]
    unspoiled_before = np.zeros(S, dtype=bool)
    unspoiled_before[1:] = ~spoil[:-1]
    first, end = _stretches(ok, _chained(ok, same_chunk & unspoiled_before))
    # a sampled free segment right after a stretch (same dispatch, no spoil between) is its last member
    e = np.minimum(end, S - 1)
    takes = (end < S) & stepped[e] & ~is_rf[e] & adc[e] & same_chunk[e] & unspoiled_before[e]
    end = end + takes
    keep = end - first >= 2
    first, end = first[keep], end[keep]
    if first.size == 0:
        return
    bounds = np.stack([first, end], 1).reshape(-1)
    sums = np.add.reduceat(rec["d"], bounds[bounds < S], axis=0)[::2]
    rec["d"][first] = sums
    rec["f"][first, 0] = sums[:, 0].astype(np.float32)
    flags[first] |= MERGED | (flags[end - 1] & (SPOIL | ADC))
    rec["pad"][first, 3] = end - first
    covered = np.zeros(S + 1, dtype=np.int64)
    np.add.at(covered, first + 1, 1)
    np.add.at(covered, end, -1)
    stepped[np.cumsum(covered[:-1]) > 0] = False


def zap_spoiled(flags, rec, is_rf, spoil, chunk_id, stepped) -> np.ndarray:
    """SPOILED_FREE on the stepped free records without ADC (single segments or merged runs) whose transverse
    magnetization never reaches a sample: spoiled at their end, or starting at Mxy = 0 (right after a spoil). Only
    Mz matters there: the kernel relaxes it over the record and zeroes Mxy (no precession factors, so such records
    neither fill nor use the cache). The free records that follow such a record back to back in the same dispatch
    also start at Mxy = 0 and are folded into it (MERGED: relaxation over the summed duration). Returns the marked
    records."""
    S = flags.size
    r = np.flatnonzero(stepped)  # records, in order
    zapped = np.zeros(S, dtype=bool)
    if r.size == 0:
        return zapped
    length = np.where(flags[r] & MERGED, rec["pad"][r, 3], 1)
    end = r + length
    free = ~is_rf[r] & (flags[r] & ADC == 0)
    after_spoil = np.concatenate([[False], spoil[:-1]])[r]
    can_open = free & ((flags[r] & SPOIL != 0) | after_spoil)
    # records linked to the previous one: free, starting where it ends, in its dispatch
    link = np.zeros(r.size, dtype=bool)
    link[1:] = free[1:] & (r[1:] == end[:-1]) & (chunk_id[r[1:]] == chunk_id[r[:-1]])
    # in a group of linked records, a chain opens at the first record that can open one and runs to the group's end
    group = np.cumsum(~link) - 1
    group_first = np.flatnonzero(~link)
    candidate = np.where(can_open, np.arange(r.size), r.size)
    opens_at = np.minimum.reduceat(candidate, group_first)[group]
    head = np.arange(r.size) == opens_at
    folded = np.arange(r.size) > opens_at
    if folded.any():
        h = r[opens_at[folded]]  # each folded record's chain head (a segment index)
        d_add = np.zeros((S, rec["d"].shape[1]))
        np.add.at(d_add, h, rec["d"][r[folded]])
        heads_with = np.unique(h)
        rec["d"][heads_with] += d_add[heads_with]
        rec["f"][heads_with, 0] = rec["d"][heads_with, 0].astype(np.float32)
        chain_end = np.zeros(S, dtype=np.int64)
        np.maximum.at(chain_end, h, end[folded])
        flags[heads_with] |= MERGED
        rec["pad"][heads_with, 3] = chain_end[heads_with] - heads_with
        stepped[r[folded]] = False
    zapped[r[head]] = True
    flags[zapped] |= SPOILED_FREE
    return zapped


def mark_trains(flags, rec, is_rf, adc, spoil, chunk_id, stepped) -> None:
    """HB_TRAIN on the first of every maximal run (>= MIN_TRAIN) of consecutive stepped RF segments without ADC or
    spoil in one dispatch, its length in pad[3]. The forward kernel steps a train in a loop of its own, with the same
    arithmetic as segment by segment (so checkpointed forwards may use trains too); the segments keep their hints."""
    ok = is_rf & ~adc & ~spoil & stepped
    same_chunk = np.zeros(ok.size, dtype=bool)
    same_chunk[1:] = chunk_id[1:] == chunk_id[:-1]
    first, end = _stretches(ok, _chained(ok, same_chunk))
    keep = end - first >= MIN_TRAIN
    flags[first[keep]] |= TRAIN
    rec["pad"][first[keep], 3] = (end - first)[keep]


def dispatch_starts(seq: Sequence, chunk: int, rows_cap: int, spans=(), rows_w=None, heads=(),
                    runs_cap: int = NUFFT_RUNS) -> list[int]:
    """Dispatch boundaries: at most ``chunk`` segments, ``rows_cap`` partial rows (``rows_w``: the segments whose
    sample has a partial row; default every ADC segment) and ``runs_cap`` NUFFT runs (``heads``: their first
    segments) each, never inside one of ``spans`` ((first, length) of reused-pulse occurrences and NUFFT runs)."""
    adc = flag_array(seq.adc).astype(np.int64) if rows_w is None else np.asarray(rows_w, dtype=np.int64)
    S = len(seq)
    inside = np.zeros(S + 1, dtype=bool)  # a cut before segment e would split a span
    for a, k in spans:
        inside[a + 1 : a + k] = True
    hw = np.zeros(S, dtype=np.int64)
    hw[np.asarray(list(heads), dtype=np.int64)] = 1
    cumh = np.concatenate([[0], np.cumsum(hw)])
    starts, s = [], 0
    cum = np.concatenate([[0], np.cumsum(adc)])
    while s < S:
        starts.append(s)
        e = min(S, s + chunk)
        if cum[e] - cum[s] > rows_cap:  # cut at the segment holding sample number rows_cap of this dispatch
            e = int(np.searchsorted(cum, cum[s] + rows_cap, side="left"))
        if cumh[e] - cumh[s] > runs_cap:  # likewise at the runs_cap-th run head
            e = min(e, int(np.searchsorted(cumh, cumh[s] + runs_cap, side="left")))
        while s < e < S and inside[e]:
            e -= 1
        if e <= s:  # a span longer than the chunk: cut after it
            e = s + 1
            while e < S and inside[e]:
                e += 1
        s = e
    return starts


def pow2_scale(bound: float) -> float:
    """The power of two 2^-e, e = ceil(log2(bound)), that brings values up to ``bound`` into [-1, 1] (1 for a bound
    of 0 or not finite)."""
    if not (bound > 0 and math.isfinite(bound)):
        return 1.0
    return math.ldexp(1.0, -min(100, max(-100, math.ceil(math.log2(bound)))))


def sens_operand(coils: torch.Tensor, n: int, n_pad: int, n_tiles: int) -> torch.Tensor:
    """(N, C) complex sensitivities in the matrix-core A order: (n_pad / 64, tiles, 16 steps, 64 lanes), where lane
    l = 16 g + i of step u holds row i (Re S_0, Im S_0, Re S_1, Im S_1, ... of the tile) at spin 16 g + u of the
    wave."""
    C = coils.shape[1]
    full = torch.zeros(n_pad, 8 * n_tiles, dtype=torch.complex128)
    full[:n, :C] = coils.detach().to(torch.complex128).cpu()
    w = n_pad // 64
    re = full.real.reshape(w, 64, n_tiles, 8)
    im = full.imag.reshape(w, 64, n_tiles, 8)
    m = torch.stack([re, im], dim=4).reshape(w, 64, n_tiles, 16).permute(0, 2, 3, 1)  # (w, tiles, 16 rows, 64 k)
    m = m.reshape(w, n_tiles, 16, 4, 16).permute(0, 1, 4, 3, 2)  # k = 16 g + u -> (w, tiles, u, g, i)
    return m.reshape(w, n_tiles, 16, 64).float().contiguous()


@dataclass
class DeviceCase:
    """Everything the kernels read, resident on the device, and the per-dispatch plan."""

    n: int
    n_pad: int
    n_coils: int
    n_tiles: int
    n_adc: int
    spins: torch.Tensor
    pos_lo: torch.Tensor
    sens_a: torch.Tensor
    segs: torch.Tensor  # uint8 view of the records
    state: torch.Tensor  # (6, n_pad)
    state0: torch.Tensor  # the initial state (kept to rerun)
    psi_rows: torch.Tensor  # (n_adc,) float64
    starts: list[int]
    ends: list[int]
    row_starts: list[int]
    rows: list[int]
    partial: torch.Tensor
    red_work: torch.Tensor  # hb_signal_reduce's float64 workspace
    sig: torch.Tensor  # (n_adc, C, 2) float64
    n_pulses: int = 0
    maps: torch.Tensor | None = None  # (P, 12, n_pad) reused-pulse maps
    pulse_table: torch.Tensor | None = None  # (rows, 8) float64
    pulse_index: torch.Tensor | None = None  # (P, 2) int32
    pulses: list = field(default_factory=list)
    extra: dict = field(default_factory=dict)
    nu: dict | None = None  # NUFFT readouts: plan, device tables, work buffers, per-dispatch runs
    perm: np.ndarray | None = None  # the spins' order on the device (sorted for NUFFT readouts), or None
    rf_free: list[bool] = field(default_factory=list)  # per dispatch: no segment stepped as RF (hb_case.rf_free)
    sens_scale: float = 1.0  # powers of two: |sensitivity| sens_scale <= 1, |M| x_scale <= 1 (hb_case)
    x_scale: float = 1.0

    def struct(self) -> hip.HbCase:
        return hip.HbCase(self.spins.data_ptr(), self.pos_lo.data_ptr(), self.sens_a.data_ptr(),
                          self.segs.data_ptr(), self.state.data_ptr(),
                          None if self.maps is None else self.maps.data_ptr(), self.n, self.n_pad, self.n_coils,
                          self.n_tiles, self.n_segments, self.n_pulses, 0, self.sens_scale, self.x_scale, 0)

    @property
    def n_segments(self) -> int:
        return self.segs.numel() // SEG_DTYPE.itemsize


def check_inputs(phantom: Phantom, coils) -> None:
    for what, present in (("free pools", bool(phantom.pools)), ("diffusion", phantom.diffusion is not None),
                          ("motion", phantom.motion is not None)):
        if present:
            raise NotImplementedError(f"hip engine: {what} is not implemented")
    if coils is not None and coils.shape[1] > hip.MAX_COILS:
        raise ValueError(f"at most {hip.MAX_COILS} coils on hip, got {coils.shape[1]}")


def finite_inputs(phantom: Phantom, seq: Sequence, coils=None, m_init=None) -> bool:
    """Every number of the case is finite: the premise of ``read_result``'s output guard."""
    ts = [phantom.pos, phantom.pd, phantom.t1, phantom.t2, phantom.db0, phantom.b1, seq.dt, seq.rf, seq.grad,
          seq.adc_phase, seq.b0_offset, seq.rf_freq, coils, m_init]
    if phantom.bound is not None:
        ts += [getattr(phantom.bound, k) for k in ("f", "k", "t1", "t2", "dw")]
    for t in ts:
        if isinstance(t, torch.Tensor):
            if not bool(torch.isfinite(t.detach()).all()):
                return False
        elif isinstance(t, float) and not math.isfinite(t):
            return False
    return True


def check_output(case: DeviceCase, *outs: torch.Tensor, premise: bool = True) -> None:
    """Failure detection: a non-finite result from finite inputs (the case's, and ``premise``: the incoming
    gradients' of a reverse sweep) is an engine fault and raises; otherwise the result is returned as computed."""
    if premise and case.extra.get("finite_inputs") and not all(bool(torch.isfinite(o).all()) for o in outs):
        raise RuntimeError("hip engine: non-finite result from finite inputs")


def _with_flag_memo(fn):
    """Run ``fn`` with each of the sequence's flag lists converted to an array once (``sequence.flag_memo``)."""

    @functools.wraps(fn)
    def call(*args, **kwargs):
        with flag_memo():
            return fn(*args, **kwargs)

    return call


@_with_flag_memo
def prepare_case(phantom: Phantom, seq: Sequence, coils=None, m_init=None, *, chunk: int = DEFAULT_CHUNK,
                 device=None, partial_bytes: int = PARTIAL_BYTES, aligned: bool = False,
                 reuse: bool = False, surrogate: bool = False, nufft=None, conc=None) -> DeviceCase:
    """Upload a case: the buffers of the module docstring, the dispatch plan and the work buffers. ``aligned``:
    dispatches are exactly [c chunk, (c + 1) chunk) (the adjoint's checkpoint intervals), the partial buffer sized
    for the largest one. ``reuse``: reused pulses applied as per-spin maps; ``surrogate``: those maps from the
    learned surrogate (trained here on first use); ``nufft``: (width, terms) of NUFFT readouts (the spins are then
    sorted along the readout's phase; ``read_result`` restores their order). ``conc``: (coefficients (N, 6 or 7),
    mode) of a scanner's concomitant fields (``fields.spin_fields``), or None."""
    from . import mt
    from . import nufft as nu_mod

    reuse = reuse or surrogate
    check_inputs(phantom, coils)
    full = mt.needed(phantom, seq, conc)  # the full-variant kernels (engines/mt.py): plain records only
    if full and (reuse or nufft is not None):
        raise NotImplementedError("hip engine: a bound pool, concomitant fields or bound-only RF run with "
                                  "pulses='step' and readouts='direct'")
    dev = device if device is not None else hip.HIP.device
    n = len(phantom)
    n_pad = max(BLOCK, math.ceil(n / BLOCK) * BLOCK)
    if coils is None:
        coils = torch.ones(n, 1, dtype=torch.complex128)
    plan, order = None, None
    if nufft is not None:
        plan = nu_mod.make_plan(seq, *nufft)
        if plan is not None:
            pos_in = phantom.pos.detach().double().cpu()
            # the kernels' float64 position: float32 hi + float32 lo
            hi32 = pos_in.float().double()
            pos_k = (hi32 + (pos_in - hi32).float().double()).numpy()
            db0_k = phantom.db0.detach().float().double().cpu().numpy()
            order = nu_mod.sort_order(plan, pos_k, db0_k, seq)
            if not nu_mod.windows(plan, pos_k[order], db0_k[order], seq, n, n_pad):
                plan, order = None, None  # windows too wide: the per-sample path
        if order is not None:
            idx = torch.from_numpy(order)
            phantom = Phantom(phantom.pos[idx], phantom.pd[idx], phantom.t1[idx], phantom.t2[idx], phantom.db0[idx],
                              b1=None if phantom.b1 is None else phantom.b1[idx])
            coils = coils[idx]
            m_init = None if m_init is None else m_init[idx]
    n_coils = coils.shape[1]
    n_tiles = math.ceil(n_coils / 8)
    pad = lambda t: torch.cat([t, t.new_zeros(n_pad - n, *t.shape[1:])])
    b1 = phantom.b1 if phantom.b1 is not None else torch.ones(n, dtype=torch.float64)
    pos64 = phantom.pos.detach().double().cpu()
    pos_hi = pos64.float()
    cols = [pos_hi[:, 0].double(), pos_hi[:, 1].double(), pos_hi[:, 2].double(), phantom.pd, 1 / phantom.t1,
            1 / phantom.t2, phantom.db0, b1]
    spins = torch.stack([pad(c.detach().double().cpu()) for c in cols]).float()
    spins[4:6, n:] = 1.0  # padding spins: finite rates, zero PD and state
    pos_lo = pad(pos64 - pos_hi.double()).T.float()
    if m_init is None:
        m64 = torch.stack([torch.zeros(n, dtype=torch.float64), torch.zeros(n, dtype=torch.float64),
                           phantom.pd.detach().double().cpu()])
    else:
        m64 = m_init.detach().double().cpu()[:, :3].T
    hi = m64.float()
    state = torch.zeros(6, n_pad, dtype=torch.float32)
    state[:3, :n] = hi
    state[3:, :n] = (m64 - hi.double()).float()
    n_adc = seq.n_adc
    n_blocks = n_pad // BLOCK
    vals = 16 * n_tiles
    rows_cap = max(8, (partial_bytes // (4 * vals * n_blocks)) // 8 * 8)
    reused = pulses.reused_pulses(seq) if reuse else []
    spans = [(a, k) for _, k, occ in reused for a, _ in occ]
    adc_np = flag_array(seq.adc).astype(np.int64)
    rows_w = adc_np.copy()  # samples with a partial row (not in NUFFT runs)
    if plan is not None:
        for a, k in plan.runs:
            rows_w[a : a + k] = 0
        spans += list(plan.runs)
    starts = (list(range(0, len(seq), int(chunk))) if aligned
              else dispatch_starts(seq, int(chunk), rows_cap, spans, rows_w,
                                   [a for a, _ in plan.runs] if plan is not None else ()))
    ends = starts[1:] + [len(seq)]
    adc_cum = np.concatenate([[0], np.cumsum(rows_w)])
    row_starts = [int(adc_cum[s]) for s in starts]
    rows = [int(adc_cum[e] - adc_cum[s]) for s, e in zip(starts, ends)]
    rec = segment_table(seq, starts, reused, runs=not aligned and not full, nufft=plan)
    if full:
        mt.mark_bound_only(rec, seq)
    covered = np.zeros(len(seq), dtype=bool)  # segments applied as reused-pulse maps
    for _, k, occ in reused:
        for a, _ in occ:
            covered[a : a + k] = True
    rf_cum = np.concatenate([[0], np.cumsum(flag_array(seq.is_rf) & ~covered)])
    # the coil contraction's scales: |M| <= max(PD, |m_init|) along any sequence (rotations keep |M|, relaxation
    # moves it toward (0, 0, PD))
    sens_a = sens_operand(coils, n, n_pad, n_tiles)
    sens_scale = pow2_scale(float(sens_a.abs().max()) if sens_a.numel() else 0.0)
    x_scale = pow2_scale(max(float(phantom.pd.detach().abs().max()) if n else 0.0,
                             float((m64[0] * m64[0] + m64[1] * m64[1] + m64[2] * m64[2]).max().sqrt()) if n
                             else 0.0))
    psi_rows = np.remainder(seq.adc_phase.detach().double().cpu().numpy()[flag_array(seq.adc)],
                            2 * math.pi)
    max_rows = max(rows + [0])
    case = DeviceCase(
        n=n, n_pad=n_pad, n_coils=n_coils, n_tiles=n_tiles, n_adc=n_adc,
        spins=spins.to(dev).contiguous(),
        pos_lo=pos_lo.to(dev).contiguous(),
        sens_a=sens_a.to(dev),
        segs=torch.from_numpy(rec.view(np.uint8).copy()).to(dev),
        state=state.to(dev).contiguous(),
        state0=state.to(dev).contiguous(),
        psi_rows=torch.from_numpy(psi_rows.copy()).to(dev),
        starts=starts, ends=ends, row_starts=row_starts, rows=rows,
        rf_free=[bool(rf_cum[e] == rf_cum[s]) for s, e in zip(starts, ends)],
        sens_scale=sens_scale, x_scale=x_scale,
        partial=torch.empty(max(1, n_blocks * max_rows * vals), dtype=torch.float32, device=dev),
        red_work=torch.empty(max(1, -(-n_blocks // REDUCE_CHUNK) * max_rows * vals), dtype=torch.float64,
                             device=dev),
        sig=torch.zeros(max(n_adc, 1), n_coils, 2, dtype=torch.float64, device=dev),
        **pulse_buffers(seq, reused, n_pad, dev),
    )
    case.extra["finite_inputs"] = finite_inputs(phantom, seq, coils, m_init)
    if plan is not None:
        attach_nufft(case, plan, order, seq, adc_np, starts, ends, dev)
    if full:
        mt.attach(case, phantom, seq, coils, m_init, n, n_pad, dev, conc)
    if surrogate and case.n_pulses:
        from . import surrogate as sur

        sur.prepare(case, seq)
    return case


def attach_nufft(case: DeviceCase, plan, order, seq: Sequence, adc_np, starts, ends, dev) -> None:
    """The NUFFT plan's device buffers, the partial rows' map to signal rows, and per dispatch its runs."""
    from . import nufft as nu_mod

    sig_row = np.cumsum(adc_np) - 1  # signal row of every ADC segment
    in_run = np.zeros(len(seq), dtype=bool)
    for a, k in plan.runs:
        in_run[a : a + k] = True
    mapped = sig_row[(adc_np == 1) & ~in_run]  # partial row -> signal row
    heads = np.asarray([a for a, _ in plan.runs], dtype=np.int64)
    per_disp = [np.flatnonzero((heads >= s0) & (heads < s1)).astype(np.int32) for s0, s1 in zip(starts, ends)]
    runs_max = max(1, max(len(x) for x in per_disp))
    wmax = int(plan.win[..., 1].max())
    cols = wmax * plan.terms
    vals = 16 * case.n_tiles
    n_blocks = case.n_pad // BLOCK
    case.perm = order
    case.nu = dict(nu_mod.to_device(plan, dev), plan=plan, cols=cols, runs_max=runs_max,
                   row_map=torch.from_numpy(mapped.astype(np.int32)).to(dev),
                   row_of=torch.from_numpy(sig_row[heads].astype(np.int32)).to(dev),
                   kind_of=torch.from_numpy(plan.kind).to(dev),
                   slots=[torch.from_numpy(x).to(dev) for x in per_disp], n_slots=[len(x) for x in per_disp],
                   part=torch.empty(n_blocks * runs_max * cols * vals, dtype=torch.float32, device=dev),
                   grid=torch.empty(runs_max * plan.terms * plan.ng * vals, dtype=torch.float64, device=dev))


def pulse_buffers(seq: Sequence, reused, n_pad: int, dev) -> dict:
    """The reused pulses' float64 table and index, and room for their per-spin maps."""
    if not reused:
        return {}
    table, index = pulses.pulse_table(seq, reused)
    return dict(n_pulses=len(reused), pulses=reused,
                maps=torch.empty(len(reused) * 12 * n_pad, dtype=torch.float32, device=dev),
                pulse_table=torch.from_numpy(table).to(dev), pulse_index=torch.from_numpy(index).to(dev))


def run_forward(case: DeviceCase, lib=None, keep: bool = False, timer=None) -> list[torch.Tensor] | None:
    """All dispatches on the current stream from ``case.state0``: the state ends in ``case.state``, the signal in
    ``case.sig``. ``keep``: also return the state before every dispatch (the adjoint's checkpoints). ``timer``:
    called as timer("begin") / timer("end") around every forward-kernel launch (the benchmark records events)."""
    lib = lib or hip.HIP.lib()
    case.state.copy_(case.state0)
    mt_bufs = case.extra.get("mt")
    if mt_bufs is not None:
        from . import mt

        if mt_bufs["mb"] is not None:
            mt_bufs["mb"].copy_(mt_bufs["mb0"])
    st = case.struct()
    n_blocks = case.n_pad // BLOCK
    ckpts = [] if keep else None
    if case.extra.get("surrogate") is not None:
        from . import surrogate

        surrogate.compose_maps(case, hip.load())
    elif case.n_pulses:
        lib.pulse_maps(hip.ctypes.byref(st), case.pulse_table.data_ptr(), case.pulse_index.data_ptr(),
                       case.n_pulses, None, 0, case.maps.data_ptr())
    nu = case.nu
    nu_arg = None
    if nu is not None:
        nu_arg = hip.HbNufft(nu["nu"].data_ptr(), nu["nu_win"].data_ptr(), nu["part"].data_ptr(), nu["runs_max"],
                             nu["cols"], nu["plan"].terms, 0)
    for d, (s0, s1, r0, nr) in enumerate(zip(case.starts, case.ends, case.row_starts, case.rows)):
        if keep:  # with a bound pool, its Mzb and low part as rows 6, 7
            mb = None if mt_bufs is None else mt_bufs["mb"]
            ckpts.append(case.state.clone() if mb is None else torch.cat([case.state, mb]))
        st.rf_free = int(bool(case.rf_free) and case.rf_free[d])
        if timer:
            timer("begin")
        if mt_bufs is not None:
            mt.forward_dispatch(case, st, s0, s1, nr, r0)
        else:
            lib.forward(hip.ctypes.byref(st), s0, s1, case.partial.data_ptr(), nr,
                        None if nu_arg is None else hip.ctypes.byref(nu_arg))
        if timer:
            timer("end")
        if nr:
            lib.signal_reduce(case.partial.data_ptr(), n_blocks, nr, case.n_tiles, case.n_coils,
                              case.psi_rows.data_ptr(), r0,
                              None if nu is None else nu["row_map"].data_ptr(), case.red_work.data_ptr(),
                              case.sig.data_ptr())
        if nu is not None and nu["n_slots"][d]:
            pl = nu["plan"]
            lib.nufft_grid(nu["part"].data_ptr(), n_blocks, nu["runs_max"], nu["cols"], case.n_tiles, nu["n_slots"][d],
                           nu["slots"][d].data_ptr(), nu["kind_of"].data_ptr(), pl.ng, pl.terms,
                           nu["nu_csr_off"].data_ptr(), nu["nu_csr"].data_ptr(), nu["nu_csr_base"].data_ptr(),
                           nu["grid"].data_ptr())
            lib.nufft_signal(nu["grid"].data_ptr(), nu["n_slots"][d], nu["slots"][d].data_ptr(),
                             nu["row_of"].data_ptr(), pl.n, pl.ng, pl.terms, case.n_tiles, case.n_coils,
                             nu["nu_tw"].data_ptr(), nu["nu_lw"].data_ptr(), case.psi_rows.data_ptr(),
                             case.sig.data_ptr())
    return ckpts


def read_result(case: DeviceCase) -> SimResult:
    """(C, n_adc) complex64 signal and (N, 3) float32 magnetization on the CPU (the GPU engines' contract), with a
    bound pool its Mz in ``m_bound`` (N,) (the reference's driver.py:266)."""
    sig = case.sig[: case.n_adc].cpu()
    signal = torch.complex(sig[..., 0], sig[..., 1]).T.to(torch.complex64).contiguous()
    st = case.state.cpu().double()
    m = (st[:3, : case.n] + st[3:, : case.n]).T.float().contiguous()
    m_bound = None
    if case.extra.get("mt") is not None and case.extra["mt"]["mb"] is not None:
        mb = case.extra["mt"]["mb"].cpu().double()
        m_bound = (mb[0, : case.n] + mb[1, : case.n]).float().contiguous()
    if case.perm is not None:  # back to the caller's spin order (NUFFT readouts; never with a bound pool)
        inv = np.empty_like(case.perm)
        inv[case.perm] = np.arange(len(case.perm))
        m = m[torch.from_numpy(inv)].contiguous()
    check_output(case, sig, st, *(() if m_bound is None else (m_bound,)))
    res = SimResult(signal, m, m_bound)
    if case.extra.get("surrogate") is not None:
        from . import surrogate

        res.surrogate = surrogate.report(case)
    return res


def forward(phantom: Phantom, seq: Sequence, coils=None, m_init=None, *, chunk: int = DEFAULT_CHUNK,
            reuse: bool = False, surrogate: bool = False, nufft=None, conc=None) -> SimResult:
    case = prepare_case(phantom, seq, coils, m_init, chunk=chunk, reuse=reuse, surrogate=surrogate, nufft=nufft,
                        conc=conc)
    run_forward(case)
    hip.HIP.synchronize()
    return read_result(case)
