"""Reused pulses (``simulate(..., pulses="reuse")``): a shaped RF pulse that recurs, identical up to a global RF
phase, is composed once per spin into its affine map m -> m + D m + v, in float64 on the GPU, and applied at every
occurrence conjugated by the occurrence's phase: an RF phase phi is a rotation of the frame about z, so the
occurrence maps m to m + Rz(phi) (D Rz(-phi) m + v).

Reference behaviour: ``src/blochparty/engines/gpu/driver.py:109-164`` (which runs count as one pulse) and
``cuda/pulse.cu.in`` + ``kernels.cu.in:141-165`` (maps and their application). The reference's are forward only;
here the adjoint applies them too (``engines/adjoint.py``).

What counts as one pulse: maximal runs of >= 2 consecutive RF segments without ADC (a spoiled segment ends its run)
that have the same length, and whose dt, gradient, b0_offset and spoil flags equal a first run's, and whose RF equals
the first run's times e^{i phi} for one phase phi, each to ``tol`` of the larger peak; phi is the least-squares phase.

How (``reused_pulses``): runs are found with array operations, then, per run length, grouped by prototype
elimination: the earliest ungrouped run is a prototype, all remaining runs of its length are compared with it at
once (one matrix-vector product gives every candidate's least-squares phase, one reduction each comparison), the
matches join it, and the next ungrouped run is the next prototype. A pulse's occurrences are in time order, the
first (the prototype) with phase 0; pulses come in the order of their first occurrence.
"""

from __future__ import annotations

import numpy as np

from ..sequence import Sequence, flag_array


def rf_runs(is_rf: np.ndarray, adc: np.ndarray, spoil: np.ndarray) -> list[tuple[int, int]]:
    """(first, length) of every maximal run of RF segments without ADC; a spoiled segment is the last of its run."""
    live = np.asarray(is_rf, dtype=bool) & ~np.asarray(adc, dtype=bool)
    if not live.any():
        return []
    after_spoil = np.zeros_like(live)
    after_spoil[1:] = np.asarray(spoil, dtype=bool)[:-1]
    prev = np.concatenate([[False], live[:-1]])
    starts = np.flatnonzero(live & (~prev | after_spoil))
    # a run ends before the next start or at the first segment that is not live
    dead = np.append(np.flatnonzero(~live), live.size)  # with a sentinel past the end
    nxt_dead = dead[np.searchsorted(dead, starts)]
    nxt_start = np.append(starts[1:], live.size)
    ends = np.minimum(nxt_dead, nxt_start)
    return list(zip(starts.tolist(), (ends - starts).tolist()))


def _within(diff_peak: np.ndarray, peak_a: np.ndarray, peak_b: np.ndarray, tol: float) -> np.ndarray:
    return diff_peak <= tol * np.maximum(np.maximum(peak_a, peak_b), 1e-300)


def reused_pulses(seq: Sequence, tol: float = 1e-6) -> list[tuple[int, int, list[tuple[int, float]]]]:
    """[(first segment, length, [(occurrence start, phase [rad]), ...])] of every pulse that occurs at least twice
    (the module's rules), in the order of first occurrence."""
    is_rf, adc, spoil = (flag_array(f) for f in (seq.is_rf, seq.adc, seq.spoil))
    dt = seq.dt.detach().double().cpu().numpy()
    rf = seq.rf.detach().double().cpu().numpy()
    z_all = rf[:, 0] + 1j * rf[:, 1]
    g_all = seq.grad.detach().double().cpu().numpy()
    b0_all = np.zeros(len(dt)) if seq.b0_offset is None else seq.b0_offset.detach().double().cpu().numpy()
    runs = [(a, k) for a, k in rf_runs(is_rf, adc, spoil) if k >= 2]
    by_len: dict[int, list[int]] = {}
    for a, k in runs:
        by_len.setdefault(k, []).append(a)
    found = []
    for k, starts in by_len.items():
        idx = np.asarray(starts)[:, None] + np.arange(k)[None, :]  # (R, k) segment indices
        Z, D, G, B, P = z_all[idx], dt[idx], g_all[idx], b0_all[idx], spoil[idx]
        peak = {name: np.abs(x).reshape(len(starts), -1).max(1) for name, x in (("z", Z), ("d", D), ("g", G),
                                                                                  ("b", B))}
        open_ = np.ones(len(starts), dtype=bool)
        for p in range(len(starts)):
            if not open_[p]:
                continue
            cand = np.flatnonzero(open_)
            c = Z[cand] @ np.conj(Z[p])  # least-squares phase of each candidate against the prototype
            phi = np.where(c != 0, np.angle(c), 0.0)
            fit = Z[p][None, :] * np.exp(1j * phi)[:, None]
            ok = (P[cand] == P[p]).all(1)
            for name, X in (("d", D), ("g", G), ("b", B)):
                gap = np.abs(X[cand] - X[p][None]).reshape(cand.size, -1).max(1)
                ok &= _within(gap, peak[name][cand], peak[name][p], tol)
            ok &= _within(np.abs(Z[cand] - fit).max(1), peak["z"][cand], peak["z"][p], tol)
            ok[cand == p] = True
            members = cand[ok]
            open_[members] = False
            if members.size > 1:
                occ = [(int(starts[m]), 0.0 if m == p else float(phi[np.searchsorted(cand, m)])) for m in members]
                found.append((int(starts[p]), k, occ))
    return sorted(found, key=lambda pulse: pulse[0])


def pulse_table(seq: Sequence, pulses) -> tuple[np.ndarray, np.ndarray]:
    """The float64 segments of every distinct pulse's first occurrence, (sum of lengths, 8): dt, rf_x, rf_y, g_x,
    g_y, g_z, b0_offset, 0; and the (P, 2) int32 (row offset, length) of each pulse in it."""
    dt = seq.dt.detach().double().cpu().numpy()
    rf = seq.rf.detach().double().cpu().numpy()
    g = seq.grad.detach().double().cpu().numpy()
    b0 = np.zeros(len(dt)) if seq.b0_offset is None else seq.b0_offset.detach().double().cpu().numpy()
    rows, index, off = [], [], 0
    for a, k, _ in pulses:
        s = slice(a, a + k)
        rows.append(np.stack([dt[s], rf[s, 0], rf[s, 1], g[s, 0], g[s, 1], g[s, 2], b0[s], np.zeros(k)], 1))
        index.append((off, k))
        off += k
    table = np.concatenate(rows) if rows else np.zeros((1, 8))
    return np.ascontiguousarray(table), np.asarray(index if index else [(0, 0)], dtype=np.int32)
