"""Reused pulses (``simulate(..., pulses="reuse")``): a shaped RF pulse that recurs, identical up to a global RF
phase, is composed once per spin into its affine map m -> m + D m + v, in float64 on the GPU, and applied at every
occurrence conjugated by the occurrence's phase: an RF phase phi is a rotation of the frame about z, so the
occurrence maps m to m + Rz(phi) (D Rz(-phi) m + v).

Reference: ``src/blochparty/engines/gpu/driver.py:109-164`` (grouping: runs of >= 2 consecutive RF segments without
ADC, a spoil only on the last, with an earlier run's dt, gradient, b0_offset and flags and its RF times e^{i phase}
to ``tol`` of the peak, phase by least squares) and ``cuda/pulse.cu.in`` + ``kernels.cu.in:141-165`` (maps and their
application). The reference's are forward only; here the adjoint applies them too (``engines/adjoint.py``).
"""

from __future__ import annotations

import numpy as np

from ..sequence import Sequence


def rf_runs(is_rf: np.ndarray, adc: np.ndarray, spoil: np.ndarray) -> list[tuple[int, int]]:
    """(first, length) of every maximal run of RF segments without ADC, cut after a spoiled segment."""
    runs, start = [], None
    live = is_rf & ~adc
    for s in range(len(is_rf)):
        if live[s]:
            if start is None:
                start = s
            if spoil[s]:
                runs.append((start, s + 1 - start))
                start = None
        elif start is not None:
            runs.append((start, s - start))
            start = None
    if start is not None:
        runs.append((start, len(is_rf) - start))
    return runs


def reused_pulses(seq: Sequence, tol: float = 1e-6) -> list[tuple[int, int, list[tuple[int, float]]]]:
    """[(first segment, length, [(occurrence start, phase [rad]), ...])] of every pulse that occurs at least twice,
    most segments saved first."""
    is_rf = np.asarray(seq.is_rf, dtype=bool)
    adc = np.asarray(seq.adc, dtype=bool)
    spoil = np.asarray(seq.spoil, dtype=bool)
    S = len(is_rf)
    dt = seq.dt.detach().double().cpu().numpy()
    rf = seq.rf.detach().double().cpu().numpy()
    z = rf[:, 0] + 1j * rf[:, 1]
    g = seq.grad.detach().double().cpu().numpy()
    b0 = np.zeros(S) if seq.b0_offset is None else seq.b0_offset.detach().double().cpu().numpy()

    def close(x, y):
        return np.abs(x - y).max() <= tol * max(np.abs(x).max(), np.abs(y).max(), 1e-300)

    groups: list[tuple[int, int, list]] = []
    by_shape: dict[tuple, list[int]] = {}
    for a, k in rf_runs(is_rf, adc, spoil):
        if k < 2:
            continue
        cur = slice(a, a + k)
        key = (k, round(float(dt[cur].sum()) * 1e9))
        for gi in by_shape.get(key, ()):
            first = groups[gi][0]
            ref = slice(first, first + k)
            if not (np.array_equal(spoil[cur], spoil[ref]) and close(dt[cur], dt[ref]) and close(g[cur], g[ref])
                    and close(b0[cur], b0[ref])):
                continue
            c = np.vdot(z[ref], z[cur])  # least-squares phase of this run against the first
            ph = float(np.angle(c)) if c != 0 else 0.0
            if close(z[cur], z[ref] * np.exp(1j * ph)):
                groups[gi][2].append((a, ph))
                break
        else:
            by_shape.setdefault(key, []).append(len(groups))
            groups.append((a, k, [(a, 0.0)]))
    groups = [p for p in groups if len(p[2]) > 1]
    return sorted(groups, key=lambda p: -(len(p[2]) - 1) * p[1])


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
