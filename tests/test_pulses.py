"""Reused-pulse grouping and dispatch planning on the CPU (the reference's tests/engines/test_gpu_reuse.py cases)."""

import math

import numpy as np
import pytest
import torch

import blochparty_amd as bp
from blochparty_amd.engines import driver, pulses


@pytest.mark.parametrize("make, distinct", [
    (lambda: bp.tse_2d(48, etl=8, tr=0.8, slice_thickness=5e-3, spoiling="gradient", n_dummy=1), 2),
    (lambda: bp.gre_2d(48, ramps=True, spoiling="rf", slice_thickness=5e-3, n_dummy=2), 1),
    (lambda: bp.gre_cartesian((6, 6, 6), flip_deg=10.0, tr=8e-3, slice_thickness=0.22, rf_duration=2e-3,
                              system=bp.System(rf_raster=10e-6)), 1),
])
def test_grouping_finds_the_distinct_pulses(make, distinct):
    seq = make()
    groups = pulses.reused_pulses(seq)
    assert len(groups) == distinct
    rf = seq.rf.numpy()
    for a, k, occ in groups:
        assert k >= 2 and all(seq.is_rf[a : a + k]) and not any(seq.adc[a : a + k])
        for s, ph in occ:  # every occurrence is the first one turned by its phase
            z0 = rf[a : a + k, 0] + 1j * rf[a : a + k, 1]
            z = rf[s : s + k, 0] + 1j * rf[s : s + k, 1]
            assert np.abs(z - z0 * np.exp(1j * ph)).max() <= 1e-6 * np.abs(z0).max()


def test_rf_spoiling_phases_are_recovered():
    seq = bp.gre_2d(12, ramps=True, spoiling="rf", slice_thickness=5e-3, n_dummy=2)
    ((a, k, occ),) = pulses.reused_pulses(seq)
    phases = np.array([ph for _, ph in occ])
    expected = np.array([math.radians((117.0 * j * (j + 1) / 2) % 360.0) for j in range(len(occ))])
    d = np.angle(np.exp(1j * (phases - (expected - expected[0]))))
    assert np.abs(d).max() < 1e-9


def test_dispatches_never_split_an_occurrence_and_records_skip_them():
    seq = bp.tse_2d(16, etl=4, tr=0.3, slice_thickness=5e-3, rf_duration=1e-3, system=bp.System(rf_raster=20e-6))
    groups = pulses.reused_pulses(seq)
    spans = [(s, k) for _, k, occ in groups for s, _ in occ]
    for chunk in (7, 33, 301):
        starts = driver.dispatch_starts(seq, chunk, 1 << 20, spans)
        for s0 in starts:
            assert not any(a < s0 < a + k for a, k in spans), (chunk, s0)
    rec = driver.segment_table(seq, [0], groups)
    fl = rec["flags"]
    for pid, (_, k, occ) in enumerate(groups):
        for s, _ in occ:
            assert fl[s] & driver.PULSE and rec["pad"][s, 1] == pid and rec["pad"][s, 2] == k
    inner = np.zeros(len(seq), dtype=bool)
    for a, k in spans:
        inner[a : a + k] = True
    # cache hints name stepped segments only: no hint on a segment inside an occurrence
    assert not np.any(fl[inner] & (driver.SAME_FREE | driver.SAME_RF_DT | driver.NEXT_FREE | driver.NEXT_RF_DT))


def test_rf_runs_against_a_plain_scan():
    """rf_runs by array operations against a segment-by-segment scan of the definition, on random flag patterns."""

    def scan(is_rf, adc, spoil):
        runs, start = [], None
        for s in range(len(is_rf)):
            live = is_rf[s] and not adc[s]
            if start is not None and not live:
                runs.append((start, s - start))
                start = None
            if live:
                start = s if start is None else start
                if spoil[s]:
                    runs.append((start, s + 1 - start))
                    start = None
        if start is not None:
            runs.append((start, len(is_rf) - start))
        return runs

    rng = np.random.default_rng(0)
    for trial in range(200):
        n = int(rng.integers(1, 60))
        is_rf, adc, spoil = rng.random(n) < 0.7, rng.random(n) < 0.1, rng.random(n) < 0.15
        assert pulses.rf_runs(is_rf, adc, spoil) == scan(is_rf, adc, spoil), trial
