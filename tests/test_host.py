"""Host side of the GPU engine on the CPU: the C-ABI library loads and exports what the header declares, the
matrix-core operand layout and the dispatch plan are right, and the API refuses what it should. No GPU calls."""

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest
import torch

import blochparty_amd as bp
from blochparty_amd.engines import capabilities, driver, hip
from cases import busy_sequence, random_phantom

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    out = set()
    for h in (ROOT / "include").glob("*.h"):
        out |= set(re.findall(r"\b(hb_\w+)\s*\(", h.read_text()))
    return out


def test_library_exports_every_declared_symbol():
    if not hip.LIB_PATH.exists():
        pytest.skip("library not built (run python -m blochparty_amd.build)")
    lib = ctypes.CDLL(str(hip.LIB_PATH))
    names = declared_symbols()
    assert {"hb_forward", "hb_signal_reduce", "hb_error_string", "hb_abi_version"} <= names
    for name in names:
        assert hasattr(lib, name), name
    assert hip.load().hb_abi_version() == hip.ABI_VERSION
    assert lib.hb_error_string  # callable without a device


def test_segment_record_layout():
    assert driver.SEG_DTYPE.itemsize == 128
    assert driver.SEG_DTYPE.fields["d"][1] == 32 and driver.SEG_DTYPE.fields["flags"][1] == 72
    assert driver.SEG_DTYPE.fields["w"][1] == 96


def mfma_16x16x4(a_lanes, b_lanes):
    """v_mfma_f32_16x16x4f32 on one wave: lane l gives A[l % 16][l // 16] and B[l // 16][l % 16]; returns D (16, 16)."""
    A = np.zeros((16, 4))
    B = np.zeros((4, 16))
    for lane in range(64):
        A[lane % 16, lane // 16] = a_lanes[lane]
        B[lane // 16, lane % 16] = b_lanes[lane]
    return A @ B


@pytest.mark.parametrize("n, C", [(200, 13), (64, 8), (100, 1)])
def test_sensitivity_operand_order_gives_the_coil_sums(n, C):
    """The kernel's flush, emulated lane by lane: staged Mxy columns, the operand from sens_operand, the P / Q
    combination across lanes; must equal sum_spins S_c Mxy."""
    nt = -(-C // 8)
    n_pad = 256
    g = torch.Generator().manual_seed(0)
    coils = torch.randn(n, C, dtype=torch.complex128, generator=g)
    A = driver.sens_operand(coils, n, n_pad, nt).double().numpy()
    M = np.zeros((n_pad, 8), complex)
    M[:n] = torch.randn(n, 8, dtype=torch.complex128, generator=g).numpy()
    out = np.zeros((8, 8 * nt), complex)
    for w in range(n_pad // 64):
        stg = np.concatenate([M[64 * w: 64 * w + 64].real.T, M[64 * w: 64 * w + 64].imag.T])  # (16 cols, 64)
        for t in range(nt):
            D = sum(mfma_16x16x4(A[w, t, u], [stg[lane % 16, 16 * (lane // 16) + u] for lane in range(64)])
                    for u in range(16))
            acc = np.array([[D[4 * (lane // 16) + r, lane % 16] for r in range(4)] for lane in range(64)])
            for lane in range(64):  # rows of lane group gg: Re S_2gg, Im S_2gg, Re S_2gg+1, Im S_2gg+1
                jj, gg = lane % 16, lane // 16
                rot = acc[16 * gg + (jj + 8) % 16]  # DPP row_ror:8
                if jj < 8:
                    p, q = acc[lane, 0] - rot[1], rot[0] + acc[lane, 1]
                else:
                    p, q = rot[2] - acc[lane, 3], acc[lane, 2] + rot[3]
                out[jj & 7, 8 * t + 2 * gg + (jj >> 3)] += p + 1j * q
    ref = M[:n].T @ np.concatenate([coils.numpy(), np.zeros((n, 8 * nt - C))], 1)
    assert np.abs(out - ref).max() < 1e-5 * np.abs(ref).max()


def test_dispatch_plan_and_cache_hints():
    seq = busy_sequence()
    starts = driver.dispatch_starts(seq, 7, 8)
    ends = starts[1:] + [len(seq)]
    adc = np.asarray(seq.adc)
    assert starts[0] == 0 and all(e > s for s, e in zip(starts, ends))
    assert all(adc[s:e].sum() <= 8 for s, e in zip(starts, ends))
    rec = driver.segment_table(seq, starts)
    fl = rec["flags"]
    for s in starts:  # no cache hint may reach across a dispatch start
        nxt_free = next((i for i in range(s, len(seq)) if not seq.is_rf[i]), None)
        nxt_rf = next((i for i in range(s, len(seq)) if seq.is_rf[i]), None)
        if nxt_free is not None and all(i >= s for i in [nxt_free]):
            assert not (fl[nxt_free] & driver.SAME_FREE) or any(
                not seq.is_rf[i] for i in range(s, nxt_free)), s
        if nxt_rf is not None:
            assert not (fl[nxt_rf] & driver.SAME_RF_DT) or any(seq.is_rf[i] for i in range(s, nxt_rf))
    psi = rec["f"][:, 7]
    assert (psi >= 0).all() and (psi < 2 * np.pi + 1e-6).all()


def walk(rec, s0, s1):
    """The ADC rows a dispatch stages, in the forward kernel's control flow (runs, pulses, single segments)."""
    rows, s = [], s0
    while s < s1:
        f = rec["flags"][s]
        if f & driver.RUN:
            L = rec["pad"][s, 3]
            rows += list(range(rec["pad"][s, 0], rec["pad"][s, 0] + L))
            s += L
        elif f & driver.PULSE:
            s += rec["pad"][s, 2]
        elif f & driver.TRAIN:
            L = rec["pad"][s, 3]
            assert not (rec["flags"][s : s + L] & (driver.ADC | driver.SPOIL)).any()
            assert (rec["flags"][s : s + L] & driver.RF).all()
            s += L
        elif f & driver.MERGED:  # free segments, sampled at most at the end
            L = rec["pad"][s, 3]
            assert not (rec["flags"][s : s + L] & driver.RF).any()
            assert not (rec["flags"][s + 1 : s + L - 1] & driver.ADC).any()
            assert not (f & driver.ADC and f & driver.SPOILED_FREE)
            if f & driver.ADC:
                assert rec["flags"][s + L - 1] & driver.ADC
                rows.append(rec["pad"][s, 0])
            s += L
        else:
            if f & driver.ADC:
                rows.append(rec["pad"][s, 0])
            s += 1
    assert s == s1
    return rows


@pytest.mark.parametrize("chunk", [10**6, 7, 64, 150])
def test_readout_runs_and_rf_trains_cover_every_sample_once(chunk):
    from blochparty_amd import workloads

    seq = workloads.first_trs(workloads.gre3d_sequence(8), 6)  # 200-substep sinc pulses, 8-sample readouts
    starts = driver.dispatch_starts(seq, chunk, 1 << 20)
    rec = driver.segment_table(seq, starts, runs=True)
    fl, adc = rec["flags"], np.asarray(seq.adc)
    ends = starts[1:] + [len(seq)]
    assert sum((walk(rec, a, b) for a, b in zip(starts, ends)), []) == list(range(int(adc.sum())))
    key = np.stack([seq.dt.numpy(), *seq.grad.numpy().T], 1)
    first = np.flatnonzero(fl & driver.RUN)
    assert len(first) > 0 if chunk > 16 else True
    for s in first:
        L = rec["pad"][s, 3]
        assert L >= driver.MIN_RUN
        assert not np.asarray(seq.is_rf)[s : s + L].any() and adc[s : s + L].all()
        assert not np.asarray(seq.spoil)[s : s + L].any()
        assert np.allclose(key[s : s + L], key[s], rtol=1e-12, atol=0)
        assert not any(a in range(s + 1, s + L) for a in starts)  # never across a dispatch start
    # merged free runs: the head carries the sums of its segments' duration and moments; only the last spoiled or
    # sampled, unless the run never reaches a sample (SPOILED_FREE)
    raw = driver.segment_table(seq, [0])
    spoil, is_rf = np.asarray(seq.spoil), np.asarray(seq.is_rf)
    for s in np.flatnonzero(fl & driver.MERGED):
        L = rec["pad"][s, 3]
        assert L >= 2 and not is_rf[s : s + L].any() and not adc[s : s + L - 1].any()
        np.testing.assert_allclose(rec["d"][s], raw["d"][s : s + L].sum(0), rtol=1e-12, atol=1e-18)
        assert np.float32(rec["d"][s, 0]) == rec["f"][s, 0]
        if not fl[s] & driver.SPOILED_FREE:
            assert not spoil[s : s + L - 1].any()
            assert bool(fl[s] & driver.SPOIL) == bool(spoil[s + L - 1])
            assert bool(fl[s] & driver.ADC) == bool(adc[s + L - 1])
        assert not any(a in range(s + 1, s + L) for a in starts)
    assert ((fl & driver.MERGED != 0) & (fl & driver.ADC != 0)).any() if chunk > 16 else True
    # SPOILED_FREE: no sample and no RF; spoiled at its end or starting right after a spoil (Mxy = 0)
    zapped = np.flatnonzero(fl & driver.SPOILED_FREE)
    assert len(zapped) > 0 and not (fl[zapped] & (driver.ADC | driver.RF | driver.SAME_FREE)).any()
    for s in zapped:
        L = rec["pad"][s, 3] if fl[s] & driver.MERGED else 1
        assert not adc[s : s + L].any() and not is_rf[s : s + L].any()
        assert spoil[s + L - 1] or (s > 0 and spoil[s - 1]) or spoil[s : s + L].any()
    # without runs (the adjoint's replay) no record carries the run or merge bits; RF trains are marked either way
    plain = driver.segment_table(seq, starts)
    assert not (plain["flags"] & (driver.RUN | driver.SAME_RUN | driver.MERGED | driver.SPOILED_FREE)).any()
    assert np.array_equal(plain["flags"] & driver.TRAIN, fl & driver.TRAIN)
    assert sum((walk(plain, a, b) for a, b in zip(starts, ends)), []) == list(range(int(adc.sum())))
    if chunk > 16:
        assert (fl & driver.TRAIN).any()
    assert (rec["pad"][fl & driver.TRAIN != 0, 3] >= driver.MIN_TRAIN).all()


def test_case_packing_splits_float64_to_48_bits():
    ph = random_phantom(300, seed=1)
    m0 = torch.randn(300, 3, dtype=torch.float64)
    case = driver.prepare_case(ph, busy_sequence(), None, m0, device="cpu")
    st = case.state.double()
    assert ((st[:3, :300] + st[3:, :300]) - m0.T).abs().max() < 2.0**-46 * m0.abs().max()
    rec = case.segs.numpy().view(driver.SEG_DTYPE)
    dt = busy_sequence().dt.numpy()
    np.testing.assert_allclose(rec["w"][:, 0], -2 * np.pi * dt, rtol=1e-7)
    pos = case.spins[:3, :300].double() + case.pos_lo[:, :300].double()
    assert (pos - ph.pos.T).abs().max() < 2.0**-46 * ph.pos.abs().max()
    assert torch.all(case.spins[3, 300:] == 0) and torch.all(case.state[:, 300:] == 0)


def test_api_refusals_without_gpu():
    ph = random_phantom(10)
    seq = busy_sequence()
    with pytest.raises(ValueError):
        bp.simulate(ph, seq, engine="core")
    with pytest.raises(ValueError):
        bp.simulate(ph, seq, engine="tpu")
    if not bp.available():
        with pytest.raises(RuntimeError):
            bp.simulate(ph, seq, engine="gpu")
    with pytest.raises(NotImplementedError):
        capabilities.check("hip", capabilities.features(bp.Phantom(ph.pos, ph.pd, ph.t1, ph.t2, ph.db0,
                                                                   diffusion=torch.ones(10)), seq))
    capabilities.check("hip", frozenset({"b1", "m_init"}), grad=True)


def test_output_guard_raises_on_non_finite_results_of_finite_inputs():
    import dataclasses
    import types

    from blochparty_amd.exchange import BoundPool

    ph, seq = random_phantom(10), busy_sequence()
    coils = torch.ones(10, 2, dtype=torch.complex128)
    assert driver.finite_inputs(ph, seq, coils)
    t1 = ph.t1.clone()
    t1[3] = float("nan")
    assert not driver.finite_inputs(dataclasses.replace(ph, t1=t1), seq, coils)
    rf = seq.rf.clone()
    rf[0, 1] = float("inf")
    assert not driver.finite_inputs(ph, dataclasses.replace(seq, rf=rf), coils)
    assert not driver.finite_inputs(ph, seq, coils * float("nan"))
    assert not driver.finite_inputs(dataclasses.replace(ph, bound=BoundPool(0.1, float("inf"), 1.0, 1e-5)), seq)
    assert driver.finite_inputs(dataclasses.replace(ph, bound=BoundPool(0.1, 20.0, 1.0, 1e-5)), seq)

    bad, good = torch.tensor([1.0, float("nan")]), torch.ones(3)
    case = types.SimpleNamespace(extra={"finite_inputs": True})
    with pytest.raises(RuntimeError, match="non-finite"):
        driver.check_output(case, good, bad)
    driver.check_output(case, good, good)
    driver.check_output(case, bad, premise=False)  # a non-finite incoming gradient
    driver.check_output(types.SimpleNamespace(extra={"finite_inputs": False}), bad)


def test_default_checkpoint_interval():
    """About sqrt(segments) in multiples of 64; 4x that when reused pulses or NUFFT runs make most segments identity
    records (the 64^3 GRE: 1,110,016 segments -> 1,088 and 4,352)."""
    from blochparty_amd.engines import adjoint

    assert adjoint.default_checkpoint_every(1_110_016) == 1088
    assert adjoint.default_checkpoint_every(1_110_016, bodies=True) == 4352
    assert adjoint.default_checkpoint_every(10) == 64 and adjoint.default_checkpoint_every(0) == 64
    for n in (100, 5000, 123457):
        e = adjoint.default_checkpoint_every(n)
        assert e % 64 == 0 and e >= n**0.5 and e - 64 < max(n**0.5, 64)


def test_bench_runs_on_one_gpu_only():
    """bench.py --gpus N > 1 exits with a message (this build measures one GPU's share)."""
    import subprocess
    import sys
    from pathlib import Path

    bench = Path(__file__).resolve().parent.parent / "bench.py"
    r = subprocess.run([sys.executable, str(bench), "--gpus", "2"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 2 and "one GPU" in r.stderr, (r.returncode, r.stderr[-300:])
