"""The NUFFT readouts' host plan (engines/nufft.py), CPU only: run selection, the spins' sort and windows, and an
emulation of the kernels' use of the plan's tables (window columns, gather lists, DFT twiddles, Lagrange weights)
against the direct sum over spins and samples."""

import numpy as np
import pytest

from blochparty_amd import workloads as W
from blochparty_amd.engines import driver, nufft


def test_readout_runs_start_at_the_second_sample():
    seq = W.first_trs(W.gre3d_sequence(32), 4)
    runs = nufft.readout_runs(seq)
    assert len(runs) == 4
    adc, dt = np.asarray(seq.adc), seq.dt.numpy()
    for a, k in runs:
        assert k == 31  # 32 samples: the half-dwell first one stays per sample
        assert adc[a - 1] and dt[a - 1] < dt[a]  # the half-dwell sample before the run
        assert adc[a : a + k].all() and np.allclose(dt[a : a + k], dt[a])
    short = W.first_trs(W.gre3d_sequence(16), 2)  # 16-sample readouts leave 15: too short
    assert nufft.readout_runs(short) == []


def emulate(plan, q, r2, c, n_pad):
    """The kernels' arithmetic on the plan's tables (float64, one coil): window columns per workgroup, the cells'
    gather, the DFT at the modes, the decay terms."""
    n_blocks = n_pad // 256
    w, M, ng = plan.width, plan.terms, plan.ng
    fr = q - np.floor(q)
    x = fr * ng
    cs = c * np.exp(-2j * np.pi * fr * plan.shift)
    em = np.exp(-np.outer(r2, plan.nodes))  # (N, M)
    wb, Wd = plan.win[0, :, 0], plan.win[0, :, 1]
    part = np.zeros((n_blocks, Wd.max(), M), complex)
    for b in range(n_blocks):
        sl = slice(256 * b, min(256 * (b + 1), len(q)))
        for lc in range(Wd[b]):
            z = (wb[b] + lc - x[sl]) * 2 / w
            kv = nufft.es_kernel(z, plan.beta)
            part[b, lc] = (kv[:, None] * em[sl] * cs[sl, None]).sum(0)
    off, ent = plan.csr_off[0], plan.csr[plan.csr_base[0] :]
    grid = np.zeros((M, ng), complex)
    for u in range(ng):
        for e in range(off[u], off[u + 1]):
            grid[:, u] += part[ent[e, 0], ent[e, 1]]
    tw, lw = nufft.dft_tables(plan)
    T = tw[..., 0] + 1j * tw[..., 1]  # (n, ng)
    return (lw * (grid @ T.T)).sum(0)  # (n,)


@pytest.mark.parametrize("width,terms", [(6, 3), (8, 4)])
def test_plan_tables_reproduce_the_direct_sum(width, terms):
    seq = W.first_trs(W.gre3d_sequence(32), 2)
    plan = nufft.make_plan(seq, width, terms)
    a, n = plan.runs[0]
    rng = np.random.default_rng(0)
    N = 3000
    pos = rng.uniform(-0.11, 0.11, (N, 3))
    db0 = rng.uniform(-100, 100, N)
    order = nufft.sort_order(plan, pos, db0, seq)
    pos, db0 = pos[order], db0[order]
    n_pad = -(-N // 256) * 256
    assert nufft.windows(plan, pos, db0, seq, N, n_pad)
    q = nufft.sample_cycles(pos, db0, seq, plan.kinds[0])
    # every spin's kernel support lies in its workgroup's window
    base = np.floor((q - np.floor(q)) * plan.ng - 0.5 * width).astype(int) + 1
    blk = np.arange(N) // 256
    assert (base >= plan.win[0, blk, 0]).all()
    assert (base + width <= plan.win[0, blk, 0] + plan.win[0, blk, 1]).all()
    assert plan.win[0, :, 1].max() <= nufft.WMAX
    r2 = 1 / rng.uniform(0.03, 2.0, N)
    c = rng.normal(size=N) + 1j * rng.normal(size=N)
    t = (np.arange(n) + 1) * float(seq.dt[a])
    direct = (c[None, :] * np.exp(-2j * np.pi * np.outer(np.arange(1, n + 1), q)) * np.exp(-np.outer(t, r2))).sum(1)
    got = emulate(plan, q, r2, c, n_pad)
    err = np.abs(got - direct).max() / np.abs(direct).max()
    assert err < (3e-5 if width == 6 else 2e-6), err


def test_records_and_dispatches_keep_runs_whole():
    seq = W.first_trs(W.gre3d_sequence(32), 6)
    plan = nufft.make_plan(seq)
    rows_w = np.asarray(seq.adc, dtype=np.int64).copy()
    for a, k in plan.runs:
        rows_w[a : a + k] = 0
    starts = driver.dispatch_starts(seq, 300, 1 << 20, list(plan.runs), rows_w, [a for a, _ in plan.runs], 2)
    ends = starts[1:] + [len(seq)]
    heads = [a for a, _ in plan.runs]
    for s0, s1 in zip(starts, ends):
        assert sum(s0 <= h < s1 for h in heads) <= 2
        for a, k in plan.runs:
            assert not (a < s0 < a + k)  # never inside a run
    rec = driver.segment_table(seq, starts, runs=True, nufft=plan)
    fl = rec["flags"]
    for (a, k), kind in zip(plan.runs, plan.kind):
        assert fl[a] & driver.NUFFT and rec["pad"][a, 3] == k and rec["pad"][a, 2] == kind
        assert not (fl[a + 1 : a + k] & (driver.RUN | driver.MERGED | driver.NUFFT)).any()
    # a run whose Mxy is never read (the next record relaxes Mz only) skips the precession of its jump
    dead = [a for a, k in plan.runs if fl[a] & driver.XY_DEAD]
    assert dead and all(fl[a + k] & driver.SPOILED_FREE for a, k in plan.runs if fl[a] & driver.XY_DEAD)
    # identical runs after the first in a dispatch reuse its Mz factor; free records with the previous free one's
    # duration reuse its relaxation factors (each TR's prephaser + first sample has the same duration)
    for s0, s1 in zip(starts, ends):
        hs = [h for h in heads if s0 <= h < s1]
        assert [bool(fl[h] & driver.SAME_RUN) for h in hs] == [False] + [True] * (len(hs) - 1)
    same_dt = np.flatnonzero(fl & driver.SAME_FREE_DT)
    assert len(same_dt) > 0 and not (fl[same_dt] & (driver.RF | driver.SPOILED_FREE)).any()
    # slots count up within each dispatch
    for s0, s1 in zip(starts, ends):
        sl = [rec["pad"][h, 1] for h in heads if s0 <= h < s1]
        assert sl == list(range(len(sl)))


def test_checkpoint_tables_mark_occurrence_and_run_bodies():
    """With checkpoint intervals (fixed multiples, which may split a reused pulse or a readout run), an occurrence's
    and a run's first segment carries the record and the rest are identity bodies naming it; the kernels' start rule
    (a dispatch starting on a body starts after its record) lands on the record's end."""
    from blochparty_amd.engines import pulses

    seq = W.first_trs(W.gre3d_sequence(32), 5)
    plan = nufft.make_plan(seq)
    reused = pulses.reused_pulses(seq)
    every = 128
    starts = list(range(0, len(seq), every))
    rec = driver.segment_table(seq, starts, reused, runs=False, nufft=plan)
    fl, pad = rec["flags"], rec["pad"]
    spans = [(a, k, "pulse") for _, k, occ in reused for a, _ in occ] + [(a, k, "run") for a, k in plan.runs]
    assert reused and plan.runs
    covered = np.zeros(len(seq), dtype=bool)
    for a, k, what in spans:
        assert fl[a] & (driver.PULSE if what == "pulse" else driver.NUFFT)
        assert (fl[a + 1 : a + k] == driver.PULSE_BODY).all() and (pad[a + 1 : a + k, 3] == a).all()
        assert not covered[a : a + k].any()
        covered[a : a + k] = True
    for r, (a, k) in enumerate(plan.runs):
        assert pad[a, 4] == r and pad[a, 3] == k
    bodies = np.flatnonzero(fl & driver.PULSE_BODY)
    assert len(bodies) == sum(k - 1 for _, k, _ in spans)
    split = 0
    for s0 in starts:
        if fl[s0] & driver.PULSE_BODY:
            head = pad[s0, 3]
            length = pad[head, 2] if fl[head] & driver.PULSE else pad[head, 3]
            assert head < s0 < head + length and (head, length) in [(a, k) for a, k, _ in spans]
            split += 1
    assert split > 0  # the intervals do split records here


def test_pulse_rf_projection_and_difference_steps():
    """The rf gradient of a reused pulse's occurrence is the least-norm one with its two sums: its inner products with
    the occurrence's rf and with i rf are dL/d(ln amplitude) and dL/d(phase), it lies in their span, the other slots
    are zero. The central-difference steps turn the pulse by the stated angle and skip axes without a gradient."""
    import math

    import torch

    from blochparty_amd.engines import adjoint

    rng = np.random.default_rng(4)
    S = 40
    rf = rng.normal(size=(S, 2))
    g = torch.from_numpy(rng.normal(size=(S, 8)))
    pulses_ = [(5, 10, [(5, 0.0), (25, 1.3)])]
    c = [float(g[5, 1]), float(g[25, 1])]
    d = [float(g[5, 2]), float(g[25, 2])]
    before = g.clone()
    adjoint.project_pulse_rf(g, rf, pulses_)
    G = g.numpy()
    for (a, _), ci, di in zip(pulses_[0][2], c, d):
        z = rf[a : a + 10]
        gz = G[a : a + 10, 1:3]
        assert abs((gz * z).sum() - ci) < 1e-12 * max(1, abs(ci))
        jz = np.stack([-z[:, 1], z[:, 0]], 1)
        assert abs((gz * jz).sum() - di) < 1e-12 * max(1, abs(di))
        # in the span of rf and i rf
        coef, *_ = np.linalg.lstsq(np.stack([z.ravel(), jz.ravel()], 1), gz.ravel(), rcond=None)
        assert np.allclose(coef[0] * z + coef[1] * jz, gz, atol=1e-12)
        assert (G[a : a + 10, [0, 3, 4, 5, 6, 7]] == 0).all()
    untouched = np.r_[0:5, 15:25, 35:40]
    assert torch.equal(g[untouched], before[untouched])
    # steps: a pulse with a z gradient only, T = 1 ms, angle pi / 2 (rf in Hz: angle = 2 pi sum dt |rf|)
    K = 100
    table = np.zeros((K, 8))
    table[:, 0] = 1e-5
    table[:, 1] = 0.25 / (K * 1e-5)
    table[:, 5] = 2000.0
    st = adjoint.map_fd_steps(table, np.array([[0, K]], dtype=np.int32))[0]
    T = K * 1e-5
    assert st[0] == 0 and st[1] == 0
    assert math.isclose(2 * math.pi * T * 2000.0 * st[2], adjoint.FD_TURN)  # z: the phase over the pulse
    assert math.isclose(2 * math.pi * T * st[3], adjoint.FD_TURN)  # db0
    assert math.isclose(st[5] * T, adjoint.FD_TURN) and st[5] == st[6]  # r1, r2: the exponent
    assert math.isclose(st[7] * math.pi / 2, adjoint.FD_TURN) and st[8] == st[7]  # b1: the rotation angle


def emulate_adjoint(plan, q, r2, h):
    """The reverse sweep's arithmetic for a run (nufft_adjoint_grid_kernel, then nufft_reverse's window loop; float64,
    one coil, unit sensitivity): per spin A = dL/dc for L = Re sum_i conj(h_i) s_i, and its derivatives in R2 and in
    the fractional phase fr, as the kernel forms them."""
    w, ng = plan.width, plan.ng
    tw, lw = nufft.dft_tables(plan)
    T = tw[..., 0] + 1j * tw[..., 1]  # (n, ng)
    G = lw @ (np.conj(T) * h[:, None])  # (terms, ng): sum_i lw[m][i] conj(tw[i][u]) h_i
    fr = q - np.floor(q)
    x = fr * ng
    E = np.exp(2j * np.pi * fr * plan.shift)
    em = np.exp(-np.outer(r2, plan.nodes))  # (N, M)
    hz = 2.0 / w
    A = np.zeros(len(q), complex)
    dR = np.zeros(len(q), complex)
    dF = np.zeros(len(q), complex)
    for j in range(len(q)):
        b = j // 256
        wb, Wd = plan.win[0, b]
        Q = R = X = 0j
        for lc in range(Wd):
            z = (wb + lc - x[j]) * hz
            u = 1 - z * z
            if u <= 0:
                continue
            kv = np.exp(plan.beta * (np.sqrt(u) - 1))
            kd = -kv * plan.beta * z / np.sqrt(u)
            g = G[:, (wb + lc) % ng]
            Q += kv * (em[j] * g).sum()
            R += kv * (-plan.nodes * em[j] * g).sum()
            X += kd * (em[j] * g).sum()
        A[j] = E[j] * Q
        dR[j] = E[j] * R
        dF[j] = E[j] * (2j * np.pi * plan.shift * Q - ng * hz * X)
    return A, dR, dF


def test_nufft_adjoint_is_the_transpose_of_the_forward():
    """The reverse sweep's run arithmetic against the forward's (emulate): <s, h> = <c, A> for random weights c and
    sample gradients h (A is the exact transpose: a wrong conjugate, phase sign or cell index fails), and its R2 and
    phase derivatives against central differences of L = Re sum conj(h) s in one spin's R2 and phase."""
    seq = W.first_trs(W.gre3d_sequence(32), 2)
    plan = nufft.make_plan(seq, 6, 3)
    rng = np.random.default_rng(5)
    N = 1500
    pos = rng.uniform(-0.11, 0.11, (N, 3))
    db0 = rng.uniform(-100, 100, N)
    order = nufft.sort_order(plan, pos, db0, seq)
    pos, db0 = pos[order], db0[order]
    n_pad = -(-N // 256) * 256
    assert nufft.windows(plan, pos, db0, seq, N, n_pad)
    q = nufft.sample_cycles(pos, db0, seq, plan.kinds[0])
    r2 = 1 / rng.uniform(0.03, 2.0, N)
    c = rng.normal(size=N) + 1j * rng.normal(size=N)
    h = rng.normal(size=plan.n) + 1j * rng.normal(size=plan.n)
    s = emulate(plan, q, r2, c, n_pad)
    A, dR, dF = emulate_adjoint(plan, q, r2, h)
    lhs, rhs = (np.conj(s) * h).sum(), (np.conj(c) * A).sum()
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    L = lambda qq, rr: float(np.real((np.conj(h) * emulate(plan, qq, rr, c, n_pad)).sum()))
    for j in (3, 700, 1499):
        e = np.zeros(N)
        e[j] = 1e-4
        fd_r = (L(q, r2 + e) - L(q, r2 - e)) / 2e-4
        an_r = float(np.real(c[j] * np.conj(dR[j])))
        assert abs(fd_r - an_r) < 1e-6 * max(abs(an_r), 1e-3), (j, fd_r, an_r)
        e[j] = 1e-7
        fd_q = (L(q + e, r2) - L(q - e, r2)) / 2e-7
        an_q = float(np.real(c[j] * np.conj(dF[j])))
        assert abs(fd_q - an_q) < 1e-5 * max(abs(an_q), 1.0), (j, fd_q, an_q)
