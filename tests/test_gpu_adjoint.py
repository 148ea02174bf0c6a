"""The HIP adjoint against the FP64 oracle's autograd and finite differences (mirrors the reference's
tests/engines/test_gpu_grad.py and tests/helpers.py:77-128)."""

import math
from dataclasses import replace

import pytest
import torch

import blochparty_amd as bp
from cases import busy_sequence, map_gradients, random_case, random_phantom, rel, weighted_loss
from oracle import core

pytestmark = pytest.mark.gpu
f64 = torch.float64


def gpu_sim(ph, seq, coils=None, m_init=None, **kw):
    r = bp.simulate(ph, seq, coils, m_init=m_init, engine="gpu", **kw)
    return r.signal, r.m


def oracle_sim(ph, seq, coils=None, m_init=None, **kw):
    return core.simulate(ph, seq, coils, m_init=m_init)


def input_gradients(sim, ph, seq, coils=None, m_init=None, **kw):
    leaves = {}

    def leaf(name, t):
        t = t.detach().clone().requires_grad_()
        leaves[name] = t
        return t

    p = replace(ph, **{k: leaf(k, getattr(ph, k)) for k in ("pos", "pd", "t1", "t2", "db0")})
    if ph.b1 is not None:
        p = replace(p, b1=leaf("b1", ph.b1))
    s = replace(seq, **{k: leaf(k, getattr(seq, k)) for k in ("dt", "rf", "grad", "adc_phase")})
    if seq.b0_offset is not None:
        s = replace(s, b0_offset=leaf("b0_offset", seq.b0_offset))
    c = None if coils is None else leaf("coils", coils)
    m0 = None if m_init is None else leaf("m_init", m_init)
    sig, m = sim(p, s, c, m0, **kw)
    grads = torch.autograd.grad(weighted_loss(sig, m), list(leaves.values()), allow_unused=True)
    return dict(zip(leaves, grads))


def compare(ph, seq, coils=None, m_init=None, tol=1e-4, **kw):
    ref = input_gradients(oracle_sim, ph, seq, coils, m_init)
    out = input_gradients(gpu_sim, ph, seq, coils, m_init, **kw)
    errors = {k: rel(out[k], ref[k]) for k in ref if ref[k] is not None and ref[k].abs().max() > 0}
    print({k: f"{v:.1e}" for k, v in errors.items()}, flush=True)
    assert max(errors.values()) < tol, errors
    return errors


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_random_cases_every_input(gpu, seed):
    ph, seq = random_case(seed)
    ph = replace(ph, b1=torch.linspace(0.9, 1.1, len(ph), dtype=f64))
    seq.b0_offset = 2.0 * torch.sin(torch.arange(len(seq), dtype=f64) * 0.05)
    m0 = torch.randn(len(ph), 3, generator=torch.Generator().manual_seed(seed), dtype=f64) * 0.5
    coils = bp.birdcage_sensitivities(ph.pos, 3)
    compare(ph, seq, coils, m0)


@pytest.mark.parametrize("n_coils", [1, 3, 8, 16, 20, 33])
def test_gre_coils(gpu, n_coils):
    """Up to 16 coils in one sweep; 20 and 33 coils in sweeps over tiles of 16, summed (the sweep is linear in
    dL/dsignal; the reference tests 20 coils, tests/engines/test_gpu_grad.py:38-44)."""
    ph = bp.shepp_logan(12)
    seq = bp.gre_2d(12, n_dummy=5, te=4e-3, tr=8e-3)
    coils = bp.birdcage_sensitivities(ph.pos, n_coils)
    compare(ph, seq, coils)


def test_busy_sequence_many_intervals(gpu):
    ph = random_phantom(300, seed=2)
    seq = busy_sequence()
    coils = bp.birdcage_sensitivities(ph.pos, 4)
    compare(ph, seq, coils, checkpoint_every=64)


def test_gradients_bit_identical_for_any_checkpoint_interval(gpu):
    ph = random_phantom(700, seed=5)
    seq = busy_sequence()
    coils = bp.birdcage_sensitivities(ph.pos, 4)
    runs = [input_gradients(gpu_sim, ph, seq, coils, checkpoint_every=k) for k in (64, 128, 192)]
    for k in runs[0]:
        assert torch.equal(runs[0][k], runs[1][k]) and torch.equal(runs[0][k], runs[2][k]), k


def test_rf_only_segment_gradients_equal_the_full_ones(gpu):
    """With only rf requiring a gradient the sweep reduces segment gradients only in blocks that hold an RF record:
    rf's gradient is the same bit for bit as when every segment gradient is formed."""
    ph = random_phantom(500, seed=6)
    seq = busy_sequence()
    rf_a = seq.rf.clone().requires_grad_()
    sig, m = gpu_sim(ph, replace(seq, rf=rf_a), checkpoint_every=64)
    (g_only,) = torch.autograd.grad(weighted_loss(sig, m), [rf_a])
    rf_b, dt_b = seq.rf.clone().requires_grad_(), seq.dt.clone().requires_grad_()
    sig, m = gpu_sim(ph, replace(seq, rf=rf_b, dt=dt_b), checkpoint_every=64)
    g_full, _ = torch.autograd.grad(weighted_loss(sig, m), [rf_b, dt_b])
    assert g_only.abs().max() > 0 and torch.equal(g_only, g_full)


def test_tse_refocusing_train_and_maps_vs_finite_differences(gpu):
    """A small TSE with sinc pulses: d loss / d(refocusing flips, T1, T2, PD maps) against central differences of
    the FP64 oracle (h = 1e-3) and against the oracle's autograd."""
    maps = bp.shepp_logan_maps(16)
    flips = torch.tensor([170.0, 150.0, 160.0, 180.0], dtype=f64, requires_grad=True)
    t1 = maps["t1"].clone().requires_grad_()
    t2 = maps["t2"].clone().requires_grad_()
    pd = maps["pd"].clone().requires_grad_()
    sysm = bp.System(rf_raster=20e-6)
    g = torch.Generator().manual_seed(3)

    def run(sim, flips_, t1_, t2_, pd_):
        ph = bp.phantom_from_maps(pd_, t1_, t2_, 0.22)
        seq = bp.tse_2d(16, etl=4, tr=0.3, refocus_flip_deg=flips_, slice_thickness=5e-3, rf_duration=1e-3,
                        system=sysm)
        coils = bp.birdcage_sensitivities(ph.pos, 8)
        if sim is oracle_sim:
            s, _ = core.simulate(ph, seq, coils)
        else:
            s = bp.simulate(ph, seq, coils, engine="gpu", checkpoint_every=64).signal
        target = torch.randn(s.shape, generator=torch.Generator().manual_seed(3), dtype=torch.complex128)
        return ((s.to(torch.complex128) - target).abs() ** 2).sum()

    loss = run(gpu_sim, flips, t1, t2, pd)
    gf, g1, g2, gp = torch.autograd.grad(loss, [flips, t1, t2, pd])
    ref = torch.autograd.grad(run(oracle_sim, flips, t1, t2, pd), [flips, t1, t2, pd])
    # sequence-parameter gradients through shaped pulses sum many RF-substep terms: float32 against float64 within
    # 1e-4 (the reference accepts 3e-3 for flip gradients through sincs, its float64 core run in float32 is 5e-3 off);
    # a reverse step that rounded its factors 1 + d and cos|w| to float32 biased lambda's norm at every substep and
    # missed this by 8x
    for mine, r, name, tol in zip((gf, g1, g2, gp), ref, ("flips", "t1", "t2", "pd"), (1e-4, 1e-4, 1e-4, 1e-4)):
        e = rel(mine, r)
        print(f"{name}: rel vs oracle autograd {e:.2e}", flush=True)
        assert e < tol, name
    h = 1e-3
    with torch.no_grad():
        for i in range(4):
            d = torch.zeros(4, dtype=f64)
            d[i] = h
            fd = (run(oracle_sim, flips + d, t1, t2, pd) - run(oracle_sim, flips - d, t1, t2, pd)) / (2 * h)
            assert abs(float(fd) - float(gf[i])) < 1e-3 * float(gf.abs().max()), (i, float(fd), float(gf[i]))


def test_ernst_angle_by_gradient_descent(gpu):
    """Maximise the steady-state signal of a spoiled GRE over the flip angle with the GPU gradient: the optimum is
    the Ernst angle acos(exp(-TR / T1)) (the reference's tests/engines/test_gpu_grad.py:105-131)."""
    t1, tr, te = 0.1, 5e-3, 2e-3  # 150 TRs = 7.5 T1: the last TR is in the steady state
    ph = bp.Phantom(pos=torch.zeros(1, 3, dtype=f64), pd=torch.ones(1, dtype=f64), t1=torch.full((1,), t1, dtype=f64),
                    t2=torch.full((1,), 0.05, dtype=f64), db0=torch.zeros(1, dtype=f64))
    flip = torch.tensor(60.0, dtype=f64, requires_grad=True)
    opt = torch.optim.LBFGS([flip], lr=1.0, max_iter=50, line_search_fn="strong_wolfe")

    def closure():
        opt.zero_grad()
        b = bp.SequenceBuilder()
        for _ in range(150):
            b.rf(1e-6, flip_deg=flip).delay(te).spoil().delay(tr - te - 1e-6)
        b.rf(1e-6, flip_deg=flip).delay(te).readout(1, 1e-9)
        loss = -bp.simulate(ph, b.build(), engine="gpu").signal.abs().sum()
        loss.backward()
        return loss

    opt.step(closure)
    ernst = math.degrees(math.acos(math.exp(-tr / t1)))
    # |signal| is even in the flip angle: the descent may end at -ernst (as the reference's test allows)
    assert abs(abs(float(flip.detach())) - ernst) < 0.01, (float(flip.detach()), ernst)


# ---- reused pulses as maps in the adjoint (pulses="reuse"; the reference's reused pulses are forward only)


def pulse_case(n=8, seed=3, n_phantom=None):
    """A 2D GRE n x n with a sinc slice pulse (ramps, RF spoiling) on spins spread through the slice, B1 varying
    (the phantom n_phantom x n_phantom, default n x n)."""
    ph = bp.shepp_logan(n_phantom or n)
    g = torch.Generator().manual_seed(seed)
    pos = ph.pos.clone()
    pos[:, 2] = (torch.rand(len(ph), generator=g, dtype=f64) - 0.5) * 4e-3
    ph = replace(ph, pos=pos, b1=torch.linspace(0.9, 1.1, len(ph), dtype=f64),
                 db0=(torch.rand(len(ph), generator=g, dtype=f64) - 0.5) * 60)
    seq = bp.gre_2d(n, ramps=True, spoiling="rf", slice_thickness=5e-3, n_dummy=3)
    m0 = torch.randn(len(ph), 3, generator=g, dtype=f64) * 0.3
    return ph, seq, bp.birdcage_sensitivities(ph.pos, 3), m0


def test_reused_pulse_maps_in_the_adjoint(gpu):
    ph, seq, coils, m0 = pulse_case()
    ref = map_gradients(oracle_sim, ph, seq, coils, m0)
    out = map_gradients(gpu_sim, ph, seq, coils, m0, pulses="reuse", rf_grad="scale", checkpoint_every=64)
    step = map_gradients(gpu_sim, ph, seq, coils, m0, checkpoint_every=64)
    errors = {k: rel(out[k], ref[k]) for k in ref if ref[k].abs().max() > 0}
    errors_step = {k: rel(step[k], ref[k]) for k in ref if ref[k].abs().max() > 0}
    print("maps", {k: f"{v:.1e}" for k, v in errors.items()}, flush=True)
    print("steps", {k: f"{v:.1e}" for k, v in errors_step.items()}, flush=True)
    assert set(errors) == set(ref)  # every gradient is exercised (none is zero in the oracle)
    assert max(errors.values()) < 1e-4, errors


def test_reused_pulse_maps_in_the_adjoint_two_pulses(gpu):
    """A TSE with sinc pulses (a 90 excitation and an equal 180 refocusing pulse, each reused: two maps, two seed
    sets): every spin input's gradient and the flip scale's (one amplitude for both pulses) against the oracle."""
    maps = bp.shepp_logan_maps(12)
    ph = bp.phantom_from_maps(maps["pd"], maps["t1"], maps["t2"], 0.22)
    g = torch.Generator().manual_seed(5)
    pos = ph.pos.clone()
    pos[:, 2] = (torch.rand(len(ph), generator=g, dtype=f64) - 0.5) * 4e-3
    ph = replace(ph, pos=pos, b1=torch.linspace(0.95, 1.05, len(ph), dtype=f64))
    seq = bp.tse_2d(12, etl=4, tr=0.3, slice_thickness=5e-3, rf_duration=1e-3, system=bp.System(rf_raster=20e-6))
    from blochparty_amd.engines import pulses

    assert len(pulses.reused_pulses(seq)) == 2
    coils = bp.birdcage_sensitivities(ph.pos, 4)
    ref = map_gradients(oracle_sim, ph, seq, coils, None)
    out = map_gradients(gpu_sim, ph, seq, coils, None, pulses="reuse", rf_grad="scale", checkpoint_every=128)
    err = {k: rel(out[k], ref[k]) for k in ref if ref[k].abs().max() > 0}
    print("two pulses", {k: f"{v:.1e}" for k, v in err.items()}, flush=True)
    assert {"pos", "pd", "t1", "t2", "b1", "flip", "coils"} <= set(err)
    assert max(err.values()) < 1e-4, err


def test_reused_pulse_gradients_bit_identical_for_any_checkpoint_interval(gpu):
    """Occurrences straddle the intervals differently for each interval length (the map is applied at an
    occurrence's first segment, the rest is the identity)."""
    ph, seq, coils, m0 = pulse_case(6, seed=4)
    runs = [map_gradients(gpu_sim, ph, seq, coils, m0, pulses="reuse", rf_grad="scale", checkpoint_every=k)
            for k in (64, 192)]
    for k in runs[0]:
        assert torch.equal(runs[0][k], runs[1][k]), k


def test_reused_pulse_gradients_refuse_timing_and_gradient_inputs(gpu):
    ph, seq, coils, m0 = pulse_case(4)
    for name in ("dt", "grad"):
        s = replace(seq, **{name: getattr(seq, name).clone().requires_grad_()})
        with pytest.raises(ValueError, match="pulses='step'"):
            bp.simulate(ph, s, coils, engine="gpu", pulses="reuse")


def test_reused_pulse_rf_gradient_needs_the_scale_opt_in(gpu):
    """With maps, rf's gradient is each occurrence's complex-amplitude component only: a waveform leaf that requires
    grad is refused unless rf_grad='scale' accepts that; the spin inputs alone need no opt-in."""
    ph, seq, coils, m0 = pulse_case(4)
    s = replace(seq, rf=seq.rf.clone().requires_grad_())
    for pulses in ("reuse", "surrogate"):
        with pytest.raises(ValueError, match="rf_grad='scale'"):
            bp.simulate(ph, s, coils, engine="gpu", pulses=pulses)
    t1 = ph.t1.clone().requires_grad_()
    out = bp.simulate(replace(ph, t1=t1), seq, coils, engine="gpu", pulses="reuse")
    (out.signal.abs() ** 2).sum().backward()
    assert torch.isfinite(t1.grad).all() and t1.grad.abs().max() > 0


# ---- NUFFT readouts in the adjoint (readouts="nufft"; the reference's are forward only)


def nufft_case(n=32, seed=7, side=64):
    """A 2D GRE n x n with a hard pulse and RF spoiling (readout runs of n - 1 samples) on a side x side grid of
    random spins filling the field of view (workgroups of 256 spins sorted along the readout's phase then span a few
    grid cells: the NUFFT path, not the per-sample fallback); B1, dB0 and z varying, 4 coils, a random m_init."""
    g = torch.Generator().manual_seed(seed)
    rnd = lambda *shape: torch.rand(*shape, generator=g, dtype=f64)
    xs = (torch.arange(side, dtype=f64) + 0.5) / side * 0.22 - 0.11
    X, Y = torch.meshgrid(xs, xs, indexing="ij")
    N = side * side
    pos = torch.stack([X.flatten(), Y.flatten(), (rnd(N) - 0.5) * 4e-3], 1)
    ph = bp.Phantom(pos=pos, pd=0.5 + 0.5 * rnd(N), t1=0.4 + rnd(N), t2=0.04 + 0.1 * rnd(N), db0=(rnd(N) - 0.5) * 60,
                    b1=0.9 + 0.2 * rnd(N))
    seq = bp.gre_2d(n, spoiling="rf", n_dummy=3)
    m0 = torch.randn(N, 3, generator=g, dtype=f64) * 0.3
    return ph, seq, bp.birdcage_sensitivities(ph.pos, 4), m0


def test_nufft_readouts_in_the_adjoint(gpu):
    ph, seq, coils, m0 = nufft_case()
    from blochparty_amd.engines import driver

    case = driver.prepare_case(ph, seq, coils, nufft=(6, 3), aligned=True, chunk=128)
    assert case.nu is not None  # the readouts are NUFFT runs (not the per-sample fallback)
    ref = map_gradients(oracle_sim, ph, seq, coils, m0)
    out = map_gradients(gpu_sim, ph, seq, coils, m0, readouts="nufft", checkpoint_every=128)
    direct = map_gradients(gpu_sim, ph, seq, coils, m0, checkpoint_every=128)
    err = {k: rel(out[k], ref[k]) for k in ref if ref[k].abs().max() > 0}
    err_d = {k: rel(direct[k], ref[k]) for k in ref if ref[k].abs().max() > 0}
    print("nufft", {k: f"{v:.1e}" for k, v in err.items()}, flush=True)
    print("direct", {k: f"{v:.1e}" for k, v in err_d.items()}, flush=True)
    assert set(err) == set(ref)
    assert max(err.values()) < 1e-3, err


def test_nufft_readouts_in_the_adjoint_16_coils(gpu):
    """16 coils (the adjoint's widest register set) with NUFFT runs: every gradient against the oracle."""
    ph, seq, _, m0 = nufft_case(side=48)
    coils = bp.birdcage_sensitivities(ph.pos, 16)
    ref = map_gradients(oracle_sim, ph, seq, coils, m0)
    out = map_gradients(gpu_sim, ph, seq, coils, m0, readouts="nufft", checkpoint_every=128)
    err = {k: rel(out[k], ref[k]) for k in ref if ref[k].abs().max() > 0}
    print("16 coils nufft", {k: f"{v:.1e}" for k, v in err.items()}, flush=True)
    assert max(err.values()) < 1e-4, err


def test_nufft_readouts_in_the_adjoint_20_coils_and_reused_pulses(gpu):
    """More than one coil tile with NUFFT runs and reused pulse maps in the sweep: each tile's runs reversed from its
    own adjoint grid, the maps' seeds summed over the passes."""
    ph, seq, coils, m0 = pulse_case(8, seed=6, n_phantom=20)
    coils = bp.birdcage_sensitivities(ph.pos, 20)
    ref = map_gradients(oracle_sim, ph, seq, coils, m0)
    out = map_gradients(gpu_sim, ph, seq, coils, m0, pulses="reuse", rf_grad="scale", checkpoint_every=64)
    err = {k: rel(out[k], ref[k]) for k in ref if ref[k].abs().max() > 0}
    print("20 coils, reused pulses", {k: f"{v:.1e}" for k, v in err.items()}, flush=True)
    assert max(err.values()) < 1e-4, err
    ph2, seq2, _, m02 = nufft_case(side=48)
    coils2 = bp.birdcage_sensitivities(ph2.pos, 20)
    ref2 = map_gradients(oracle_sim, ph2, seq2, coils2, m02)
    out2 = map_gradients(gpu_sim, ph2, seq2, coils2, m02, readouts="nufft", checkpoint_every=128)
    err2 = {k: rel(out2[k], ref2[k]) for k in ref2 if ref2[k].abs().max() > 0}
    print("20 coils, nufft", {k: f"{v:.1e}" for k, v in err2.items()}, flush=True)
    assert max(err2.values()) < 1e-3, err2


def test_nufft_gradients_bit_identical_for_any_checkpoint_interval(gpu):
    """Runs straddle the intervals differently for each interval length (a run is applied at its first segment, the
    rest is the identity)."""
    ph, seq, coils, m0 = nufft_case(32, seed=8)
    runs = [map_gradients(gpu_sim, ph, seq, coils, m0, readouts="nufft", checkpoint_every=k) for k in (64, 192)]
    for k in runs[0]:
        assert torch.equal(runs[0][k], runs[1][k]), k
