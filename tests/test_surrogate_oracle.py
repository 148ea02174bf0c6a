"""Pins of the surrogate's FP64 oracle (oracle/surrogate.py) against the core oracle's rotation and RF step, closed
forms, finite differences and torch.optim.Adam. CPU only."""

import math

import numpy as np
import pytest
import torch

from oracle import core
from oracle import surrogate as S

f64 = torch.float64


def sinc_table(k=60, duration=1.2e-3, flip=40.0, phase=0.7):
    """(K, 3) dt, rf_x, rf_y of a windowed sinc (TBW 4) with a linear phase ramp: complex RF."""
    dt = duration / k
    t = (np.arange(k) + 0.5) * dt - 0.5 * duration
    env = (0.5 + 0.5 * np.cos(2 * np.pi * t / duration)) * np.sinc(4 / duration * t)
    amp = flip / 360.0 * env / (env.sum() * dt)
    ph = phase + 900.0 * t
    return np.stack([np.full(k, dt), amp * np.cos(ph), amp * np.sin(ph)], 1)


def test_quaternion_rotation_is_the_core_rotation():
    rng = np.random.default_rng(0)
    w = rng.normal(size=(200, 3)) * 2.0
    m = rng.normal(size=(200, 3))
    mine = S.quat_rotate(S.rot_quat(*w.T), m)
    ref = core.rotate(*(torch.tensor(a) for a in m.T), *(torch.tensor(a) for a in w.T))
    assert np.abs(mine - torch.stack(ref, 1).numpy()).max() < 1e-12
    q = S.qmul(S.rot_quat(*w[:100].T), S.rot_quat(*w[100:].T))  # composition: second rotation, then first
    both = S.quat_rotate(S.rot_quat(*w[:100].T), S.quat_rotate(S.rot_quat(*w[100:].T), m[:100]))
    assert np.abs(S.quat_rotate(q, m[:100]) - both).max() < 1e-12


def test_hard_pulse_on_resonance_closed_form():
    # one substep of area flip/360 cycles on x: rotation by -flip about x (w = -2 pi dt b1 rf)
    for flip in (10.0, 90.0, 180.0):
        dt = 1e-4
        q = S.substep_quat(dt, flip / 360.0 / dt, 0.0, np.zeros(1), np.ones(1))[0]
        th = math.radians(flip)
        assert np.allclose(q, [math.cos(th / 2), -math.sin(th / 2), 0, 0], atol=1e-14)
        m = S.quat_rotate(q, np.array([0.0, 0.0, 1.0]))
        assert np.allclose(m, [0.0, math.sin(th), math.cos(th)], atol=1e-14)  # tips toward +y


def compose_core(table, f, b1, pd, r1, r2, m0, relax=True):
    """core.rf_segment applied to every substep (rotation only when relax=False)."""
    t = lambda a: torch.as_tensor(a, dtype=f64)
    inf = np.full(len(f), np.inf)
    sp = core.Spins(pos=torch.zeros(len(f), 3, dtype=f64), db0=t(f), pd=t(pd), t1=t(1 / r1 if relax else inf),
                    t2=t(1 / r2 if relax else inf), sens_re=None, sens_im=None, b1=t(b1))
    mx, my, mz = (t(m0[:, i]).clone() for i in range(3))
    g = torch.zeros(3, dtype=f64)
    for dt, rx, ry in table:
        mx, my, mz = core.rf_segment(mx, my, mz, t(dt), t([rx, ry]), g, None, sp)
    return torch.stack([mx, my, mz], 1).numpy()


@pytest.fixture(scope="module")
def spins():
    rng = np.random.default_rng(1)
    n = 64
    return dict(f=rng.uniform(-3000, 3000, n), b1=rng.uniform(0.7, 1.3, n), pd=rng.uniform(0.5, 1.0, n),
                r1=1 / rng.uniform(0.3, 2.0, n), r2=1 / rng.uniform(0.02, 0.2, n), m0=rng.normal(size=(n, 3)) * 0.6)


def test_block_quaternions_compose_to_the_pulse_rotation(spins):
    table = sinc_table()
    k = len(table)
    for m in (1, 6, k):
        bounds = S.block_bounds(k, m)
        qs = S.block_quats(table, bounds, spins["f"], spins["b1"])
        assert np.abs(np.linalg.norm(qs, axis=-1) - 1).max() < 1e-13
        taus = np.add.reduceat(table[:, 0], bounds[:-1])
        zero = np.zeros_like(spins["f"])
        D, v = S.block_map(qs, taus, spins["f"], spins["pd"], zero, zero)  # no relaxation: D + I is the rotation
        assert np.abs(v).max() == 0
        mine = spins["m0"] + np.einsum("nij,nj->ni", D, spins["m0"])
        ref = compose_core(table, spins["f"], spins["b1"], spins["pd"], zero, zero, spins["m0"], relax=False)
        assert np.abs(mine - ref).max() < 1e-12, m


def test_b0_offset_during_the_pulse(spins):
    # a b0_offset column shifts every substep's offset: the same map as the spins' own offsets shifted
    table = sinc_table(k=30)
    b0 = np.linspace(-200.0, 300.0, len(table))
    with_b0 = np.concatenate([table, np.zeros((len(table), 3)), b0[:, None]], 1)
    bounds = S.block_bounds(len(table), 5)
    taus = np.add.reduceat(table[:, 0], bounds[:-1])
    areas = np.add.reduceat(table[:, 0] * b0, bounds[:-1])
    qs = S.block_quats(with_b0, bounds, spins["f"], spins["b1"])
    D, v = S.block_map(qs, taus, spins["f"], spins["pd"], spins["r1"], spins["r2"], areas)
    # reference: one substep per block with the offsets folded into f substep by substep (core.rf_segment's b0)
    t = lambda a: torch.as_tensor(a, dtype=f64)
    sp = core.Spins(pos=torch.zeros(len(spins["f"]), 3, dtype=f64), db0=t(spins["f"]), pd=t(spins["pd"]),
                    t1=t(1 / spins["r1"]), t2=t(1 / spins["r2"]), sens_re=None, sens_im=None, b1=t(spins["b1"]))
    m0 = spins["m0"]
    mx, my, mz = (t(m0[:, i]).clone() for i in range(3))
    for (dt, rx, ry), off in zip(table, b0):
        mx, my, mz = core.rf_segment(mx, my, mz, t(dt), t([rx, ry]), torch.zeros(3, dtype=f64), t(off), sp)
    ref = torch.stack([mx, my, mz], 1).numpy()
    exact_blocks = S.block_bounds(len(table), len(table))
    qe = S.block_quats(with_b0, exact_blocks, spins["f"], spins["b1"])
    De, ve = S.block_map(qe, table[:, 0], spins["f"], spins["pd"], spins["r1"], spins["r2"], table[:, 0] * b0)
    assert np.abs(m0 + np.einsum("nij,nj->ni", De, m0) + ve - ref).max() < 1e-12
    assert np.abs(m0 + np.einsum("nij,nj->ni", D, m0) + v - ref).max() < 1e-3  # the 5-block split


def test_block_map_with_relaxation(spins):
    table = sinc_table()
    k = len(table)
    ref = compose_core(table, spins["f"], spins["b1"], spins["pd"], spins["r1"], spins["r2"], spins["m0"])

    def err(m):
        bounds = S.block_bounds(k, m)
        qs = S.block_quats(table, bounds, spins["f"], spins["b1"])
        taus = np.add.reduceat(table[:, 0], bounds[:-1])
        D, v = S.block_map(qs, taus, spins["f"], spins["pd"], spins["r1"], spins["r2"])
        return np.abs(spins["m0"] + np.einsum("nij,nj->ni", D, spins["m0"]) + v - ref).max()

    assert err(k) < 1e-12  # one substep per block: the exact Strang sequence
    assert err(10) < 1e-5
    # second order in the block length once the blocks' precession angles are small (2 pi f tau < 1)
    assert 3.0 < err(15) / err(30) < 6.0


def test_mlp_gradient_by_finite_differences():
    rng = np.random.default_rng(2)
    h, n_out = 8, 5
    p = rng.normal(size=S.mlp_size(h, n_out, 3)) * 0.5
    x = rng.uniform(-1, 1, (40, 3))
    y = rng.normal(size=(40, n_out))
    g = S.mlp_grad(p, x, y, h, n_out)
    bad = S.mlp_grad(p, x, y, h, n_out, tanh_derivative=False)
    eps = 1e-6
    fd = np.array([(S.mlp_loss(p + eps * e, x, y, h, n_out) - S.mlp_loss(p - eps * e, x, y, h, n_out)) / (2 * eps)
                   for e in np.eye(len(p))])
    assert np.abs(g - fd).max() < 1e-7 * np.abs(fd).max()
    assert np.abs(bad - fd).max() > 1e-3 * np.abs(fd).max()  # the planted mistake is caught


def test_adam_matches_torch():
    rng = np.random.default_rng(3)
    p = rng.normal(size=30)
    tp = torch.tensor(p, requires_grad=True)
    opt = torch.optim.Adam([tp], lr=1e-2)
    m = v = np.zeros(30)
    for t in range(1, 6):
        g = rng.normal(size=30)
        p, m, v = S.adam(p, g, m, v, t, 1e-2)
        tp.grad = torch.tensor(g)
        opt.step()
    assert np.abs(p - tp.detach().numpy()).max() < 1e-12


def test_pulse_shape_key_and_norm_fold_the_amplitude():
    """Pulses that differ only in amplitude share a shape key; the per-pulse normalisation turns a spin's B1 into
    B1 x peak (CPU)."""
    import torch

    from blochparty_amd import workloads as W
    from blochparty_amd.engines import pulses
    from blochparty_amd.engines import surrogate as sur

    shapes = []
    for flip in (10.0, 13.7):
        seq = W.first_trs(W.gre3d_sequence(4, flip_deg=flip), 2)
        table, index = pulses.pulse_table(seq, pulses.reused_pulses(seq))
        off, k = index[0]
        shapes.append(sur.shape_of(table[off : off + k]))
    a, b = shapes
    assert abs(b.scale / a.scale - 13.7 / 10.0) < 1e-12
    assert a.key() == b.key()
    np.testing.assert_allclose(a.table, b.table, rtol=0, atol=1e-15)
    other = sur.shape_of(np.concatenate([a.table[:, :1], 2 * a.table[:, 1:3], a.table[:, 3:]], 1) * 1.0)
    assert other.key() == a.key()  # twice the amplitude, same shape
    bent = a.table.copy()
    bent[5, 1] *= 1.01
    assert sur.shape_of(bent).key() != a.key()
    net = sur.Net(torch.zeros(1), torch.zeros(1), torch.zeros(1), torch.tensor([1.0, 4.0, 30.0, 3.0, 1e-5, 0.0]),
                  ((1.0, 30.0), (4.0, 3.0)))
    norm = sur.pulse_norm(net, b)
    b1 = 0.93
    assert abs((b1 - float(norm[2])) / float(norm[3]) - (b1 * b.scale - 30.0) / 3.0) < 1e-5  # float32 norm
    assert float(norm[0]) == 1.0 and float(norm[1]) == 4.0 and float(norm[4]) == np.float32(1e-5)


def test_features_pinned_by_corners_harmonics_and_quarter_turns():
    head = (100.0, 400.0, 50.0, 5.0, 2e-5)
    q_max = 4
    x = S.features([100.0 - 400.0, 100.0 + 400.0], [45.0, 55.0], head, q_max)
    np.testing.assert_allclose(x[:, :2], [[-1.0, -1.0], [1.0, 1.0]], rtol=0, atol=1e-15)
    f = np.linspace(-3000.0, 3000.0, 7)
    for q in range(1, q_max + 1):
        a = S.features(f, np.full(7, 50.0), head, q_max)[:, 2 * q : 2 * q + 2]
        one = S.features(q * f, np.full(7, 50.0), head, q_max)[:, 2:4]
        np.testing.assert_allclose(a, one, rtol=0, atol=1e-12)  # harmonic q at f = harmonic 1 at q f
        quarter = S.features([1.0 / (4 * q * head[4])], [50.0], head, q_max)[0, 2 * q : 2 * q + 2]
        np.testing.assert_allclose(quarter, [1.0, 0.0], rtol=0, atol=1e-12)


def test_mlp_forward_and_grad_against_torch_nn():
    """The oracle's network against torch.nn's Linear / Tanh stack loaded with the same flat parameters (the weight
    layout W1, b1, W2, b2, W3, b3, each Linear's (out, in) weight): outputs to 1e-12, and the flat gradient of the
    mean squared error against autograd's, in the same order (CPU)."""
    import torch

    h, n_out = 16, 10
    rng = np.random.default_rng(7)
    p = rng.normal(size=S.mlp_size(h, n_out, 3)) * 0.4
    x = rng.uniform(-1, 1, (50, 3))
    y = rng.normal(size=(50, n_out))
    net = torch.nn.Sequential(torch.nn.Linear(3, h), torch.nn.Tanh(), torch.nn.Linear(h, h), torch.nn.Tanh(),
                              torch.nn.Linear(h, n_out)).double()
    with torch.no_grad():
        for t, v in zip(net.parameters(), S.mlp_unpack(p, h, n_out, 3)):
            t.copy_(torch.from_numpy(v))
    out, _ = S.mlp_forward(p, x, h, n_out)
    ref = net(torch.from_numpy(x))
    np.testing.assert_allclose(out, ref.detach().numpy(), rtol=0, atol=1e-12)
    loss = torch.nn.functional.mse_loss(ref, torch.from_numpy(y))
    loss.backward()
    auto = np.concatenate([t.grad.numpy().ravel() for t in net.parameters()])
    np.testing.assert_allclose(S.mlp_grad(p, x, y, h, n_out), auto, rtol=0, atol=1e-12)
    assert abs(S.mlp_loss(p, x, y, h, n_out) - float(loss)) < 1e-12


def test_first_order_pinned_by_a_hard_pulse_and_the_residual_order():
    """A hard pulse on resonance: the first order is exact. A shaped pulse off resonance: block_quats - first_quats
    falls by more than 3.5 per halving of b (the first order is right), while the planted mistake (the precession
    phase's sign flipped) leaves a first-order residual (about 2 per halving)."""
    k, m = 40, 4
    hard = np.zeros((k, 8))
    hard[:, 0], hard[:, 1], hard[:, 2] = 5e-6, 0.6, -0.3
    bounds = S.block_bounds(k, m)
    f0 = np.zeros(3)
    b = np.array([100.0, 250.0, 400.0])
    exact = S.block_quats(hard, bounds, f0, b)
    mine = S.first_quats(S.first_order(hard, bounds, f0), b)
    assert np.abs(exact - mine).max() < 1e-13
    from blochparty_amd import workloads as W
    from blochparty_amd.engines import pulses, surrogate as sur

    seq = W.first_trs(W.gre3d_sequence(4), 2)
    table, index = pulses.pulse_table(seq, pulses.reused_pulses(seq))
    off, kk = index[0]
    shape = sur.shape_of(table[off : off + kk])
    f = np.array([-900.0, -250.0, 130.0, 700.0])
    bounds = S.block_bounds(len(shape.table), 10)

    def residual(bb, sign=1.0):
        q = S.block_quats(shape.table, bounds, f, np.full(len(f), bb))
        q1 = S.first_quats(S.first_order(shape.table, bounds, f, sign), np.full(len(f), bb))
        return np.abs(q - q1).max()

    peak = shape.scale
    assert residual(peak) / residual(peak / 2) > 3.5
    bad = residual(peak, -1.0) / residual(peak / 2, -1.0)
    assert 1.6 < bad < 2.4  # the planted mistake is caught


def test_block_map_split_correction_pinned_against_the_substeps():
    """A slice-selective pulse at offsets up to 6.6 kHz (blocks that precess many turns) with T2 = 80 ms: the corrected
    split is at least 10x closer to the substep-exact map than the plain one and at least second order in the block
    length, and the planted mistake (the correction's sign flipped) is further off than the plain split. The
    symmetric slab pulse near its band: within 1e-6 of the substeps."""
    from blochparty_amd import workloads as W
    from blochparty_amd.engines import pulses, surrogate as sur

    seq = W.multislice_sequence(16, lines=1, slices=5)
    table, index = pulses.pulse_table(seq, pulses.reused_pulses(W.concat([seq, seq])))
    off, k = index[4]
    shape = sur.shape_of(table[off : off + k])
    f = np.linspace(-6600.0, 6600.0, 133)
    n = len(f)
    b = np.full(n, shape.scale)
    pd, r1, r2 = np.ones(n), np.full(n, 1 / 0.8), np.full(n, 1 / 0.08)

    def split(m, moments=True, sign=1.0):
        bounds = S.block_bounds(k, m)
        q = S.block_quats(shape.table, bounds, f, b)
        taus = np.add.reduceat(shape.table[:, 0], bounds[:-1])
        areas = np.add.reduceat(shape.table[:, 0] * shape.table[:, 6], bounds[:-1])
        mom = None
        if moments:
            mom = (S.first_moment(shape.table, bounds, f) * b[:, None, None],
                   S.second_moment(shape.table, bounds, f) * (b * b)[:, None, None])
        return S.block_map(q, taus, f, pd, r1, r2, areas, mom, sign)

    exact = split(k, moments=False)  # one substep per block: the substeps themselves
    err = lambda out: max(np.abs(out[0] - exact[0]).max(), np.abs(out[1] - exact[1]).max())
    plain, fixed = err(split(10, moments=False)), err(split(10))
    assert fixed < plain / 10
    assert fixed / err(split(20)) > 3.0
    assert err(split(10, sign=-1.0)) > plain  # the planted mistake is caught
    slab_seq = W.first_trs(W.gre3d_sequence(8), 2)
    table, index = pulses.pulse_table(slab_seq, pulses.reused_pulses(slab_seq))
    off, k = index[0]
    shape = sur.shape_of(table[off : off + k])
    f = np.linspace(-1300.0, 1300.0, 53)
    n = len(f)
    b = np.full(n, 1.3 * shape.scale)
    pd, r1, r2 = np.ones(n), np.full(n, 1 / 0.5), np.full(n, 1 / 0.05)
    exact = split(k, moments=False)
    assert err(split(10)) < 1e-6
