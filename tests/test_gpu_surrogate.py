"""The learned RF surrogate on the GPU: its kernels against the FP64 oracle (oracle/surrogate.py), the trained
network against the exact composition (the judge), the end-to-end signal against the FP64 core oracle, and an
ablation (what training buys; what the network saves over composing the substeps)."""

import numpy as np
import pytest
import torch

import blochparty_amd as bp
from blochparty_amd import workloads as W
from blochparty_amd.engines import driver, pulses, surrogate as sur
from cases import map_gradients, rel
from oracle import core
from oracle import surrogate as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def slab():
    """The 3D GRE's reused slab pulse (2 ms sinc, TBW 4, 200 substeps) as the surrogate sees it, on a 16^3 grid."""
    ph, seq, coils = W.gre3d(16)
    reused = pulses.reused_pulses(seq)
    table, index = pulses.pulse_table(seq, reused)
    off, k = index[0]
    return ph, seq, coils, sur.shape_of(table[off : off + k])


def box_inputs(box, n, seed):
    """(n, 2) float32 inputs (f, b1 x peak) drawn uniformly from a box."""
    g = np.random.default_rng(seed)
    c, h = np.array(box[0]), np.array(box[1])
    return ((g.uniform(-1, 1, (n, 2)) * h + c)).astype(np.float32)


def net_features(net, x):
    """The oracle's feature map of float32 inputs x (n, 2) with the network's float32 normalisation head."""
    head = net.norm[:5].double().cpu().numpy()
    return S.features(x[:, 0].astype(np.float64), x[:, 1].astype(np.float64), head, sur.FOURIER)


def test_spinor_blocks_match_the_oracle(gpu, slab):
    *_, shape = slab
    x = box_inputs(((0.0, 1.0), (1350.0, 0.3)), 2000, 0)
    y = sur.spinor_blocks(shape, torch.from_numpy(x).cuda()).cpu().numpy()
    xd = x.astype(np.float64)
    ref = S.block_quats(shape.table, shape.bounds, xd[:, 0], xd[:, 1]).reshape(len(x), -1)
    assert np.abs(y - ref).max() < 2e-7  # float64 products rounded to float32


def test_first_order_table_and_quaternions_match_the_oracle(gpu, slab):
    """hb_sur_first_order (F_j, H_j, G_j) at its grid points and hb_sur_first_quats between them (linear
    interpolation on 2049 points) against the oracle's float64 first order and time moments."""
    *_, shape = slab
    box = ((-40.0, shape.scale), (1300.0, 0.3 * shape.scale))
    first = sur.first_order(shape, box, "cuda")
    lo, hi = np.float32(box[0][0]) - np.float32(box[1][0]), np.float32(box[0][0]) + np.float32(box[1][0])
    grid = np.linspace(float(lo), float(hi), sur.FIRST_GRID)
    mine = first.double().cpu().numpy().transpose(1, 0, 2)  # (G, M, 4)
    ref = S.first_order(shape.table, shape.bounds, grid)  # (G, M, 2)
    assert np.abs(mine[..., :2] - ref).max() < 1e-6 * np.abs(ref).max()
    ref_h = S.first_moment(shape.table, shape.bounds, grid)
    assert np.abs(mine[..., 2:4] - ref_h).max() < 1e-6 * np.abs(ref_h).max()
    ref_g = S.second_moment(shape.table, shape.bounds, grid)
    assert np.abs(mine[..., 4:7] - ref_g).max() < 1e-5 * np.abs(ref_g).max()
    assert not mine[..., 7].any()
    net = sur.new_net(box, shape.fourier_base, torch.zeros(sur.N_OUT), torch.ones(sur.N_OUT), 0, "cuda", first)
    x = box_inputs(box, 3000, 11)
    q = sur.first_quats(net, torch.from_numpy(x).cuda()).double().cpu().numpy().reshape(len(x), -1, 4)
    xd = x.astype(np.float64)
    qr = S.first_quats(S.first_order(shape.table, shape.bounds, xd[:, 0]), xd[:, 1])
    assert np.abs(q - qr).max() < 2e-6


def random_net(seed=1):
    box = ((-17.0, 1.0), (1020.0, 0.25))
    rng = np.random.default_rng(seed)
    mu = torch.from_numpy(rng.normal(size=sur.N_OUT)).float()
    sd = torch.from_numpy(rng.uniform(0.01, 0.5, sur.N_OUT)).float()
    net = sur.new_net(box, 2.5e-5, mu.cuda(), sd.cuda(), seed, "cuda")
    net.params.copy_(torch.from_numpy(rng.normal(size=sur.N_PARAMS) * 0.3).float())
    return net, box, mu.double().numpy(), sd.double().numpy()


def test_mlp_eval_and_gradient_match_the_oracle(gpu):
    net, box, mu, sd = random_net()
    n = 3000  # more than one 512-sample chunk and a ragged last one
    x = box_inputs(box, n, 2)
    y = np.random.default_rng(3).normal(size=(n, sur.N_OUT)).astype(np.float32)
    p = net.params.double().cpu().numpy()
    xn = net_features(net, x)
    out, _ = S.mlp_forward(p, xn, sur.HIDDEN, sur.N_OUT)
    mine = sur.evaluate(net, torch.from_numpy(x).cuda()).double().cpu().numpy()
    assert rel(torch.from_numpy(mine), torch.from_numpy(mu + sd * out)) < 1e-5
    g = sur.gradient(net, torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()).double().cpu().numpy()
    ref = S.mlp_grad(p, xn, (y - mu) / sd, sur.HIDDEN, sur.N_OUT)
    assert np.abs(g - ref).max() < 2e-5 * np.abs(ref).max()
    # deterministic: the same gradient bit for bit
    g2 = sur.gradient(net, torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()).double().cpu().numpy()
    assert np.array_equal(g, g2)


def test_adam_steps_match_the_oracle(gpu):
    import ctypes

    from blochparty_amd.engines import hip

    net, box, mu, sd = random_net(4)
    n = 1024
    x = torch.from_numpy(box_inputs(box, n, 5)).cuda()
    y = torch.from_numpy(np.random.default_rng(6).normal(size=(n, sur.N_OUT)).astype(np.float32)).cuda()
    lib = hip.load()
    work = torch.empty(lib.hb_mlp_workspace(n), dtype=torch.float32, device="cuda")
    p = net.params.double().cpu().numpy()
    m = v = np.zeros_like(p)
    for t, lr in ((1, 1e-3), (2, 5e-4), (3, 2e-4)):
        g = sur.gradient(net, x, y).double().cpu().numpy()
        p, m, v = S.adam(p, g, m, v, t, lr)
        rc = lib.hb_mlp_train_step(ctypes.byref(net.struct()), x.data_ptr(), y.data_ptr(), n, work.data_ptr(), t, lr,
                                   sur.stream())
        assert rc == 0
        mine = net.params.double().cpu().numpy()
        assert np.abs(mine - p).max() < 1e-6, t
        p = mine  # continue from the device's float32 parameters


@pytest.fixture(scope="module")
def trained(slab):
    from blochparty_amd.engines import hip

    if not hip.available():
        pytest.skip("needs a ROCm GPU and the built library")
    ph, seq, coils, shape = slab
    case = driver.prepare_case(ph, seq, coils, reuse=True)
    box = sur.padded_box(sur.input_range(case.spins, case.pos_lo, case.n, shape.g, shape.scale))
    return sur.train(shape, box, dev="cuda"), box, shape


def test_trained_network_against_the_exact_maps(gpu, trained):
    net, box, shape = trained
    print(f"trained in {net.train_s:.1f} s, held-out loss {net.final_loss:.2e}")
    assert net.final_loss < 1e-3
    x = box_inputs(box, 4000, 7)
    xd = x.astype(np.float64)
    f = xd[:, 0]
    q = sur.block_quats(net, torch.from_numpy(x).cuda()).double().cpu().numpy().reshape(len(x), -1, 4)
    rng = np.random.default_rng(8)
    pd, r1, r2 = rng.uniform(0.7, 1.0, len(x)), 1 / rng.uniform(0.8, 4.0, len(x)), 1 / rng.uniform(0.08, 2, len(x))
    D, v = S.block_map(q, shape.blocks[:, 0], f, pd, r1, r2, shape.blocks[:, 1])
    exact_q = S.block_quats(shape.table, shape.bounds, f, xd[:, 1])
    De, ve = S.block_map(exact_q, shape.blocks[:, 0], f, pd, r1, r2, shape.blocks[:, 1])
    err = max(np.abs(D - De).max(), np.abs(v - ve).max())
    print(f"map error on {len(x)} box inputs: {err:.2e}")
    assert err < 1e-3


def test_surrogate_simulation_matches_the_oracle(gpu, trained, slab):
    ph, seq, coils, _ = slab
    seq4 = W.first_trs(seq, 4)
    res = bp.simulate(ph, seq4, coils, engine="gpu", pulses="surrogate")
    rep = res.surrogate
    print(rep)
    assert rep["covered"] == 1 and not rep["rejected"] and rep["pulses"][0]["outside_box"] == 0
    assert rep["pulses"][0]["judge_max_map_err"] < 1e-3
    ref_sig, ref_m = core.simulate(ph, seq4, coils)
    e_sig = rel(res.signal, ref_sig)
    exact = bp.simulate(ph, seq4, coils, engine="gpu", pulses="reuse")
    print(f"k-space error vs oracle: surrogate {e_sig:.2e}, exact maps {rel(exact.signal, ref_sig):.2e}")
    assert e_sig < 1e-4
    assert rel(res.m, ref_m) < 1e-3


def test_gradients_through_the_surrogate_maps(gpu, trained, slab):
    """Forward + adjoint with the network's maps (each occurrence one map in both sweeps; the maps' parameter
    derivatives from the exact float64 composition): every gradient against the oracle's float64 autograd through
    all 200 substeps, on a seeded subset of the slab's spins over 4 TRs."""
    ph, seq, coils, _ = slab
    ph_s, c_s = W.subset(ph, coils, 1024)
    seq4 = W.first_trs(seq, 4)

    def gpu(p, s, c, m0, **kw):
        r = bp.simulate(p, s, c, m_init=m0, engine="gpu", **kw)
        return r.signal, r.m

    ref = map_gradients(lambda p, s, c, m0: core.simulate(p, s, c, m_init=m0), ph_s, seq4, c_s, None)
    sur_g = map_gradients(gpu, ph_s, seq4, c_s, None, pulses="surrogate", rf_grad="scale", checkpoint_every=128)
    exact = map_gradients(gpu, ph_s, seq4, c_s, None, pulses="reuse", rf_grad="scale", checkpoint_every=128)
    err = {k: rel(sur_g[k], ref[k]) for k in ref if ref[k].abs().max() > 0}
    err_x = {k: rel(exact[k], ref[k]) for k in ref if ref[k].abs().max() > 0}
    print("surrogate", {k: f"{v:.1e}" for k, v in err.items()}, flush=True)
    print("exact maps", {k: f"{v:.1e}" for k, v in err_x.items()}, flush=True)
    assert max(err_x.values()) < 1e-4, err_x
    assert max(err.values()) < 1e-3, err


def test_gradients_through_the_surrogate_maps_and_nufft_readouts(gpu):
    """The headline forward + adjoint configuration: the network's maps and NUFFT readouts in both sweeps, every
    gradient against the oracle's float64 autograd, on 8,192 seeded spins of the 3D GRE 24^3 (readout runs of 23
    samples) over 4 TRs."""
    ph, seq, coils = W.gre3d(24)
    ph_s, c_s = W.subset(ph, coils, 8192, seed=3)
    seq4 = W.first_trs(seq, 4)
    case = driver.prepare_case(ph_s, seq4, c_s, nufft=(6, 3), aligned=True, chunk=256, surrogate=True)
    assert case.nu is not None and case.extra["surrogate"]  # NUFFT runs, and the network covers the pulse

    def gpu(p, s, c, m0, **kw):
        r = bp.simulate(p, s, c, m_init=m0, engine="gpu", **kw)
        return r.signal, r.m

    ref = map_gradients(lambda p, s, c, m0: core.simulate(p, s, c, m_init=m0), ph_s, seq4, c_s, None)
    kw = dict(rf_grad="scale", readouts="nufft", checkpoint_every=256)
    sur_nu = map_gradients(gpu, ph_s, seq4, c_s, None, pulses="surrogate", **kw)
    exact_nu = map_gradients(gpu, ph_s, seq4, c_s, None, pulses="reuse", **kw)
    err = {k: rel(sur_nu[k], ref[k]) for k in ref if ref[k].abs().max() > 0}
    err_x = {k: rel(exact_nu[k], ref[k]) for k in ref if ref[k].abs().max() > 0}
    print("surrogate + nufft", {k: f"{v:.1e}" for k, v in err.items()}, flush=True)
    print("exact maps + nufft", {k: f"{v:.1e}" for k, v in err_x.items()}, flush=True)
    assert max(err_x.values()) < 1e-4, err_x
    assert max(err.values()) < 1e-3, err


def test_spins_outside_the_box_fall_back_to_the_exact_maps(gpu, trained, slab):
    ph, seq, coils, _ = slab
    seq2 = W.first_trs(seq, 2)
    far = bp.Phantom(pos=ph.pos.clone(), pd=ph.pd, t1=ph.t1, t2=ph.t2, db0=ph.db0.clone())
    far.db0[::3] += 5000.0  # far outside any box trained on the unshifted phantom
    case = driver.prepare_case(far, seq2, coils, surrogate=True)
    driver.run_forward(case)
    out = driver.read_result(case)
    rep = out.surrogate["pulses"][0]
    assert rep["outside_box"] == 0  # the box follows the phantom: a new network covers the shifted spins
    ex = bp.simulate(far, seq2, coils, engine="gpu", pulses="reuse")
    assert rel(out.signal, ex.signal) < 1e-3
    # with the network of the unshifted phantom the shifted spins lie outside its box: exact maps for them
    net, box, shape = trained
    case.extra["surrogate"] = [(0, shape, net)]
    driver.run_forward(case)
    out2 = driver.read_result(case)
    assert out2.surrogate["pulses"][0]["outside_box"] == (len(far) + 2) // 3
    assert rel(out2.signal, ex.signal) < 1e-3


def test_ablation_training_and_cost(gpu, trained, slab):
    """What the trained network gains: against the same network untrained, the map error; against composing the
    200 substeps per spin exactly (the method without it), the time to form every spin's maps."""
    ph, seq, coils, _ = slab
    net, box, shape = trained
    case = driver.prepare_case(ph, seq, coils, surrogate=True)
    driver.run_forward(case)
    trained_err = driver.read_result(case).surrogate["pulses"][0]["judge_max_map_err"]
    mu, sd = net.norm[sur.NORM_HEAD : sur.NORM_HEAD + sur.N_OUT], net.norm[sur.NORM_HEAD + sur.N_OUT :]
    raw = sur.new_net(box, shape.fourier_base, mu, sd, 0, "cuda", net.first)
    case.extra["surrogate"] = [(0, shape, raw)]
    driver.run_forward(case)
    raw_err = driver.read_result(case).surrogate["pulses"][0]["judge_max_map_err"]
    first_only = sur.new_net(box, shape.fourier_base, mu, sd, 0, "cuda", net.first)
    first_only.params.zero_()  # outputs = the residuals' means: the first-order part alone
    case.extra["surrogate"] = [(0, shape, first_only)]
    driver.run_forward(case)
    first_err = driver.read_result(case).surrogate["pulses"][0]["judge_max_map_err"]
    print(f"judge map error: trained {trained_err:.2e}, first order alone {first_err:.2e}, untrained {raw_err:.2e}")
    assert raw_err > 100 * trained_err
    assert first_err > 10 * trained_err
    # time: network maps vs exact substep composition over a 1 M-spin case
    big_ph, big_seq, big_c = W.gre3d(64)
    big = driver.prepare_case(big_ph, W.first_trs(big_seq, 2), big_c, reuse=True)
    from blochparty_amd.engines import hip
    import ctypes

    lib = hip.load()
    st = big.struct()
    outside = torch.zeros(big.n_pad, dtype=torch.int32, device="cuda")
    blocks = torch.from_numpy(shape.blocks).cuda()
    sp = sur.HbSurPulse(blocks.data_ptr(), (ctypes.c_double * 3)(*shape.g), 0, 0)

    def timed(fn):
        fn()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(3):
            fn()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / 3

    t_exact = timed(lambda: lib.hb_pulse_maps(ctypes.byref(st), big.pulse_table.data_ptr(), big.pulse_index.data_ptr(),
                                              big.n_pulses, None, 0, big.maps.data_ptr(), sur.stream()))
    norm = sur.pulse_norm(net, shape)
    mlp = sur.HbMlp(net.params.data_ptr(), net.adam_m.data_ptr(), net.adam_v.data_ptr(), norm.data_ptr(), sur.HIDDEN,
                    sur.N_OUT)
    t_net = timed(lambda: lib.hb_surrogate_maps(ctypes.byref(st), ctypes.byref(mlp), ctypes.byref(sp),
                                                big.maps.data_ptr(), outside.data_ptr(), sur.stream()))
    print(f"maps of {big.n} spins: exact composition {t_exact:.3f} ms, network {t_net:.3f} ms "
          f"({t_exact / t_net:.1f}x)")
    assert t_net < t_exact


def test_a_flip_change_reuses_the_network(gpu, slab):
    """The network learns the pulse's shape (RF over its peak) and takes B1 x peak: a different flip angle within the
    box's margin uses the same network, trained once, and stays within the judge's bound; so do the 16 refocusing
    flips of a TSE train (one network for all of them)."""
    ph, seq, coils, _ = slab
    seq2 = W.first_trs(seq, 2)
    first = driver.prepare_case(ph, seq2, coils, surrogate=True)
    net = first.extra["surrogate"][0][2]
    bumped = W.first_trs(W.gre3d_sequence(16, flip_deg=10.05), 2)
    again = driver.prepare_case(ph, bumped, coils, surrogate=True)
    assert again.extra["surrogate"][0][2] is net  # no retraining for a 0.5 % flip change
    driver.run_forward(again)
    out = driver.read_result(again)
    assert out.surrogate["pulses"][0]["judge_max_map_err"] < 1e-3
    exact = bp.simulate(ph, bumped, coils, engine="gpu", pulses="reuse")
    assert rel(out.signal, exact.signal) < 1e-3
    flips = torch.linspace(178.5, 180.0, 4, dtype=torch.float64)
    tse = bp.tse_2d(16, etl=4, tr=0.3, refocus_flip_deg=flips, slice_thickness=5e-3, rf_duration=2e-3,
                    system=bp.System(rf_raster=20e-6))
    tph = bp.shepp_logan(16)
    tcase = driver.prepare_case(tph, tse, bp.birdcage_sensitivities(tph.pos, 4), surrogate=True)
    nets = {id(n) for _, _, n in tcase.extra["surrogate"]}
    assert tcase.n_pulses == 5 and len(tcase.extra["surrogate"]) == 5
    assert len(nets) == 2  # the excitation, and one network for the four refocusing flips


def test_pulse_flip_against_the_oracle(gpu, slab):
    """The flip angle the training sampler follows: the network-free composition of the exact block quaternions in
    the lab frame against the oracle's block map without relaxation (R = I + D, flip = arccos R_zz), and the slab
    pulse's 10 degrees on resonance."""
    *_, shape = slab
    f = np.linspace(-1500.0, 1500.0, 61).astype(np.float32)
    flip = sur.pulse_flip(shape, torch.from_numpy(f).cuda(), shape.scale)
    fd = f.astype(np.float64)
    b = np.full(len(f), np.float32(shape.scale), dtype=np.float64)
    qs = S.block_quats(shape.table, shape.bounds, fd, b)
    D, _ = S.block_map(qs, shape.blocks[:, 0], fd, np.ones(len(f)), np.zeros(len(f)), np.zeros(len(f)),
                       shape.blocks[:, 1])
    ref = np.arccos(np.clip(1.0 + D[:, 2, 2], -1.0, 1.0))
    assert np.abs(flip - ref).max() < 1e-5
    assert abs(flip[30] - np.deg2rad(10.0)) < 1e-4  # f = 0


def test_surrogate_maps_compose_like_the_oracle(gpu):
    """The maps kernel's composition against the oracle's block_map with the split correction: a network whose
    outputs are zero stands for the first-order quaternions alone, on the multi-slice workload's edge-slice pulse
    (offsets up to +-6.7 kHz, T2 down to 80 ms, B1 0.7-1.3)."""
    import ctypes

    from blochparty_amd.engines import hip

    ph, seq, coils = W.multislice(16, lines=2)
    case = driver.prepare_case(ph, seq, coils, reuse=True)
    table = case.pulse_table.cpu().numpy()
    off, k = case.pulse_index.cpu().numpy()[4]
    shape = sur.shape_of(table[off : off + k])
    box = sur.padded_box(sur.input_range(case.spins, case.pos_lo, case.n, shape.g, shape.scale))
    net = sur.new_net(box, shape.fourier_base, torch.zeros(sur.N_OUT), torch.ones(sur.N_OUT), 0, "cuda",
                      sur.first_order(shape, box, "cuda"))
    net.params.zero_()
    norm = sur.pulse_norm(net, shape)
    maps = torch.zeros_like(case.maps)
    outside = torch.zeros(case.n_pad, dtype=torch.int32, device="cuda")
    blocks = torch.from_numpy(shape.blocks).cuda()
    sp = sur.HbSurPulse(blocks.data_ptr(), (ctypes.c_double * 3)(*shape.g), 4, 0)
    lib = hip.load()
    assert lib.hb_surrogate_maps(ctypes.byref(case.struct()), ctypes.byref(net.struct(norm)), ctypes.byref(sp),
                                 maps.data_ptr(), outside.data_ptr(), sur.stream()) == 0
    n = case.n
    mine = maps.view(case.n_pulses, 12, case.n_pad)[4, :, :n].double().cpu().numpy()
    assert int(outside[:n].sum()) == 0
    spins = case.spins[:, :n].double().cpu().numpy()
    pos = spins[:3] + case.pos_lo[:, :n].double().cpu().numpy()
    f = spins[6] + (shape.g[:, None] * pos).sum(0)
    b = spins[7] * shape.scale
    pd, r1, r2 = spins[3], spins[4], spins[5]
    qs = S.first_quats(S.first_order(shape.table, shape.bounds, f), b)
    mom = (S.first_moment(shape.table, shape.bounds, f) * b[:, None, None],
           S.second_moment(shape.table, shape.bounds, f) * (b * b)[:, None, None])
    D, v = S.block_map(qs, shape.blocks[:, 0], f, pd, r1, r2, shape.blocks[:, 1], mom)
    ref = np.concatenate([D.reshape(n, 9), v], 1).T
    err = np.abs(mine - ref).max()
    print(f"maps kernel vs oracle block map (first order, split correction): {err:.2e}")
    assert err < 3e-6
