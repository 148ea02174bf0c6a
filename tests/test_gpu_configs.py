"""The benchmark configurations at their own workloads, against the FP64 oracle (GPU).

- 3D GRE 64^3 x 4 sub-voxel isochromats, sinc slab, 8 coils: a seeded subset of the real phantom's spins on the real
  sequence over its first TRs, every RF path with per-sample and NUFFT readouts; and all 1,048,576 spins through
  the real dispatch plan (finite, repeatable bit for bit, stepping and reused maps agreeing).
- 2D TSE 128^2, ETL 16, 90/180 sincs of 100 substeps, a 16-flip refocusing train and the T1/T2/PD maps requiring
  grad, 8 coils, checkpoints every 64 segments, L2 loss to a seeded target: the adjoint's gradients against the
  oracle's float64 autograd and the oracle's central differences (h = 1e-3 deg along a seeded direction of the
  train), on a seeded subset of the phantom's spins over the whole sequence; and the full phantom's flip gradient
  against central differences of the GPU forward.
- 3D GRE 128^3, one isochromat per voxel, 16 coils: forward and the 16-coil adjoint against the oracle on a seeded
  subset over the first TRs, stepping and with the learned surrogate's maps.

The oracle runs on this host's CPU as the checker; the reference's own GPU tests follow the same pattern
(tests/engines/test_gpu_grad.py:38-131).
"""

import math
from dataclasses import replace

import pytest
import torch

import blochparty_amd as bp
from blochparty_amd import workloads as W
from blochparty_amd.engines import driver
from cases import rel
from oracle import core

pytestmark = pytest.mark.gpu
f64 = torch.float64


def signal_err(out, ref):
    return float((out.signal.to(torch.complex128) - ref).abs().max() / ref.abs().max())


# ---- 3D GRE 64^3 x 4 isochromats, sinc slab, 8 coils


@pytest.fixture(scope="module")
def gre64():
    ph, seq, coils = W.gre3d(64)
    return ph, seq, coils


@pytest.fixture(scope="module")
def gre64_sample(gre64):
    ph, seq, coils = gre64
    ph_s, c_s = W.subset(ph, coils, 4096, seed=11)
    seq_s = W.first_trs(seq, 8)
    ref, ref_m = core.simulate(ph_s, seq_s, c_s)
    return ph_s, seq_s, c_s, ref, ref_m


@pytest.mark.parametrize("pulses, readouts, tol", [
    ("step", "direct", 1e-5), ("reuse", "direct", 1e-5), ("surrogate", "direct", 1e-4),
    ("step", "nufft", 1e-4), ("reuse", "nufft", 1e-4), ("surrogate", "nufft", 1e-4),
])
def test_gre64_sample_every_path(gpu, gre64_sample, pulses, readouts, tol):
    ph, seq, coils, ref, ref_m = gre64_sample
    if readouts == "nufft":  # the readouts really take the NUFFT path on this sample
        case = driver.prepare_case(ph, seq, coils, nufft=(6, 3), reuse=pulses == "reuse")
        assert case.nu is not None
        del case
    out = bp.simulate(ph, seq, coils, engine="gpu", pulses=pulses, readouts=readouts)
    err = signal_err(out, ref)
    err_m = float((out.m.double() - ref_m).abs().max() / ref_m.abs().max())
    print(f"64^3 sample {pulses}/{readouts}: signal {err:.2e}, m {err_m:.2e}", flush=True)
    assert err < tol and err_m < tol, (err, err_m)
    if pulses == "surrogate":
        rep = out.surrogate
        assert rep["covered"] == 1 and not rep["rejected"] and rep["pulses"][0]["judge_max_map_err"] < 1e-4, rep


def test_gre64_full_size_dispatch_plan(gpu, gre64):
    """All 1,048,576 spins over the whole sequence: every output finite, the reused-map forward repeatable bit for
    bit, and stepping every substep within 1e-5 of it (relative to the largest sample)."""
    ph, seq, coils = gre64
    case = driver.prepare_case(ph, seq, coils, reuse=True)
    assert len(case.starts) > 1  # the real plan: several dispatches
    runs = []
    for _ in range(2):
        driver.run_forward(case)
        runs.append(driver.read_result(case))
    assert torch.equal(runs[0].signal, runs[1].signal) and torch.equal(runs[0].m, runs[1].m)
    assert torch.isfinite(runs[0].signal).all() and torch.isfinite(runs[0].m).all()
    del case
    torch.cuda.empty_cache()
    step = bp.simulate(ph, seq, coils, engine="gpu")
    e = rel(step.signal.to(torch.complex128), runs[0].signal.to(torch.complex128))
    e_m = rel(step.m.double(), runs[0].m.double())
    print(f"64^3 full size: step vs reuse signal {e:.2e}, m {e_m:.2e}", flush=True)
    assert e < 1e-5 and e_m < 1e-5, (e, e_m)


# ---- 2D TSE 128^2, ETL 16, adjoint


def tse128(flips):
    return bp.tse_2d(128, etl=16, tr=0.6, refocus_flip_deg=flips, slice_thickness=5e-3, rf_duration=2e-3,
                     system=bp.System(rf_raster=20e-6))


def tse_loss(sim, ph, flips, maps, coils, target):
    """L = sum |s - target|^2 with the T1 / T2 / PD maps (per spin) and the flip train as inputs."""
    t1, t2, pd = maps
    p = bp.Phantom(ph.pos, pd, t1, t2, ph.db0)
    s = sim(p, tse128(flips), coils)
    return ((s - target).abs() ** 2).sum()


def gpu_signal(p, seq, coils):
    return bp.simulate(p, seq, coils, engine="gpu", checkpoint_every=64).signal


def oracle_signal(p, seq, coils):
    return core.simulate(p, seq, coils)[0]


@pytest.fixture(scope="module")
def tse_case():
    ph_full = bp.shepp_logan(128)
    coils_full = bp.birdcage_sensitivities(ph_full.pos, 8)
    ph, coils = W.subset(ph_full, coils_full, 384, seed=3)
    g = torch.Generator().manual_seed(3)
    flips0 = torch.linspace(150.0, 180.0, 16, dtype=f64)
    target = torch.complex(torch.randn(8, 16384, generator=g, dtype=f64), torch.randn(8, 16384, generator=g,
                                                                                   dtype=f64)) * 0.05
    return ph_full, coils_full, ph, coils, flips0, target


def leaves(ph, flips0):
    return [flips0.clone().requires_grad_(), ph.t1.clone().requires_grad_(), ph.t2.clone().requires_grad_(),
            ph.pd.clone().requires_grad_()]


def test_tse128_adjoint_against_oracle_autograd_and_differences(gpu, tse_case):
    ph_full, coils_full, ph, coils, flips0, target = tse_case
    grads = {}
    for name, sim in (("gpu", gpu_signal), ("oracle", oracle_signal)):
        x = leaves(ph, flips0)
        loss = tse_loss(sim, ph, x[0], x[1:], coils, target)
        grads[name] = torch.autograd.grad(loss, x)
    errs = {k: rel(a.double(), b) for k, a, b in zip(("flips", "t1", "t2", "pd"), grads["gpu"], grads["oracle"])}
    print("TSE 128 adjoint vs oracle autograd:", {k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
    assert max(errs.values()) < 1e-4, errs
    # central differences of the oracle along a seeded direction of the flip train (h = 1e-3 deg)
    v = torch.randn(16, generator=torch.Generator().manual_seed(4), dtype=f64)
    v = v / v.norm()
    h = 1e-3
    with torch.no_grad():
        lp = tse_loss(oracle_signal, ph, flips0 + h * v, (ph.t1, ph.t2, ph.pd), coils, target)
        lm = tse_loss(oracle_signal, ph, flips0 - h * v, (ph.t1, ph.t2, ph.pd), coils, target)
    fd = float((lp - lm) / (2 * h))
    dir_gpu = float((grads["gpu"][0].double() * v).sum())
    dir_ref = float((grads["oracle"][0] * v).sum())
    print(f"TSE 128 directional dL/dflip: gpu {dir_gpu:.8e}, oracle autograd {dir_ref:.8e}, central differences "
          f"{fd:.8e}", flush=True)
    assert abs(dir_ref - fd) < 1e-6 * abs(fd)  # the reference gradient itself (truncation ~h^2)
    assert abs(dir_gpu - fd) < 1e-4 * abs(fd)


def test_tse128_full_phantom_flip_gradient(gpu, tse_case):
    """All 8,169 spins: the adjoint's gradient of the flip train along a seeded direction against central
    differences of the GPU forward (float32: h = 0.05 deg keeps the rounding of the loss below the truncation)."""
    ph_full, coils_full, _, _, flips0, target = tse_case
    x = leaves(ph_full, flips0)
    loss = tse_loss(gpu_signal, ph_full, x[0], x[1:], coils_full, target)
    (g,) = torch.autograd.grad(loss, x[:1])
    v = torch.randn(16, generator=torch.Generator().manual_seed(5), dtype=f64)
    v = v / v.norm()
    h = 0.05
    with torch.no_grad():
        lp = tse_loss(gpu_signal, ph_full, flips0 + h * v, (ph_full.t1, ph_full.t2, ph_full.pd), coils_full, target)
        lm = tse_loss(gpu_signal, ph_full, flips0 - h * v, (ph_full.t1, ph_full.t2, ph_full.pd), coils_full, target)
    fd = float((lp - lm) / (2 * h))
    d = float((g.double() * v).sum())
    print(f"TSE 128 full phantom directional dL/dflip: adjoint {d:.6e}, GPU central differences {fd:.6e}",
          flush=True)
    assert math.isfinite(d) and abs(d - fd) < 2e-3 * abs(fd), (d, fd)


# ---- 3D GRE 128^3, one isochromat per voxel, 16 coils


@pytest.fixture(scope="module")
def gre128_sample():
    ph, seq, coils = W.gre3d(128, n_coils=16, spins_per_voxel=(1, 1, 1))
    ph_s, c_s = W.subset(ph, coils, 2048, seed=12)
    return ph_s, W.first_trs(seq, 4), c_s


@pytest.mark.parametrize("pulses", ["step", "surrogate"])
def test_gre128_16_coils_forward_and_adjoint(gpu, gre128_sample, pulses):
    ph, seq, coils = gre128_sample
    ref, _ = core.simulate(ph, seq, coils)
    out = bp.simulate(ph, seq, coils, engine="gpu", pulses=pulses)
    err = signal_err(out, ref)
    print(f"128^3 sample, 16 coils, {pulses}: forward {err:.2e}", flush=True)
    assert err < (1e-5 if pulses == "step" else 1e-4), err
    # gradients w.r.t. the flip angle and the T1 / T2 maps, L = sum |s|^2
    grads = {}
    for name in ("gpu", "oracle"):
        flip = torch.tensor(10.0, dtype=f64, requires_grad=True)
        t1, t2 = ph.t1.clone().requires_grad_(), ph.t2.clone().requires_grad_()
        p = bp.Phantom(ph.pos, ph.pd, t1, t2, ph.db0)
        sq = replace(seq, rf=seq.rf * (flip / 10.0))  # the pulse's samples are linear in the flip angle
        if name == "gpu":
            s = bp.simulate(p, sq, coils, engine="gpu", pulses=pulses, rf_grad="scale", checkpoint_every=256).signal
        else:
            s = core.simulate(p, sq, coils)[0]
        grads[name] = torch.autograd.grad((s.abs() ** 2).sum(), [flip, t1, t2])
    errs = {k: rel(a.double(), b) for k, a, b in zip(("flip", "t1", "t2"), grads["gpu"], grads["oracle"])}
    print(f"128^3 sample, 16 coils, {pulses}: adjoint", {k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
    assert max(errs.values()) < (1e-4 if pulses == "step" else 1e-3), errs


# ---- multi-slice 2D GRE 256^2, 5 slices x 8 isochromats through 3 mm, the learned surrogate's workload


@pytest.fixture(scope="module")
def multislice_sample():
    """A seeded subset of the 2,621,440 spins on the real slices, 4 phase encodes per slice (20 TRs, five pulses:
    one per slice offset), with the oracle's signal."""
    ph = W.multislice_phantom(256)
    coils = bp.birdcage_sensitivities(ph.pos, 8)
    ph_s, c_s = W.subset(ph, coils, 8192, seed=13)
    seq = W.multislice_sequence(256, lines=4)
    ref, ref_m = core.simulate(ph_s, seq, c_s)
    return ph_s, seq, c_s, ref, ref_m


def test_multislice_surrogate_against_exact_and_oracle(gpu, multislice_sample):
    ph, seq, coils, ref, ref_m = multislice_sample
    sur = bp.simulate(ph, seq, coils, engine="gpu", pulses="surrogate")
    rep = sur.surrogate
    print("multislice judge:", [(r["pulse"], f"{r['judge_max_map_err']:.1e}", f"{r['train_s']:.1f}s")
                                for r in rep["pulses"]], flush=True)
    exact = bp.simulate(ph, seq, coils, engine="gpu", pulses="reuse")
    step = bp.simulate(ph, seq, coils, engine="gpu")
    errs = {name: signal_err(o, ref) for name, o in (("surrogate", sur), ("reuse", exact), ("step", step))}
    errs["surrogate_vs_exact_maps"] = rel(sur.signal.to(torch.complex128), exact.signal.to(torch.complex128))
    print("multislice signal vs oracle:", {k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
    assert rep["covered"] == 5 and not rep["rejected"]
    assert errs["step"] < 1e-5 and errs["reuse"] < 1e-5
    assert errs["surrogate"] < 1e-4 and errs["surrogate_vs_exact_maps"] < 1e-4
    # the largest map-entry difference on the judged spins, well inside the judge's tolerance
    assert max(r["judge_max_map_err"] for r in rep["pulses"]) < 0.5 * rep["tolerance"]


def test_multislice_fallback_for_spins_outside_the_box(gpu, multislice_sample):
    """Spins moved far off resonance lie outside the networks' boxes: the exact maps stand in for theirs."""
    ph, seq, coils, ref, _ = multislice_sample
    case = driver.prepare_case(ph, seq, coils, surrogate=True)
    nets = case.extra["surrogate"]
    far = bp.Phantom(ph.pos, ph.pd, ph.t1, ph.t2, ph.db0.clone(), b1=ph.b1)
    far.db0[::5] += 20000.0  # beyond the +-6.7 kHz of offsets the slice-select gradient spans over the slices
    case_far = driver.prepare_case(far, seq, coils, reuse=True)
    case_far.extra["surrogate"] = nets  # the networks of the unshifted spins, not trained anew
    driver.run_forward(case_far)
    out = driver.read_result(case_far)
    outside = [r["outside_box"] for r in out.surrogate["pulses"]]
    assert all(o == (len(far) + 4) // 5 for o in outside), outside
    exact = bp.simulate(far, seq, coils, engine="gpu", pulses="reuse")
    e = rel(out.signal.to(torch.complex128), exact.signal.to(torch.complex128))
    print(f"multislice fallback: {outside[0]} spins outside per pulse, vs exact maps {e:.1e}", flush=True)
    assert e < 1e-4
