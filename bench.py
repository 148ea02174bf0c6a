"""Benchmark of the GPU path on one MI355X: spin-samples/s of simulate(..., engine="gpu") on the 3D GRE 64^3 workload.

    python bench.py [--steps K] [--warmup W] [--n 64] [--no-cpu] [--no-adjoint] [--no-exact] [--no-e2e]

Workload: 3D GRE n^3 x 4 sub-voxel isochromats (1,048,576 spins at n = 64), a three-class smooth random phantom,
+-100 Hz dB0, 8 coils, a 2 ms sinc slab pulse (TBW 4, 200 substeps) every TR of 8 ms, n^2 lines of n samples.
Every input of a device leg is resident in HBM before its timed region starts (host packing and upload are outside
it); each timed region is bracketed by torch.cuda.synchronize(). Throughput = spins x ADC samples x steps / seconds
(the reference's metric, benchmarks/bench.py:65).

The JSON line is printed as soon as the value's leg, its roofline and the CPU baseline exist, and printed again,
extended, after every later leg: the last line printed is the complete one.

Legs (the value's leg runs K timed steps after W warmup steps; every other leg runs min(K, its cap) timed steps after
min(W, 1) warmup steps, the caps being --leg-steps and --adjoint-steps, and each leg states its own counts):
- forward with the learned RF surrogate (pulses="surrogate") and NUFFT readouts: every spin's pulse map from the
  network, the exact maps of every 257th spin composed and compared (the judge), then the event loop with every
  readout as a type-1 NUFFT -- the JSON line's ``value``. The network is trained once before the timed steps (its
  training time is reported, not timed);
- forward with reused exact maps (pulses="reuse": the sinc pulse composed once per spin in float64), NUFFT readouts;
- forward stepping every RF substep, NUFFT readouts (exact RF with the same readouts: the RF path's own ratio);
- forward stepping every RF substep, per-sample readouts (the exact-RF baseline);
- forward + adjoint of L = sum |signal|^2 w.r.t. the flip angle (through the RF gradients), the T1 map and the coil
  maps (the reference's --grad case, benchmarks/bench.py:69-95): the forward with checkpoints, dL/dsignal on the
  device, and the reverse sweep; with exact RF, then with the learned surrogate's maps (per-sample readouts), then
  with the surrogate's maps and NUFFT readouts (each readout run reversed as the transpose of its NUFFT; its ratio
  to the exact-RF leg is ``fwd_adjoint_speedup_vs_exact_rf``). Each reports dL/dflip, which should agree;
- the multi-slice workload of the learned surrogate (5 slices x 256^2 x 8 isochromats, 3 ms sinc at 30 deg, B1
  0.7-1.3): forward with the surrogate's maps and forward stepping every RF substep, same readouts (``multislice``);
- ``full_variant``: the full variant (a super-Lorentzian bound pool and a 0.55 T scanner's concomitant fields,
  workloads.mt_gre2d: 65,536 spins, 16,384 samples, 8 coils) forward, and forward + adjoint of L = sum |signal|^2
  w.r.t. the flip angle, the T1 map, the bound-pool fraction and the coil maps (float64 reverse sweep);
- ``downstream``: either side of the path -- the sequence builder and the segment table on the host (wall time), and
  the reconstruction of the workload's k-space on the GPU (grid, centred 3D inverse FFT, root-sum-of-squares)
  against numpy in float64, with its HBM roofline;
- ``end_to_end``: simulate() timed from the caller's Phantom / Sequence / coils to the CPU result, the way the
  reference's harness times it (benchmarks/bench.py:47-95; sequence and phantom built outside the timed region):
  the value's configuration forward, and forward + .backward() with the flip angle, the T1 map and the coils
  requiring grad.
Every forward-kernel launch of the value's leg is bracketed by HIP events on the stream it runs on; the roofline's
``achieved`` is the algorithmic flops per launch over their mean duration. ``traffic`` is the HBM bytes per launch
of the same kernel from the committed rocprofv3 PMC summary (profiles/), when present.

The CPU baseline is the FP64 oracle timed on this host's cores (the container quota) on a bounded sample of the same
workload (a seeded subset of spins over the first TRs); the GPU runs the same sample on every RF path and reports
their relative errors against the oracle. Progress goes to stderr.
"""

from __future__ import annotations

import argparse
import dataclasses
import json
import os
import sys
import time

import numpy as np
import torch

METRIC = "spin·samples/s fwd & fwd+adjoint, 3D GRE 64³; RF surrogate vs exact @1e-4"
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector, published
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published
FLIP_DEG = 10.0  # the workload's flip angle: rf = FLIP_DEG x (the pulse's shape per degree)

# counting convention of the algorithmic work (per spin)
FLOPS_RF = 48  # exact RF substep
FLOPS_FREE = 12  # free or ADC event
FLOPS_COIL = 8  # per coil, per spin-sample (complex multiply-add)
FLOPS_PULSE = 30  # a reused pulse's map applied at one occurrence
FLOPS_MZ = 3  # a free segment spoiled at its end: Mz relaxation only
FLOPS_SPREAD = 4  # a NUFFT grid column per spin: kernel weight x decay term x complex weight


def log(msg: str) -> None:
    print(f"# {msg}", file=sys.stderr, flush=True)


def cpu_quota() -> int:
    """Cores this process may use: the cgroup CPU quota and the affinity mask, whichever is smaller."""
    n = len(os.sched_getaffinity(0))
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()
        if quota != "max":
            n = min(n, max(1, int(float(quota) / float(period))))
    except (OSError, ValueError):
        pass
    return max(1, n)


def dispatch_flops(case, driver) -> list[float]:
    """Algorithmic flops of every forward dispatch of ``case``: stepped segments by the counting convention, reused
    pulses at FLOPS_PULSE per occurrence."""
    rec = case.segs.cpu().numpy().view(driver.SEG_DTYPE)
    flags = rec["flags"]
    rf = (flags & driver.RF) != 0
    adc = (flags & driver.ADC) != 0
    per_seg = np.where(rf, FLOPS_RF, FLOPS_FREE) + adc * FLOPS_COIL * case.n_coils
    per_seg = np.where((flags & driver.SPOILED_FREE) != 0, FLOPS_MZ, per_seg)  # Mz relaxation only
    for s in np.flatnonzero(flags & driver.MERGED):  # one free event for the merged run
        per_seg[s + 1 : s + rec["pad"][s, 3]] = 0
    for _, k, occ in case.pulses:
        for a, _ in occ:
            per_seg[a] = FLOPS_PULSE
            per_seg[a + 1 : a + k] = 0
    if case.nu is not None:  # a NUFFT run: the jump, and width x terms grid columns (kernel value, decay, coils)
        pl = case.nu["plan"]
        for a, k in pl.runs:
            per_seg[a] = FLOPS_FREE + pl.width * pl.terms * (FLOPS_COIL * case.n_coils + FLOPS_SPREAD)
            per_seg[a + 1 : a + k] = 0
    cum = np.concatenate([[0], np.cumsum(per_seg, dtype=np.float64)])
    return [float(cum[e] - cum[s]) * case.n for s, e in zip(case.starts, case.ends)]


def timed(fn, steps: int, warmup: int) -> float:
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


PMC_SUMMARY = "profiles/forward_kernel_gre3d64_nufft_pmc.txt"


def profiled_traffic() -> float | None:
    """HBM bytes per forward-kernel launch from the committed PMC summary (FETCH_SIZE + WRITE_SIZE, scaled by a copy
    of known size), or None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), PMC_SUMMARY)
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch forward_kernel:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def forward_leg(case, lib, driver, steps: int, warmup: int, work: int, label: str) -> dict:
    """Time ``steps`` forward passes of ``case``; every forward-kernel launch bracketed by events."""
    events = []

    def timer(_phase):
        ev = torch.cuda.Event(enable_timing=True)
        ev.record()
        events.append(ev)

    for _ in range(warmup):
        driver.run_forward(case, lib)
    t = timed(lambda: driver.run_forward(case, lib, timer=timer), steps, 0)
    kernel_ms = [events[2 * i].elapsed_time(events[2 * i + 1]) for i in range(len(events) // 2)]
    flops = dispatch_flops(case, driver)
    mean_ms, mean_flops = float(np.mean(kernel_ms)), float(np.mean(flops))
    out = {"value": work * steps / t, "ms_per_step": 1e3 * t / steps, "steps": steps, "warmup": warmup,
           "kernel_mean_ms": mean_ms, "launches": len(kernel_ms), "flops_per_launch": mean_flops,
           "kernel_tflops": mean_flops / (mean_ms * 1e-3) / 1e12}
    log(f"forward ({label}): {out['ms_per_step']:.1f} ms/step over {steps} steps, {out['value']:.3e} spin-samples/s; "
        f"forward kernel mean {mean_ms:.3f} ms over {len(kernel_ms)} launches, {out['kernel_tflops']:.1f} TFLOP/s "
        f"algorithmic")
    return out


def fwd_adjoint_leg(case_a, seq, coils, every, lib, adjoint, driver, steps: int, warmup: int, work: int,
                    label: str) -> dict:
    """Forward with checkpoints + reverse sweep of L = sum |s|^2 w.r.t. the flip angle, the T1 map and the coils.
    rf = flip x shape: dL/dflip = sum_s rf_s . dL/drf_s / flip for stepped RF; a reused pulse's occurrence carries
    dL/d(ln amplitude) in its first segment's rf_x slot."""
    S, n_adc = len(seq), seq.n_adc
    nc = adjoint.adjoint_width(case_a.n_coils)
    dev = case_a.spins.device
    if case_a.perm is not None:  # NUFFT readouts: the spins are sorted on the device
        coils = coils[torch.from_numpy(case_a.perm)]
    sens = adjoint.adjoint_sens(coils, case_a.n_pad, nc).to(dev)
    rf64 = seq.rf.double().to(dev)
    heads = torch.tensor([a for _, _, occ in case_a.pulses for a, _ in occ], dtype=torch.int64, device=dev)
    out = {}

    def step():
        ckpts = driver.run_forward(case_a, lib, keep=True)
        sig = case_a.sig[:n_adc]  # L = sum |s|^2: dL/dsignal = 2 s, turned by exp(+i psi)
        gr, gi = 2 * sig[..., 0], 2 * sig[..., 1]
        cp, sp = torch.cos(case_a.psi_rows)[:, None], torch.sin(case_a.psi_rows)[:, None]
        h = torch.zeros(n_adc, nc, 2, dtype=torch.float32, device=dev)
        h[:, : case_a.n_coils, 0] = (gr * cp - gi * sp).float()
        h[:, : case_a.n_coils, 1] = (gr * sp + gi * cp).float()
        lam = torch.zeros(3, case_a.n_pad, dtype=torch.float32, device=dev)
        gs, gc, segg, _ = adjoint.run_backward(case_a, ckpts, every, h, lam, sens, nc, 2, True, lib)
        sg = segg.view(S, 8)
        if len(heads):
            out["flip"] = sg[heads, 1].sum() / FLIP_DEG
        else:
            out["flip"] = (sg[:, 1:3] * rf64).sum() / FLIP_DEG
        out["t1"] = gs[5]
        out["coils"] = gc

    t_a = timed(step, steps, warmup)
    assert torch.isfinite(out["flip"]) and torch.isfinite(out["t1"]).all() and torch.isfinite(out["coils"]).all()
    leg = {"value": work * steps / t_a, "ms_per_step": 1e3 * t_a / steps, "steps": steps, "warmup": warmup,
           "rf": label, "checkpoint_every": every, "loss": "sum |signal|^2",
           "wrt": ["flip angle", "T1 map", "coil maps"], "dL_dflip_per_deg": float(out["flip"])}
    log(f"forward + adjoint ({label}): {leg['ms_per_step']:.1f} ms/step over {steps} steps, "
        f"{leg['value']:.3e} spin-samples/s, dL/dflip {leg['dL_dflip_per_deg']:.6e}")
    return leg


def value_leg(line, ph, seq, coils, nufft, lib, driver, sur, steps: int, warmup: int, work: int) -> dict:
    """The headline: the forward with the learned surrogate's maps (the network trained first, outside the timed
    region), with the forward kernel's roofline."""
    t0 = time.perf_counter()
    case_s = driver.prepare_case(ph, seq, coils, surrogate=True, nufft=nufft)
    torch.cuda.synchronize()
    nets = case_s.extra["surrogate"]
    log(f"surrogate: case prepared and {len(nets)} network(s) trained in {time.perf_counter() - t0:.1f} s")
    lead = forward_leg(case_s, lib, driver, steps, warmup, work, "learned RF surrogate")
    line["value"], line["ms_per_step"] = lead["value"], lead["ms_per_step"]
    line["roofline"] = {
        "bound": "valu", "kernel": "forward_kernel", "achieved": lead["kernel_tflops"], "peak": FP32_PEAK_TFLOPS,
        "unit": "TFLOP/s", "frac": lead["kernel_tflops"] / FP32_PEAK_TFLOPS, "traffic": profiled_traffic(),
        "mean_launch_ms": lead["kernel_mean_ms"], "launches": lead["launches"],
        "flops_per_launch": lead["flops_per_launch"],
    }
    line["surrogate"] = {"forward": lead, "judge": sur.report(case_s)}
    del case_s
    torch.cuda.empty_cache()
    return lead


def copy_probe(mib: int) -> None:
    """A plain copy of known size for scaling byte counters: x * 1 into y reads and writes ``mib`` MiB, three times."""
    x = torch.ones(mib << 18, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        torch.mul(x, 1.0, out=y)
    torch.cuda.synchronize()
    log(f"copy probe: 3 x {mib} MiB read + {mib} MiB written")


def adjoint_legs(args, line, emit, ph, seq, coils, nufft, lib, adjoint, driver, adj_k, adj_w, work) -> None:
    """The forward + adjoint legs: the surrogate's maps (or, with --adjoint-pulses reuse, the exact reused maps) with
    NUFFT readouts, then per-sample readouts, then exact RF."""
    every = args.checkpoint_every
    maps = dict(surrogate=True) if args.adjoint_pulses == "surrogate" else dict(reuse=True)
    name = "learned RF surrogate" if args.adjoint_pulses == "surrogate" else "reused exact maps"
    tag = "surrogate" if args.adjoint_pulses == "surrogate" else "reuse"
    legs = [(f"fwd_adjoint_{tag}_nufft", dict(maps, nufft=nufft), f"{name}, NUFFT readouts")] if nufft else []
    legs += [(f"fwd_adjoint_{tag}", dict(maps), name), ("fwd_adjoint", {}, "exact RF")]
    for key, kw, label in legs:
        if key == "fwd_adjoint" and args.no_exact:
            continue
        case_a = driver.prepare_case(ph, seq, coils, chunk=every, aligned=True, **kw)
        line[key] = fwd_adjoint_leg(case_a, seq, coils, every, lib, adjoint, driver, adj_k, adj_w, work, label)
        del case_a
        torch.cuda.empty_cache()
        best = line.get(f"fwd_adjoint_{tag}_nufft") or line.get(f"fwd_adjoint_{tag}")
        if line.get("fwd_adjoint") and best:
            line["fwd_adjoint_speedup_vs_exact_rf"] = best["value"] / line["fwd_adjoint"]["value"]
        emit()
    if args.copy_probe and args.adjoint_only:  # the counters' scale, as in the full run
        copy_probe(args.copy_probe)


def multislice_leg(W, driver, sur, lib, steps: int, warmup: int, nufft) -> dict:
    """The learned surrogate's own workload (multi-slice 2D GRE, 5 slices x 256^2 x 8 isochromats, 3 ms sinc at
    30 deg, B1 0.7-1.3, +-200 Hz): forward with the surrogate's maps against stepping every RF substep, both with
    the same readouts; the surrogate's signal error against the exact path's, and the judge's map errors."""
    t0 = time.perf_counter()
    ph, seq, coils = W.multislice()
    work = len(ph) * seq.n_adc
    log(f"multislice workload built in {time.perf_counter() - t0:.1f} s: N={len(ph)} S={len(seq)} "
        f"n_adc={seq.n_adc}")
    t0 = time.perf_counter()
    case = driver.prepare_case(ph, seq, coils, surrogate=True, nufft=nufft)
    torch.cuda.synchronize()
    log(f"multislice surrogate: {len(case.extra['surrogate'])} network(s) ready in {time.perf_counter() - t0:.1f} s")
    out = {"workload": "multi-slice 2D GRE, 5 slices x 256^2 x 8 isochromats through 3 mm, 3 ms sinc (TBW 4, "
                       "300 substeps) at 30 deg per slice, B1 U(0.7, 1.3), +-200 Hz, 8 coils, TR 10 ms",
           "spins": len(ph), "n_adc": seq.n_adc, "readouts": "nufft" if nufft else "direct"}
    out["surrogate"] = forward_leg(case, lib, driver, steps, warmup, work, "multislice, learned RF surrogate")
    out["judge"] = sur.report(case)
    sig_s = driver.read_result(case).signal.to(torch.complex128)
    del case
    torch.cuda.empty_cache()
    case = driver.prepare_case(ph, seq, coils, nufft=nufft)
    out["exact_rf"] = forward_leg(case, lib, driver, steps, warmup, work, "multislice, exact RF")
    sig_e = driver.read_result(case).signal.to(torch.complex128)
    del case
    torch.cuda.empty_cache()
    out["speedup_vs_exact_rf"] = out["surrogate"]["value"] / out["exact_rf"]["value"]
    out["surrogate_signal_rel_err_vs_exact"] = float((sig_s - sig_e).abs().max() / sig_e.abs().max())
    log(f"multislice: surrogate {out['speedup_vs_exact_rf']:.1f}x exact RF, signal error vs exact "
        f"{out['surrogate_signal_rel_err_vs_exact']:.2e}")
    return out


def full_variant_leg(W, driver, lib, steps: int, warmup: int, adj_steps: int, adj_warmup: int) -> dict:
    """The full variant on workloads.mt_gre2d: hb_forward_mt forward (every launch bracketed by events on its
    stream), then forward with checkpoints + hb_adjoint_full of L = sum |signal|^2 (dL/dflip, dL/dT1 map, dL/df map,
    coil gradients)."""
    from blochparty_amd import fields as F
    from blochparty_amd.engines import adjoint
    from blochparty_amd.engines import full_adjoint as FA

    ph, seq, coils, scanner = W.mt_gre2d()
    fl = F.spin_fields(scanner, ph.pos)
    conc = (fl.conc, fl.mode)
    S, n_adc = len(seq), seq.n_adc
    work = len(ph) * n_adc
    out = {"workload": "2D slice-selective spoiled GRE 128^2 x 4 isochromats through 3 mm, super-Lorentzian bound "
                       "pool by tissue class, 0.55 T leading-order concomitant fields, 8 coils, 2 ms sinc (100 "
                       "substeps) of 20 deg, TR 10 ms", "spins": len(ph), "segments": S, "n_adc": n_adc}
    case = driver.prepare_case(ph, seq, coils, conc=conc)
    events = []

    def timer(_phase):
        ev = torch.cuda.Event(enable_timing=True)
        ev.record()
        events.append(ev)

    for _ in range(warmup):
        driver.run_forward(case, lib)
    t = timed(lambda: driver.run_forward(case, lib, timer=timer), steps, 0)
    kernel_ms = [events[2 * i].elapsed_time(events[2 * i + 1]) for i in range(len(events) // 2)]
    out["forward"] = {"value": work * steps / t, "ms_per_step": 1e3 * t / steps, "steps": steps, "warmup": warmup,
                      "kernel": "forward_mt_kernel", "kernel_mean_ms": float(np.mean(kernel_ms)),
                      "launches_per_step": len(kernel_ms) // steps}
    log(f"full variant forward: {out['forward']['ms_per_step']:.1f} ms/step, {out['forward']['value']:.3e} "
        f"spin-samples/s")
    del case
    every = adjoint.default_checkpoint_every(S)
    case_a = driver.prepare_case(ph, seq, coils, chunk=every, aligned=True, conc=conc)
    dev = case_a.spins.device
    nc = adjoint.adjoint_width(case_a.n_coils)
    sens = adjoint.adjoint_sens(coils, case_a.n_pad, nc).to(dev)
    flip_shape = seq.rf.double().to(dev) / 20.0
    res = {}

    def step():
        ckpts = driver.run_forward(case_a, lib, keep=True)
        sig = case_a.sig[:n_adc]
        gr, gi = 2 * sig[..., 0], 2 * sig[..., 1]
        cp, sp = torch.cos(case_a.psi_rows)[:, None], torch.sin(case_a.psi_rows)[:, None]
        h = torch.zeros(n_adc, nc, 2, dtype=torch.float32, device=dev)
        h[:, : case_a.n_coils, 0] = (gr * cp - gi * sp).float()
        h[:, : case_a.n_coils, 1] = (gr * sp + gi * cp).float()
        lam = torch.zeros(4, case_a.n_pad, dtype=torch.float32, device=dev)
        gs, gc, segg, _ = FA.run_backward(case_a, ckpts, every, h, lam, sens, nc, True, True)
        res["flip"] = (segg.view(S, 8)[:, 1:3] * flip_shape).sum()
        res["t1"], res["f"], res["coils"] = gs[5], gs[15], gc

    t_a = timed(step, adj_steps, adj_warmup)
    assert all(torch.isfinite(v).all() for v in res.values())
    out["forward_adjoint"] = {"value": work * adj_steps / t_a, "ms_per_step": 1e3 * t_a / adj_steps,
                              "steps": adj_steps, "warmup": adj_warmup, "checkpoint_every": every,
                              "loss": "sum |signal|^2", "wrt": ["flip angle", "T1 map", "bound-pool f", "coil maps"],
                              "dL_dflip_per_deg": float(res["flip"])}
    log(f"full variant forward + adjoint: {out['forward_adjoint']['ms_per_step']:.1f} ms/step, "
        f"{out['forward_adjoint']['value']:.3e} spin-samples/s, dL/dflip {float(res['flip']):.6e}")
    return out


def gpu_ms(fn, steps: int):
    """Mean time of fn() on the current stream (HIP events), after two untimed calls; and its last result."""
    for _ in range(2):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        out = fn()
    e1.record()
    e1.synchronize()
    return e0.elapsed_time(e1) / steps, out


def recon_timing(run, n_coils: int, shape: tuple, steps: int) -> dict:
    """GPU time of ``run`` (k-space (C, n_adc) complex64 on the device -> RSS image; the HIP kernels) on seeded
    synthetic k-space, its error against numpy's centred inverse FFT and RSS in float64, its HBM roofline, and the
    same composition with torch's FFT (rocFFT) timed beside it."""
    g = torch.Generator().manual_seed(5)
    k = torch.randn(n_coils, int(np.prod(shape)), dtype=torch.complex64, generator=g)
    kd = k.cuda()
    ms, img = gpu_ms(lambda: run(kd), steps)
    axes3 = tuple(range(-len(shape), 0))
    torch_ms, _ = gpu_ms(lambda: torch.fft.fftshift(torch.fft.ifftn(torch.fft.ifftshift(
        kd.reshape(n_coils, *shape), dim=axes3), dim=axes3), dim=axes3).abs().pow(2).sum(0).sqrt(), steps)
    t0 = time.perf_counter()
    axes = (1, 2, 3)
    kk = k.numpy().astype(np.complex128).reshape(n_coils, *shape)
    ref = np.sqrt((np.abs(np.fft.fftshift(np.fft.ifftn(np.fft.ifftshift(kk, axes=axes), axes=axes), axes=axes))
                   ** 2).sum(axis=0))
    t_cpu = time.perf_counter() - t0
    err = float(np.abs(img.cpu().double().numpy() - ref).max() / np.abs(ref).max())
    floor = k.numel() * 8 + ref.size * 4
    gbs = floor / (ms * 1e-3) / 1e9
    log(f"recon (grid, ifft3c, rss) of {n_coils} x {shape}: {ms:.3f} ms on the GPU ({gbs:.0f} GB/s of the floor; "
        f"torch's FFT {torch_ms:.3f} ms), {t_cpu * 1e3:.0f} ms numpy f64, rel err {err:.2e}")
    return {"shape": [n_coils, *shape], "ms": ms, "steps": steps, "torch_fft_ms": torch_ms,
            "cpu_numpy_f64_ms": t_cpu * 1e3, "rel_err_vs_numpy_f64": err,
            "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": gbs / HBM_PEAK_GBS, "algorithmic_bytes": floor,
                         # what the HIP passes move: a read and a write of every value per axis, then the RSS
                         "traffic": len(shape) * 16 * k.numel() + 8 * k.numel() + 4 * ref.size}}


def downstream_leg(W, driver, n: int, n_coils: int, steps: int) -> dict:
    """Either side of the path: building the n^3 GRE with the sequence builder and its segment table on the host
    (their wall time), and reconstructing its k-space on the GPU (Cartesian grid, centred 3D inverse FFT,
    root-sum-of-squares over the coils; torch's FFT, i.e. rocFFT) against numpy's in float64 on the same seeded
    synthetic k-space. The reconstruction's HBM floor: read the complex64 k-space once, write the float32 image."""
    from blochparty_amd import recon

    out = {}
    t0 = time.perf_counter()
    seq = W.gre3d_sequence(n)
    t_build = time.perf_counter() - t0
    t0 = time.perf_counter()
    rec = driver.segment_table(seq, [0])
    t_table = time.perf_counter() - t0
    out["sequence_build"] = {"s": t_build, "segments": len(seq), "segments_per_s": len(seq) / t_build}
    out["segment_table"] = {"s": t_table, "records": len(rec), "records_per_s": len(rec) / t_table}
    log(f"sequence builder {t_build:.2f} s for {len(seq)} segments; segment table {t_table:.2f} s")

    shape = tuple(seq.meta["kspace_shape"])
    out["recon_rss"] = recon_timing(lambda kd: recon.rss(recon.reconstruct(kd, seq)), n_coils, shape, steps)
    # the largest grid of the workloads (the 128^3 GRE with 16 coils), where launches no longer dominate
    big = (128, 128, 128)
    out["recon_rss_128"] = recon_timing(lambda kd: recon.rss(recon.ifft3c(kd.reshape(16, *big))), 16, big, steps)
    return out


def end_to_end(bp, W, args, ph, seq, coils, work: int, steps: int, warmup: int) -> dict:
    """simulate() timed from the caller's objects to the CPU result (host packing, upload, the GPU work, readback),
    as the reference's harness times it; the grad case adds .backward() of sum |signal|^2 with the flip angle, the
    T1 map and the coils requiring grad. Inputs are built before the timed region, as the reference's are."""
    opts = dict(engine="gpu", pulses="surrogate", readouts="nufft")
    out = {"options": "pulses='surrogate', readouts='nufft' (rf_grad='scale' for the gradient case: the flip angle "
                      "is each occurrence's amplitude)", "steps": steps, "warmup": warmup}

    def fwd():
        sig = bp.simulate(ph, seq, coils, **opts).signal
        assert torch.isfinite(sig).all()

    t = timed(fwd, steps, warmup)
    out["forward"] = {"value": work * steps / t, "s_per_call": t / steps}
    log(f"end to end: simulate() forward {t / steps:.2f} s/call, {work * steps / t:.3e} spin-samples/s")
    # the grad case: the flip as a tensor through the sequence builder, the T1 map and the coils as leaves
    flip = torch.tensor(FLIP_DEG, dtype=torch.float64, requires_grad=True)
    shape = seq.rf / FLIP_DEG  # the pulse's samples per degree (the builder's rf is linear in the flip angle)
    t1 = ph.t1.detach().clone().requires_grad_()
    ph_g = bp.Phantom(ph.pos, ph.pd, t1, ph.t2, ph.db0, b1=ph.b1)
    coils_g = coils.detach().clone().requires_grad_()

    def fwd_bwd():
        for x in (flip, t1, coils_g):
            x.grad = None
        seq_g = dataclasses.replace(seq, rf=shape * flip)  # a fresh graph from the flip angle per call
        sig = bp.simulate(ph_g, seq_g, coils_g, rf_grad="scale", **opts).signal
        (sig.abs() ** 2).sum().backward()
        assert torch.isfinite(flip.grad) and torch.isfinite(t1.grad).all() and torch.isfinite(coils_g.grad).all()

    t = timed(fwd_bwd, steps, warmup)
    out["forward_backward"] = {"value": work * steps / t, "s_per_call": t / steps,
                               "dL_dflip_per_deg": float(flip.grad)}
    log(f"end to end: simulate() + backward {t / steps:.2f} s/call, {work * steps / t:.3e} spin-samples/s, "
        f"dL/dflip {float(flip.grad):.6e}")
    return out


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--leg-steps", type=int, default=3,
                    help="timed steps of each forward leg other than the value's, at most (and at most --steps)")
    ap.add_argument("--adjoint-steps", type=int, default=1,
                    help="timed steps of each forward + adjoint leg and each end-to-end case, at most")
    ap.add_argument("--n", type=int, default=64, help="grid size of the 3D GRE")
    ap.add_argument("--coils", type=int, default=8, help="receive coils")
    ap.add_argument("--iso", type=int, choices=(1, 4), default=4,
                    help="isochromats per voxel (4: 1 x 2 x 2 sub-voxel, the benchmark's; 1: one per voxel)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and the parity sample")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the forward + adjoint legs")
    ap.add_argument("--no-exact", action="store_true", help="skip the exact-RF legs")
    ap.add_argument("--no-e2e", action="store_true", help="skip the end-to-end simulate() timing")
    ap.add_argument("--no-multislice", action="store_true",
                    help="skip the multi-slice surrogate workload (surrogate vs exact RF forward)")
    ap.add_argument("--only-value", action="store_true",
                    help="the value's leg alone (no other leg, no CPU baseline): for profiler runs")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="after the legs, three elementwise copies of MIB MiB (read + write) that calibrate the HBM "
                         "byte counters of a profiler run")
    ap.add_argument("--readouts", choices=("direct", "nufft"), default="nufft",
                    help="readouts of the surrogate and reused-maps legs (nufft: type-1 NUFFT per readout run)")
    ap.add_argument("--checkpoint-every", type=int, default=4352,
                    help="segments per checkpoint interval of the forward + adjoint legs (a multiple of 64; 4 x the "
                         "library's sqrt-of-segments default: 256 checkpoints of 24 B per spin and a 6.5 GB replay "
                         "scratch fit one GPU's HBM many times, and fewer, longer launches shorten the tails)")
    ap.add_argument("--no-surrogate", action="store_true",
                    help="no network trained (counter runs): skip the surrogate's value leg, and the forward + adjoint "
                         "legs take the exact reused maps")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only the forward + adjoint legs (profiling runs: no value, CPU, multi-slice or end-to-end legs)")
    ap.add_argument("--adjoint-pulses", choices=("surrogate", "reuse"), default="surrogate",
                    help="the pulses of the non-exact forward + adjoint legs: the learned surrogate's maps, or the exact "
                         "reused maps (no network trained: for counter runs)")
    ap.add_argument("--no-full-variant", action="store_true", help="skip the full variant's legs")
    ap.add_argument("--no-downstream", action="store_true",
                    help="skip the sequence build / segment table / reconstruction timings")
    ap.add_argument("--full-variant-only", action="store_true",
                    help="only the full variant's legs (bound pool, concomitant fields), no JSON line")
    ap.add_argument("--cpu-trs", type=int, default=32)
    ap.add_argument("--cpu-spins", type=int, default=65536)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: the path shards over spins, but this build measures one GPU's share "
              f"(--gpus {args.gpus} is not supported)", file=sys.stderr)
        sys.exit(2)
    if args.only_value:
        args.no_cpu = args.no_adjoint = args.no_exact = args.no_e2e = args.no_multislice = True
        args.no_full_variant = args.no_downstream = True
    if args.adjoint_only:
        args.no_cpu = args.no_e2e = args.no_multislice = args.no_full_variant = args.no_downstream = True
    if args.no_surrogate:
        args.no_e2e = args.no_multislice = True
        args.adjoint_pulses = "reuse"
    if args.checkpoint_every % 64:
        raise SystemExit("--checkpoint-every must be a multiple of 64")

    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import blochparty_amd as bp
    from blochparty_amd import workloads as W
    from blochparty_amd.engines import adjoint, driver, hip
    from blochparty_amd.engines import surrogate as sur

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    hip.load()
    lib = hip.HIP.lib()
    if args.full_variant_only:
        fv = full_variant_leg(W, driver, lib, max(1, min(args.steps, args.leg_steps)), min(args.warmup, 1),
                              max(1, min(args.steps, args.adjoint_steps)), min(args.warmup, 1))
        log(json.dumps(fv))
        return
    t_build = time.perf_counter()
    ph, seq, coils = W.gre3d(args.n, n_coils=args.coils, spins_per_voxel=(1, 2, 2) if args.iso == 4 else (1, 1, 1))
    S = len(seq)
    log(f"workload built in {time.perf_counter() - t_build:.1f} s: N={len(ph)} S={S} n_adc={seq.n_adc} "
        f"C={coils.shape[1]}")
    work = len(ph) * seq.n_adc
    nufft = (6, 3) if args.readouts == "nufft" else None
    K, Wm = args.steps, args.warmup
    leg_k, leg_w = max(1, min(K, args.leg_steps)), min(Wm, 1)
    adj_k, adj_w = max(1, min(K, args.adjoint_steps)), min(Wm, 1)

    rf_desc = (f"learned surrogate maps (network per spin; exact maps of every {sur.JUDGE_STRIDE}th spin as the "
               f"judge); readouts {args.readouts}" + (" (type-1 NUFFT, width 6, 3 decay terms)" if nufft else ""))
    line = {
        "metric": METRIC, "value": None, "unit": "spin·samples/s", "n_gpus": 1, "steps": K, "warmup": Wm,
        "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded phantom, coils and sequence of the configuration's shape)",
        "config": {
            "workload": f"3D GRE {args.n}^3 x {args.iso} isochromat(s) per voxel, three-class random phantom, "
                        f"+-100 Hz dB0, 2 ms sinc slab (TBW 4, 200 substeps), {args.coils} coils, TR 8 ms; forward",
            "spins": len(ph), "segments": S, "n_adc": seq.n_adc, "coils": int(coils.shape[1]),
            "rf": rf_desc, "pass": "forward",
            "value_is": "device-resident kernel throughput (inputs uploaded before the timed region); simulate() "
                        "timed end to end is under end_to_end",
        },
        "roofline": None, "cpu_baseline": None,
    }

    def emit():
        print(json.dumps(line), flush=True)

    # -- the value: the surrogate leg
    if args.adjoint_only:
        return adjoint_legs(args, line, emit, ph, seq, coils, nufft, lib, adjoint, driver, adj_k, adj_w, work)
    lead = None
    if not args.no_surrogate:
        lead = value_leg(line, ph, seq, coils, nufft, lib, driver, sur, K, Wm, work)

    # -- the CPU baseline (bounded sample); the GPU's parity on the same sample comes after the device legs
    if not args.no_cpu:
        from oracle import core  # the checker / CPU baseline only

        cores = cpu_quota()
        torch.set_num_threads(cores)
        ph_s, c_s = W.subset(ph, coils, min(args.cpu_spins, len(ph)))
        seq_s = W.first_trs(seq, args.cpu_trs)
        t1 = time.perf_counter()
        ref_sig, ref_m = core.simulate(ph_s, seq_s, c_s)
        t_cpu = time.perf_counter() - t1
        line["cpu_baseline"] = {
            "value": len(ph_s) * seq_s.n_adc / t_cpu, "unit": "spin·samples/s", "cores": cores, "kind": "port",
            "sample": f"FP64 oracle, {len(ph_s)} of the {len(ph)} spins (seeded subset) over the first "
                      f"{args.cpu_trs} TRs ({seq_s.n_adc} samples, {len(seq_s)} segments) of the same sequence; "
                      f"{t_cpu:.1f} s"}
        log(f"cpu oracle {line['cpu_baseline']['value']:.3e} spin-samples/s on {cores} cores ({t_cpu:.1f} s)")
    emit()

    # -- reused exact maps, same readouts
    case_r = driver.prepare_case(ph, seq, coils, reuse=True, nufft=nufft)
    line["reuse_exact_maps_forward"] = forward_leg(case_r, lib, driver, leg_k, leg_w, work, "reused exact maps")
    del case_r
    torch.cuda.empty_cache()
    emit()

    # -- exact RF: with the value's readouts, then per-sample readouts (the baseline)
    if not args.no_exact:
        if nufft is not None:
            case_e = driver.prepare_case(ph, seq, coils, nufft=nufft)
            line["exact_rf_nufft_forward"] = forward_leg(case_e, lib, driver, leg_k, leg_w, work,
                                                         "exact RF, NUFFT readouts")
            if lead is not None:
                line["speedup_vs_exact_rf_same_readouts"] = lead["value"] / line["exact_rf_nufft_forward"]["value"]
            del case_e
            torch.cuda.empty_cache()
            emit()
        case_e = driver.prepare_case(ph, seq, coils)
        line["exact_rf_forward"] = forward_leg(case_e, lib, driver, leg_k, leg_w, work, "exact RF")
        if lead is not None:
            line["speedup_vs_exact_rf"] = lead["value"] / line["exact_rf_forward"]["value"]
        del case_e
        torch.cuda.empty_cache()
        emit()

    # -- the surrogate's own workload: multi-slice GRE, surrogate vs exact RF
    if not args.no_multislice:
        line["multislice"] = multislice_leg(W, driver, sur, lib, leg_k, leg_w, nufft)
        emit()

    # -- the full variant: a bound pool and concomitant fields, forward and forward + adjoint
    if not args.no_full_variant:
        line["full_variant"] = full_variant_leg(W, driver, lib, leg_k, leg_w, adj_k, adj_w)
        torch.cuda.empty_cache()
        emit()

    # -- either side of the path: the sequence builder and segment table, the reconstruction
    if not args.no_downstream:
        line["downstream"] = downstream_leg(W, driver, args.n, args.coils, max(3, leg_k))
        emit()

    # -- forward + adjoint: the surrogate's maps with the value's readouts first, then exact RF
    if not args.no_adjoint:
        adjoint_legs(args, line, emit, ph, seq, coils, nufft, lib, adjoint, driver, adj_k, adj_w, work)

    # -- simulate() end to end
    if not args.no_e2e:
        line["end_to_end"] = end_to_end(bp, W, args, ph, seq, coils, work, adj_k, adj_w)
        emit()

    # -- the GPU on the CPU baseline's sample, every RF path, against the oracle
    if not args.no_cpu:
        parity = {"sample": "the CPU-baseline sample"}
        for name, opt in (("surrogate", "surrogate"), ("reuse", "reuse"), ("exact", "step")):
            ro = args.readouts if opt != "step" else "direct"
            o = bp.simulate(ph_s, seq_s, c_s, engine="gpu", pulses=opt, readouts=ro)
            parity[f"{name}_signal_rel_err"] = float((o.signal.to(torch.complex128) - ref_sig).abs().max()
                                                     / ref_sig.abs().max())
            parity[f"{name}_m_rel_err"] = float((o.m.double() - ref_m).abs().max() / ref_m.abs().max())
        line["parity"] = parity
        log("GPU vs oracle on the sample: " + ", ".join(f"{k[:-15]} {v:.2e}" for k, v in parity.items()
                                                        if k.endswith("signal_rel_err")))
        emit()

    if args.copy_probe:
        copy_probe(args.copy_probe)


if __name__ == "__main__":
    main()
