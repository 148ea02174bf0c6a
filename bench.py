"""Benchmark of the GPU path on one MI355X: spin-samples/s of simulate(..., engine="gpu") on the 3D GRE 64^3 workload.

    python bench.py [--steps K] [--warmup W] [--n 64] [--no-cpu]

One step = one forward pass over the workload with every input resident in HBM (uploaded before the timed region);
the throughput is spins x ADC samples x steps / seconds (the reference's metric, benchmarks/bench.py:65). The timed
region is bracketed by torch.cuda.synchronize(). Inside it, every forward-kernel launch is bracketed by HIP events on
the stream it runs on, for the roofline of that kernel (algorithmic flops per launch over its mean duration).

The CPU baseline is the FP64 oracle timed on this host's cores (the container quota) on a bounded sample of the same
workload: a seeded subset of spins over the first TRs; the GPU runs the same sample and reports its relative error
against the oracle. One JSON line goes to stdout.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np
import torch

METRIC = "spin·samples/s fwd & fwd+adjoint, 3D GRE 64³; RF surrogate vs exact @1e-4"
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector (published)
MFMA_FP32_PEAK_TFLOPS = 157.3  # MI355X dense FP32 matrix (published)

# counting convention of the algorithmic work (per spin)
FLOPS_RF = 48  # exact RF substep
FLOPS_FREE = 12  # free or ADC event
FLOPS_COIL = 8  # per coil, per spin-sample (complex MAC)


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


def dispatch_flops(case, seq_flags: np.ndarray) -> list[float]:
    """Algorithmic flops of every forward dispatch of ``case`` (counting convention above)."""
    rf = (seq_flags & 1) != 0
    adc = (seq_flags & 2) != 0
    per_seg = np.where(rf, FLOPS_RF, FLOPS_FREE) + adc * FLOPS_COIL * case.n_coils
    cum = np.concatenate([[0], np.cumsum(per_seg, dtype=np.float64)])
    return [float(cum[e] - cum[s]) * case.n for s, e in zip(case.starts, case.ends)]


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=64, help="grid size of the 3D GRE")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and parity sample")
    ap.add_argument("--cpu-trs", type=int, default=32)
    ap.add_argument("--cpu-spins", type=int, default=65536)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: the path shards over spins, but this build measures one GPU's share "
              f"(--gpus {args.gpus} is not supported)", file=sys.stderr)
        sys.exit(2)

    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import blochparty_amd as bp
    from blochparty_amd import workloads as W
    from blochparty_amd.engines import driver, hip

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    hip.load()
    lib = hip.HIP.lib()
    t_build = time.perf_counter()
    ph, seq, coils = W.gre3d(args.n)
    case = driver.prepare_case(ph, seq, coils)
    torch.cuda.synchronize()
    print(f"# workload built and uploaded in {time.perf_counter() - t_build:.1f} s: N={case.n} S={len(seq)} "
          f"n_adc={case.n_adc} C={case.n_coils} dispatches={len(case.starts)}", file=sys.stderr, flush=True)
    flags = case.segs.cpu().numpy().view(driver.SEG_DTYPE)["flags"]
    flops = dispatch_flops(case, flags)

    for _ in range(args.warmup):
        driver.run_forward(case, lib)
    torch.cuda.synchronize()
    print("# warmup done", file=sys.stderr, flush=True)

    events = []

    def timer(phase):
        ev = torch.cuda.Event(enable_timing=True)
        ev.record()
        events.append(ev)

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        driver.run_forward(case, lib, timer=timer)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    kernel_ms = [events[2 * i].elapsed_time(events[2 * i + 1]) for i in range(len(events) // 2)]
    ms_step = 1e3 * elapsed / args.steps
    value = case.n * case.n_adc * args.steps / elapsed
    mean_ms = float(np.mean(kernel_ms))
    mean_flops = float(np.mean(flops))
    achieved = mean_flops / (mean_ms * 1e-3) / 1e12
    coil_flops = FLOPS_COIL * case.n_coils * case.n * case.n_adc
    print(f"# forward: {ms_step:.1f} ms/step, {value:.3e} spin-samples/s; forward kernel mean {mean_ms:.3f} ms over "
          f"{len(kernel_ms)} launches, {achieved:.1f} TFLOP/s algorithmic", file=sys.stderr, flush=True)

    cpu = None
    parity = None
    if not args.no_cpu:
        from oracle import core  # the checker / CPU baseline only

        cores = cpu_quota()
        torch.set_num_threads(cores)
        ph_s, c_s = W.subset(ph, coils, min(args.cpu_spins, len(ph)))
        seq_s = W.first_trs(seq, args.cpu_trs)
        t1 = time.perf_counter()
        ref_sig, ref_m = core.simulate(ph_s, seq_s, c_s)
        t_cpu = time.perf_counter() - t1
        cpu = {
            "value": len(ph_s) * seq_s.n_adc / t_cpu,
            "unit": "spin·samples/s",
            "cores": cores,
            "kind": "port",
            "sample": f"FP64 oracle, {len(ph_s)} of the {case.n} spins (seeded subset) over the first {args.cpu_trs} "
                      f"TRs ({seq_s.n_adc} samples, {len(seq_s)} segments) of the same sequence; {t_cpu:.1f} s",
        }
        out = bp.simulate(ph_s, seq_s, c_s, engine="gpu")
        err_sig = float((out.signal.to(torch.complex128) - ref_sig).abs().max() / ref_sig.abs().max())
        err_m = float((out.m.double() - ref_m).abs().max() / ref_m.abs().max())
        parity = {"signal_rel_err": err_sig, "m_rel_err": err_m, "sample": "the CPU-baseline sample"}
        print(f"# cpu oracle {cpu['value']:.3e} spin-samples/s on {cores} cores; GPU vs oracle on the sample: "
              f"signal {err_sig:.2e}, m {err_m:.2e}", file=sys.stderr, flush=True)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "spin·samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"3D GRE {args.n}^3 x 4 sub-voxel isochromats, three-class random phantom, +-100 Hz dB0, "
                        f"2 ms sinc slab (TBW 4, 200 substeps, exact RF), 8 coils, TR 8 ms, forward",
            "spins": case.n,
            "segments": len(seq),
            "n_adc": case.n_adc,
            "coils": case.n_coils,
            "rf": "exact",
            "pass": "forward",
        },
        "roofline": {
            "bound": "valu",
            "kernel": "forward_kernel",
            "achieved": achieved,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS,
            "traffic": None,
            "mean_launch_ms": mean_ms,
            "launches": len(kernel_ms),
            "flops_per_launch": mean_flops,
            "coil_mfma_tflops": coil_flops * args.steps / (sum(kernel_ms) * 1e-3) / 1e12,
        },
        "cpu_baseline": cpu,
        "parity": parity,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
