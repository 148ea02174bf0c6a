"""The GPU engine's gradients: a ``torch.autograd.Function`` around the forward kernels with checkpoints and the
hand-written adjoint kernels (reference: ``src/blochparty/engines/gpu/adjoint.py:23-323``).

The contract is the reference's ``_GPUSimulate``: inputs (pos, pd, t1, t2, db0, coils as real pairs (N, C, 2), dt, rf,
grad, adc_phase, m0, b1, b0_offset); outputs (signal (n_adc, C, 2), m (N, 3)), float32 on the CPU; first-order
gradients for every input that requires them, each in its input's dtype and device; ``checkpoint_every`` a multiple
of 64 (default about sqrt(segments)), else ``ValueError``.

Forward: the dispatches are exactly the checkpoint intervals, the state before each is kept on the device (hi + lo,
24 bytes per spin per interval). Backward: dL/dsignal turned by exp(+i psi) (the source of the sweep), then one
``hb_adjoint_interval`` per interval, last first, and ``hb_seg_reduce`` for the per-segment gradients. Parameter
gradients map as the reference's (adjoint.py:222-261): t1 = 1 / r1 (-dL/dr1 / t1^2), t2 likewise, d/dpsi from the
signal itself (dsignal/dpsi = -i signal), dL/dm0 = lambda at the start.

Beyond the reference (whose reused pulses and NUFFT readouts are forward only, driver.py:221-225): with
``pulses="reuse"`` / ``"surrogate"`` each occurrence of a reused pulse is one map in both sweeps (the sweep sums seeds
for the maps' parameter gradients, ``hb_pulse_map_grads``, and each occurrence's amplitude and RF-phase gradient,
``project_pulse_rf``); with ``readouts="nufft"`` each readout run is reversed as the transpose of its NUFFT
(``hb_nufft_adjoint_grid`` and the sweep's interpolation). Either refuses gradients for dt, grad and b0_offset.
"""

from __future__ import annotations

import ctypes
import math
from typing import NamedTuple

import numpy as np
import torch
from torch.autograd.function import once_differentiable

from ..phantom import Phantom
from ..result import SimResult
from ..sequence import Sequence
from . import driver, hip

ADJOINT_COILS = (1, 2, 4, 8, 16)  # coils held in registers by the adjoint kernel (padded up)
SEG_CHUNK = 256  # workgroups per first-pass chunk of hb_seg_reduce


def default_checkpoint_every(n_segments: int, bodies: bool = False) -> int:
    """About sqrt(segments), a multiple of 64. ``bodies``: reused pulses or NUFFT runs make most segments identity
    records that the sweep skips, so 4x that (fewer, longer launches: 7 % faster on the 64^3 GRE; the gradients are
    the same bit for bit for any interval)."""
    base = 64 * max(1, math.ceil(math.sqrt(max(n_segments, 1)) / 64))
    return 4 * base if bodies else base


class Opts(NamedTuple):
    is_rf: list
    adc: list
    spoil: list
    every: int
    pulses: str = "step"  # "reuse" / "surrogate": reused pulses as maps in the forward and the reverse sweep
    nufft: tuple | None = None  # (width, terms): readout runs as NUFFTs in the forward and the reverse sweep


class GPUSimulate(torch.autograd.Function):
    @staticmethod
    def forward(ctx, opts: Opts, pos, pd, t1, t2, db0, coils_ri, dt, rf, grad, adc_phase, m0, b1, b0_offset):
        ctx.set_materialize_grads(False)
        d = lambda t: None if t is None else t.detach()
        ph = Phantom(d(pos), d(pd), d(t1), d(t2), d(db0), b1=d(b1))
        sq = Sequence(d(dt), d(rf), d(grad), d(adc_phase), opts.is_rf, opts.adc, opts.spoil, {}, d(b0_offset))
        coils = torch.complex(coils_ri[..., 0].detach(), coils_ri[..., 1].detach())
        case = driver.prepare_case(ph, sq, coils, d(m0), chunk=opts.every, aligned=True,
                                   reuse=opts.pulses == "reuse", surrogate=opts.pulses == "surrogate",
                                   nufft=opts.nufft)
        ckpts = driver.run_forward(case, keep=True)
        hip.HIP.synchronize()
        res = driver.read_result(case)
        case.partial = None  # the forward's work buffer is not needed any more
        ctx.opts, ctx.case, ctx.ckpts = opts, case, ckpts
        ctx.rf = rf.detach().double().cpu().numpy() if case.pulses else None
        ctx.coils = coils.cpu()
        ctx.like = [None if t is None else (t.dtype, t.device)
                    for t in (pos, pd, t1, t2, db0, coils_ri, dt, rf, grad, adc_phase, m0, b1, b0_offset)]
        sig = torch.view_as_real(res.signal.T.contiguous())  # (n_adc, C, 2)
        ctx.save_for_backward(t1, t2, adc_phase, sig)
        return sig, res.m

    @staticmethod
    @once_differentiable
    def backward(ctx, g_sig, g_m):
        opts, case, ckpts = ctx.opts, ctx.case, ctx.ckpts
        t1, t2, adc_phase, sig = ctx.saved_tensors
        # pos, pd, t1, t2, db0, coils, dt, rf, grad, adc_phase, m0, b1, b0_offset
        need = list(ctx.needs_input_grad[1:])
        if g_sig is None:
            need[5] = need[9] = False
        if (g_sig is None and g_m is None) or not any(need):
            return (None,) * 14
        dev = case.spins.device
        n, n_pad, C, S, every = case.n, case.n_pad, case.n_coils, len(opts.is_rf), opts.every
        adc_idx = torch.tensor(opts.adc).nonzero()[:, 0]
        nc = adjoint_width(C) if g_sig is not None and case.n_adc else 0
        h = torch.zeros(max(case.n_adc, 1), max(nc, 1), 2, dtype=torch.float64)
        if nc:
            psi = adc_phase.detach().double().cpu()[adc_idx][:, None]
            gr, gi = g_sig[..., 0].double().cpu(), g_sig[..., 1].double().cpu()
            h[:, :C, 0] = gr * torch.cos(psi) - gi * torch.sin(psi)
            h[:, :C, 1] = gr * torch.sin(psi) + gi * torch.cos(psi)
        # with NUFFT readouts the device holds the spins sorted (case.perm: device k <- caller spin perm[k])
        perm = None if case.perm is None else torch.from_numpy(np.ascontiguousarray(case.perm))
        inv = None
        if perm is not None:
            inv = torch.empty_like(perm)
            inv[perm] = torch.arange(len(perm))
        to_caller = (lambda t: t) if inv is None else (lambda t: t[:, inv])
        sens = adjoint_sens(ctx.coils if perm is None else ctx.coils[perm], n_pad, nc)
        lam = torch.zeros(3, n_pad, dtype=torch.float32)
        if g_m is not None:
            gm = g_m.detach().float().cpu()
            lam[:, :n] = (gm if perm is None else gm[perm]).T
        need_seg = any(need[i] for i in (6, 7, 8, 12))
        # 2: only rf's gradients (the sweep forms segment gradients only where an RF record is)
        seg_mode = 0 if not need_seg else 1 if any(need[i] for i in (6, 8, 12)) else 2
        need_coil = bool(nc) and need[5]
        gs, gc, seg, lam_d = run_backward(case, ckpts, every, h.float().to(dev).contiguous(), lam.to(dev).contiguous(),
                                          sens.to(dev).contiguous(), nc, seg_mode, need_coil)
        hip.HIP.synchronize()
        GS = to_caller(gs.cpu()[:, :n])
        lam = to_caller(lam_d.cpu().double()[:, :n])
        driver.check_output(case, GS, lam, *(t for t in (gc, seg) if t is not None),
                            premise=bool(torch.isfinite(h).all()) and (g_m is None or bool(torch.isfinite(g_m).all())))
        out = [None] * 13
        if need[0]:
            out[0] = GS[0:3].T.contiguous()
        if need[1]:
            out[1] = GS[4]
        if need[2]:
            out[2] = -GS[5] / t1.detach().double().cpu() ** 2
        if need[3]:
            out[3] = -GS[6] / t2.detach().double().cpu() ** 2
        if need[4]:
            out[4] = GS[3]
        if need_coil:
            G = to_caller(gc.cpu()[: 2 * C, :n])
            out[5] = torch.stack([G[0::2].T, G[1::2].T], -1)
        if need_seg:
            sgr = seg.cpu().view(S, 8)
            if case.pulses:
                project_pulse_rf(sgr, ctx.rf, case.pulses)
            out[6], out[7], out[8], out[12] = sgr[:, 0], sgr[:, 1:3], sgr[:, 3:6], sgr[:, 6]
        if need[9]:
            g = torch.zeros(S, dtype=torch.float64)
            sg_, gsig = sig.double(), g_sig.double().cpu()
            g[adc_idx] = (gsig[..., 0] * sg_[..., 1] - gsig[..., 1] * sg_[..., 0]).sum(-1)
            out[9] = g
        if need[10]:
            out[10] = lam.T.contiguous()
        if need[11]:
            out[11] = GS[7]
        return (None, *[None if g is None or not need[i] else g.to(dtype=ctx.like[i][0], device=ctx.like[i][1])
                        for i, g in enumerate(out)])


PASS_COILS = ADJOINT_COILS[-1]  # coils per sweep; more coils take one sweep per tile of this many


def adjoint_width(n_coils: int) -> int:
    """Coil columns of dL/dsignal and sensitivity rows for ``n_coils``: the register width that holds them, or for
    more than PASS_COILS a whole number of PASS_COILS-wide tiles (one sweep each)."""
    if n_coils <= PASS_COILS:
        return next(c for c in ADJOINT_COILS if c >= n_coils)
    return PASS_COILS * -(-n_coils // PASS_COILS)


def run_backward(case: driver.DeviceCase, ckpts: list, every: int, h: torch.Tensor, lam: torch.Tensor,
                 sens: torch.Tensor, nc: int, need_seg: int, need_coil: bool, lib=None):
    """The adjoint kernels on the current stream over ``case`` (aligned dispatches, forward run with ``keep``):
    ``h`` (n_adc, nc, 2) float32 dL/dsignal turned by exp(+i psi), ``lam`` (3, n_pad) float32 dL/dm at the end (it
    ends as dL/dm0), ``sens`` (2 nc, n_pad), nc from ``adjoint_width``. ``need_seg``: 0 no segment gradients, 1 all,
    2 only rf's (the other slots come back 0). Returns the float64 accumulators (gs (8, n_pad), gc (2 nc, n_pad),
    seg (S, 8)) and lam.

    More than PASS_COILS coils: the sweep is linear in its sources (lambda and every gradient are linear in
    dL/dsignal and in dL/dm at the end), so it runs once per tile of PASS_COILS coils with that tile's columns of h
    and rows of the sensitivities, dL/dm at the end entering the first pass only; the float64 accumulators and the
    per-spin seeds of the pulse maps add up over the passes, and the maps' parameter gradients are formed once from
    the summed seeds."""
    need_seg = int(need_seg)
    lib = lib or hip.HIP.lib()
    dev = case.spins.device
    n_pad, S = case.n_pad, case.n_segments
    P = case.n_pulses
    cs = case.struct()
    seeds = torch.zeros(P * 12 * n_pad, dtype=torch.float64, device=dev) if P else None
    maps_d = None
    steps = map_fd_steps(case.pulse_table.cpu().numpy(), case.pulse_index.cpu().numpy()) if P else None
    if P and need_seg:  # the maps' derivative along each pulse's amplitude, for its occurrences' rf gradients
        maps_d = torch.empty(P * 12 * n_pad, dtype=torch.float32, device=dev)
        lib.pulse_map_dlnb1(ctypes.byref(cs), case.pulse_table.data_ptr(), case.pulse_index.data_ptr(), P,
                            float(steps[:, 8].min()), maps_d.data_ptr())
    gs = torch.zeros(8, n_pad, dtype=torch.float64, device=dev)
    gc = torch.zeros(2 * max(nc, 1), n_pad, dtype=torch.float64, device=dev)
    seg = torch.zeros(S * 8 if need_seg else 1, dtype=torch.float64, device=dev)
    width = min(nc, PASS_COILS) if nc else 0
    passes = max(1, -(-nc // PASS_COILS)) if nc else 1
    lam_sum = None
    for t in range(passes):
        cols = slice(t * width, (t + 1) * width)
        h_t = h if passes == 1 else h[:, cols].contiguous()
        sens_t = sens if passes == 1 else sens[2 * cols.start : 2 * cols.stop].contiguous()
        gc_t = gc if passes == 1 else torch.zeros(2 * width, n_pad, dtype=torch.float64, device=dev)
        seg_t = seg if t == 0 else torch.zeros_like(seg)
        lam_t = lam if t == 0 else torch.zeros_like(lam)
        _sweep(case, cs, ckpts, every, h_t, lam_t, sens_t, width, need_seg, need_coil, lib, seeds, maps_d, gs, gc_t,
               seg_t)
        if passes > 1:
            gc[2 * cols.start : 2 * cols.stop] = gc_t
            if t:
                seg += seg_t
            lam_sum = lam_t.clone() if lam_sum is None else lam_sum + lam_t
    if passes > 1:
        lam.copy_(lam_sum)
    if P:  # the spin parameters' gradients through the maps, from the seeds of the sweep(s)
        st = torch.from_numpy(np.ascontiguousarray(steps[:, :8])).to(dev)
        lib.pulse_map_grads(ctypes.byref(cs), case.pulse_table.data_ptr(), case.pulse_index.data_ptr(), P,
                            st.data_ptr(), seeds.data_ptr(), gs.data_ptr())
    return gs, gc, seg, lam


def _sweep(case, cs, ckpts, every, h, lam, sens, nc, need_seg, need_coil, lib, seeds, maps_d, gs, gc, seg) -> None:
    """One reverse sweep over every checkpoint interval, last first, with ``nc`` (<= PASS_COILS) coils."""
    dev = case.spins.device
    n_pad = case.n_pad
    n_blocks = n_pad // driver.BLOCK
    gseg = torch.zeros(n_blocks * every * 8 if need_seg else 1, dtype=torch.float32, device=dev)
    r8 = torch.empty((every + 7) // 8 * 3 * n_pad, dtype=torch.float32, device=dev)
    red = torch.empty(max(1, -(-n_blocks // SEG_CHUNK) * every * 8 if need_seg else 1), dtype=torch.float64,
                      device=dev)
    ptr = lambda t: None if t is None else t.data_ptr()
    nu, grid = case.nu, None
    if nu is not None and nc:  # the runs' adjoint grids from dL/dsignal (the transpose of hb_nufft_signal)
        pl = nu["plan"]
        grid = torch.empty(len(pl.runs) * pl.terms * pl.ng * 2 * nc, dtype=torch.float32, device=dev)
        lib.nufft_adjoint_grid(h.data_ptr(), nc, nu["row_of"].data_ptr(), len(pl.runs), pl.n, pl.ng, pl.terms,
                               nu["nu_tw"].data_ptr(), nu["nu_lw"].data_ptr(), grid.data_ptr())
    for c in reversed(range(len(ckpts))):
        s0, s1 = case.starts[c], case.ends[c]
        ad = hip.HbAdjoint(ckpts[c].data_ptr(), r8.data_ptr(), lam.data_ptr(), gs.data_ptr(), gc.data_ptr(),
                           gseg.data_ptr(), h.data_ptr(), sens.data_ptr(), ptr(maps_d), ptr(seeds),
                           None if nu is None else nu["nu"].data_ptr(), None if nu is None else nu["nu_win"].data_ptr(),
                           ptr(grid), nc, int(need_seg), int(need_coil), every)
        lib.adjoint_interval(ctypes.byref(cs), ctypes.byref(ad), s0, s1)
        if need_seg:
            lib.seg_reduce(gseg.data_ptr(), n_blocks, every, s0, s1 - s0, red.data_ptr(), seg.data_ptr(),
                           case.segs.data_ptr())


FD_TURN = 1e-4  # rad: how far a central-difference step turns a reused pulse's rotation


def map_fd_steps(table: np.ndarray, index: np.ndarray) -> np.ndarray:
    """(P, 9) central-difference steps per reused pulse: for px, py, pz, db0, pd, r1, r2, b1 (hb_pulse_map_grads'
    order; 0 where the map does not depend on the parameter), then the relative step along the amplitude. Each turns
    the pulse's rotation (or moves its relaxation exponent) by about FD_TURN: truncation ~FD_TURN^2 / 6 relative,
    rounding of the float64 composition ~1e-16 / FD_TURN."""
    steps = np.zeros((len(index), 9))
    for p, (off, k) in enumerate(index):
        rows = table[off : off + k]
        T = max(float(rows[:, 0].sum()), 1e-300)
        angle = 2 * math.pi * float((rows[:, 0] * np.hypot(rows[:, 1], rows[:, 2])).sum())
        gmax = np.abs(rows[:, 3:6]).max(0)
        for a in range(3):
            steps[p, a] = FD_TURN / (2 * math.pi * T * gmax[a]) if gmax[a] > 0 else 0.0
        steps[p, 3] = FD_TURN / (2 * math.pi * T)
        steps[p, 4] = 0.5  # the map is affine in pd: any step is exact
        steps[p, 5] = steps[p, 6] = FD_TURN / T
        rel = min(FD_TURN / max(angle, FD_TURN), 1e-3)
        steps[p, 7] = rel  # b1 ~ 1
        steps[p, 8] = rel
    return steps


def project_pulse_rf(sgr: torch.Tensor, rf: np.ndarray, pulses) -> None:
    """A reused pulse's occurrence has, from the sweep, dL/d(ln amplitude) c and dL/d(RF phase) d in its first
    segment's rf slots. Its segments' rf gradients become the least-norm ones with those two components:
    (c rf_k + d i rf_k) / sum |rf|^2, exact for a loss through each occurrence's complex amplitude (flip angle, RF
    phase); the pulse's shape, dt, gradients and b0_offset get none (simulate refuses those gradients)."""
    g = sgr.numpy()
    for _, k, occ in pulses:
        for a, _ in occ:
            c, d = float(g[a, 1]), float(g[a, 2])
            z = rf[a : a + k]
            nrm = float((z * z).sum())
            g[a : a + k] = 0.0
            if nrm > 0:
                g[a : a + k, 1] = (c * z[:, 0] - d * z[:, 1]) / nrm
                g[a : a + k, 2] = (c * z[:, 1] + d * z[:, 0]) / nrm


def adjoint_sens(coils: torch.Tensor, n_pad: int, nc: int) -> torch.Tensor:
    """(N, C) complex sensitivities as the adjoint's (2 nc, n_pad) float32 rows Re S_0, Im S_0, Re S_1, ..."""
    n, C = coils.shape
    sens = torch.zeros(2 * max(nc, 1), n_pad, dtype=torch.float32)
    if nc:
        sens[0 : 2 * C : 2, :n] = coils.real.T.float().cpu()
        sens[1 : 2 * C : 2, :n] = coils.imag.T.float().cpu()
    return sens


def simulate_grad(phantom: Phantom, seq: Sequence, coils, m_init, checkpoint_every, pulses: str = "step",
                  nufft=None, rf_grad: str | None = None) -> SimResult:
    """simulate(engine="gpu") with gradients: the inputs as tensors of ``GPUSimulate`` (adjoint.py:264-323).
    ``pulses``: "step", or "reuse" / "surrogate" (reused pulses as maps in both sweeps); ``nufft``: None or (width,
    terms), readout runs as NUFFTs in both sweeps (the reference's reused pulses and NUFFT readouts are forward only,
    driver.py:221-225). ``rf_grad``: "scale" accepts that, with maps, rf's gradient is each occurrence's
    complex-amplitude component (``project_pulse_rf``); without it a reused pulse with rf requiring grad raises
    ValueError, since a loss through the pulse's shape would get a wrong gradient."""
    n, pd = len(phantom), phantom.pd
    if coils is None:
        coils = torch.ones(n, 1, dtype=torch.complex64 if pd.dtype == torch.float32 else torch.complex128)
    if m_init is None:
        m0 = torch.stack([torch.zeros_like(pd), torch.zeros_like(pd), pd], -1)
    else:
        m0 = m_init[:, :3]
    every = checkpoint_every or default_checkpoint_every(len(seq), bodies=pulses != "step" or nufft is not None)
    if every % 64:
        raise ValueError(f"checkpoint_every must be a multiple of 64, got {every}")
    if pulses != "step" or nufft is not None:
        fixed = [n for n, t in (("dt", seq.dt), ("grad", seq.grad), ("b0_offset", seq.b0_offset))
                 if t is not None and t.requires_grad]
        if fixed:
            what = f"pulses={pulses!r}" if pulses != "step" else "readouts='nufft'"
            raise ValueError(f"{what} differentiates spin parameters, coils, m_init and RF amplitudes and phases, not "
                             f"{', '.join(fixed)}: use pulses='step' and readouts='direct'")
    if pulses != "step" and seq.rf.requires_grad and rf_grad != "scale":
        raise ValueError(f"pulses={pulses!r} gives rf only each reused pulse occurrence's complex-amplitude gradient "
                         f"(exact for flip angles and RF phases, not for a pulse's shape): pass rf_grad='scale' to "
                         f"accept that, or use pulses='step' for every rf sample's gradient")
    opts = Opts(list(seq.is_rf), list(seq.adc), list(seq.spoil), every, pulses, nufft)
    sig, m = GPUSimulate.apply(opts, phantom.pos, pd, phantom.t1, phantom.t2, phantom.db0, torch.view_as_real(coils),
                               seq.dt, seq.rf, seq.grad, seq.adc_phase, m0, phantom.b1, seq.b0_offset)
    return SimResult(torch.view_as_complex(sig.contiguous()).T, m)
