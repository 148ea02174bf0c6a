"""Gradients through the full variant: a rate bound pool, a scanner's concomitant fields and bound-only RF, through
``hb_forward_mt`` forward and ``hb_adjoint_full`` in reverse (``csrc/full_adjoint.hip``, ``csrc/full_device.h``).

Reference: ``_GPUSimulate`` / ``_simulate_grad`` with V = 1 (``src/blochparty/engines/gpu/adjoint.py:55-323``). The
contract is the reference's: inputs (pos, pd, t1, t2, db0, coils as real pairs (N, C, 2), dt, rf, grad, adc_phase,
m0 (N, 3, or N, 4 with the bound pool's Mzb), b1, b0_offset, conc (N, 6 or 7) the concomitant coefficients, bound
(N, 5) f, k, T1b, T2b, dw, rf_freq (S,)); outputs (signal (n_adc, C, 2), m (N, 3 or 4)); gradients for every input
that requires them, in its dtype and device: the encoding positions' (so a gradient nonlinearity's coefficients
through autograd), the concomitant coefficients' (so the true positions' through ``fields.spin_fields``), the bound
pool's (T1b = 1 / R1b), m0's (with the pool's default Mzb = f PD, through autograd onto f and PD), and every
segment's dt, rf, grad, b0_offset and RF frequency (so, through ``exchange.rf_frequency``, rf's phase and dt again).

Forward: the dispatches are exactly the checkpoint intervals; the state before each is kept (6 rows, plus Mzb and its
low part with a bound pool). Backward: dL/dsignal turned by exp(+i psi), then one ``hb_adjoint_full`` per interval,
last first (one sweep per tile of 16 coils, the sweep being linear in its sources), and ``hb_seg_reduce`` for the
per-segment gradients.
"""

from __future__ import annotations

import ctypes
from typing import NamedTuple

import torch
from torch.autograd.function import once_differentiable

from .. import exchange as X
from ..phantom import Phantom
from ..result import SimResult
from ..sequence import Sequence
from . import driver, hip, mt
from .adjoint import PASS_COILS, SEG_CHUNK, adjoint_sens, adjoint_width, default_checkpoint_every

N_GRADS = 20  # HB_FULL_GRADS: px py pz db0 pd r1 r2 b1, conc (7), f k r1b t2b dw


class HbFullAdjoint(ctypes.Structure):
    """``hb_full_adjoint`` of include/blochparty_amd_mt.h."""

    _fields_ = [("ckpt", ctypes.c_void_p), ("ckpt_mb", ctypes.c_void_p), ("traj", ctypes.c_void_p),
                ("lam", ctypes.c_void_p), ("gs", ctypes.c_void_p), ("gc", ctypes.c_void_p), ("gseg", ctypes.c_void_p),
                ("h", ctypes.c_void_p), ("sens", ctypes.c_void_p), ("nc", ctypes.c_int32), ("k_max", ctypes.c_int32),
                ("coil_grads", ctypes.c_int32), ("seg_grads", ctypes.c_int32)]


class Opts(NamedTuple):
    is_rf: list
    adc: list
    spoil: list
    bound_only: list | None
    every: int
    mode: int  # concomitant-field mode (0 none, 1 leading, 2 exact)
    lineshape: str
    cutoff: float | None
    memory: float


class FullSimulate(torch.autograd.Function):
    @staticmethod
    def forward(ctx, opts: Opts, pos, pd, t1, t2, db0, coils_ri, dt, rf, grad, adc_phase, m0, b1, b0_offset, conc,
                bound, rf_freq):
        ctx.set_materialize_grads(False)
        d = lambda t: None if t is None else t.detach()
        pool = None
        if bound is not None:
            bd = bound.detach()
            pool = X.BoundPool(bd[:, 0], bd[:, 1], bd[:, 2], bd[:, 3], bd[:, 4], lineshape=opts.lineshape,
                               cutoff=opts.cutoff, memory=opts.memory)
        ph = Phantom(d(pos), d(pd), d(t1), d(t2), d(db0), b1=d(b1), bound=pool)
        sq = Sequence(d(dt), d(rf), d(grad), d(adc_phase), opts.is_rf, opts.adc, opts.spoil, {}, d(b0_offset),
                      rf_freq=d(rf_freq), bound_only=opts.bound_only)
        coils = torch.complex(coils_ri[..., 0].detach(), coils_ri[..., 1].detach())
        case = driver.prepare_case(ph, sq, coils, d(m0), chunk=opts.every, aligned=True,
                                   conc=None if conc is None else (conc.detach(), opts.mode))
        ckpts = driver.run_forward(case, keep=True)
        hip.HIP.synchronize()
        res = driver.read_result(case)
        case.partial = None
        ctx.opts, ctx.case, ctx.ckpts = opts, case, ckpts
        ctx.coils = coils.cpu()
        ctx.like = [None if t is None else (t.dtype, t.device)
                    for t in (pos, pd, t1, t2, db0, coils_ri, dt, rf, grad, adc_phase, m0, b1, b0_offset, conc, bound,
                              rf_freq)]
        sig = torch.view_as_real(res.signal.T.contiguous())  # (n_adc, C, 2)
        m = res.m if res.m_bound is None else torch.cat([res.m, res.m_bound[:, None]], 1)
        ctx.save_for_backward(t1, t2, adc_phase, sig, bound, conc)
        return sig, m

    @staticmethod
    @once_differentiable
    def backward(ctx, g_sig, g_m):
        opts, case, ckpts = ctx.opts, ctx.case, ctx.ckpts
        t1, t2, adc_phase, sig, bound, conc = ctx.saved_tensors
        # pos, pd, t1, t2, db0, coils, dt, rf, grad, adc_phase, m0, b1, b0_offset, conc, bound, rf_freq
        need = list(ctx.needs_input_grad[1:])
        if g_sig is None:
            need[5] = need[9] = False
        if (g_sig is None and g_m is None) or not any(need):
            return (None,) * 17
        dev = case.spins.device
        n, n_pad, C, S = case.n, case.n_pad, case.n_coils, len(opts.is_rf)
        adc_idx = torch.tensor(opts.adc).nonzero()[:, 0]
        nc = adjoint_width(C) if g_sig is not None and case.n_adc else 0
        h = torch.zeros(max(case.n_adc, 1), max(nc, 1), 2, dtype=torch.float64)
        if nc:
            psi = adc_phase.detach().double().cpu()[adc_idx][:, None]
            gr, gi = g_sig[..., 0].double().cpu(), g_sig[..., 1].double().cpu()
            h[:, :C, 0] = gr * torch.cos(psi) - gi * torch.sin(psi)
            h[:, :C, 1] = gr * torch.sin(psi) + gi * torch.cos(psi)
        sens = adjoint_sens(ctx.coils, n_pad, nc)
        lam = torch.zeros(4, n_pad, dtype=torch.float32)
        if g_m is not None:
            lam[: g_m.shape[1], :n] = g_m.detach().float().cpu().T
        need_seg = any(need[i] for i in (6, 7, 8, 12, 15))
        need_coil = bool(nc) and need[5]
        gs, gc, seg, lam_d = run_backward(case, ckpts, opts.every, h.float().to(dev).contiguous(),
                                          lam.to(dev).contiguous(), sens.to(dev).contiguous(), nc, need_seg,
                                          need_coil)
        hip.HIP.synchronize()
        GS = gs.cpu()[:, :n]
        lam = lam_d.cpu().double()[:, :n]
        driver.check_output(case, GS, lam, *(t for t in (gc, seg) if t is not None),
                            premise=bool(torch.isfinite(h).all()) and (g_m is None or bool(torch.isfinite(g_m).all())))
        out = [None] * 16
        if need[0]:
            out[0] = GS[0:3].T.contiguous()
        if need[1]:
            out[1] = GS[4]
        if need[2]:
            out[2] = -GS[5] / t1.detach().double().cpu() ** 2  # r1 = 1 / t1
        if need[3]:
            out[3] = -GS[6] / t2.detach().double().cpu() ** 2
        if need[4]:
            out[4] = GS[3]
        if need_coil:
            G = gc.cpu()[: 2 * C, :n]
            out[5] = torch.stack([G[0::2].T, G[1::2].T], -1)
        if need_seg:
            sgr = seg.cpu().view(S, 8)
            out[6], out[7], out[8], out[12], out[15] = sgr[:, 0], sgr[:, 1:3], sgr[:, 3:6], sgr[:, 6], sgr[:, 7]
        if need[9]:  # signal = raw exp(-i psi): dsignal/dpsi = -i signal
            g = torch.zeros(S, dtype=torch.float64)
            sg_, gsig = sig.double(), g_sig.double().cpu()
            g[adc_idx] = (gsig[..., 0] * sg_[..., 1] - gsig[..., 1] * sg_[..., 0]).sum(-1)
            out[9] = g
        if need[10]:
            out[10] = lam[: 4 if bound is not None else 3].T.contiguous()
        if need[11]:
            out[11] = GS[7]
        if need[13]:
            out[13] = GS[8 : 8 + conc.shape[1]].T.contiguous()
        if need[14]:  # f, k, t1b (r1b = 1 / t1b), t2b, dw
            t1b = bound.detach().double().cpu()[:, 2]
            out[14] = torch.stack([GS[15], GS[16], -GS[17] / t1b**2, GS[18], GS[19]], -1)
        return (None, *[None if g is None or not need[i] else g.to(dtype=ctx.like[i][0], device=ctx.like[i][1])
                        for i, g in enumerate(out)])


def run_backward(case: driver.DeviceCase, ckpts: list, every: int, h: torch.Tensor, lam: torch.Tensor,
                 sens: torch.Tensor, nc: int, need_seg: bool, need_coil: bool):
    """``hb_adjoint_full`` over every interval of ``case`` (aligned dispatches, forward run with ``keep``): ``h``
    (n_adc, nc, 2) dL/dsignal turned by exp(+i psi), ``lam`` (4, n_pad) dL/d(Mx, My, Mz, Mzb) at the end (it ends as
    dL/dm0), ``sens`` (2 nc, n_pad). Returns gs (N_GRADS, n_pad), gc (2 nc, n_pad), seg (S, 8) float64 and lam. More
    than PASS_COILS coils: one sweep per tile of PASS_COILS coils, dL/dm at the end entering the first only."""
    dev = case.spins.device
    n_pad, S = case.n_pad, case.n_segments
    gs = torch.zeros(N_GRADS, n_pad, dtype=torch.float64, device=dev)
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
        _sweep(case, ckpts, every, h_t, lam_t, sens_t, width, need_seg, need_coil, gs, gc_t, seg_t)
        if passes > 1:
            gc[2 * cols.start : 2 * cols.stop] = gc_t
            if t:
                seg += seg_t
            lam_sum = lam_t.clone() if lam_sum is None else lam_sum + lam_t
    if passes > 1:
        lam.copy_(lam_sum)
    return gs, gc, seg, lam


def _sweep(case, ckpts, every, h, lam, sens, nc, need_seg, need_coil, gs, gc, seg) -> None:
    dev = case.spins.device
    n_pad = case.n_pad
    n_blocks = n_pad // driver.BLOCK
    lib = hip.load()
    mt.bind(lib)
    kern = hip.HIP.lib()
    cs = case.struct()
    mts = mt.struct(case.extra["mt"])
    pool = case.extra["mt"]["pool"] is not None
    traj = torch.empty(every * 4 * n_pad, dtype=torch.float32, device=dev)
    gseg = torch.empty(n_blocks * every * 8 if need_seg else 1, dtype=torch.float32, device=dev)
    red = torch.empty(max(1, -(-n_blocks // SEG_CHUNK) * every * 8 if need_seg else 1), dtype=torch.float64,
                      device=dev)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for c in reversed(range(len(ckpts))):
        s0, s1 = case.starts[c], case.ends[c]
        ck = ckpts[c]
        ad = HbFullAdjoint(ck.data_ptr(), ck[6].data_ptr() if pool else None, traj.data_ptr(), lam.data_ptr(),
                           gs.data_ptr(), gc.data_ptr(), gseg.data_ptr(), h.data_ptr(), sens.data_ptr(), nc, every,
                           int(need_coil), int(need_seg))
        rc = lib.hb_adjoint_full(ctypes.byref(cs), ctypes.byref(mts), ctypes.byref(ad), s0, s1, stream)
        if rc != 0:
            raise RuntimeError(f"HIP error {rc} in hb_adjoint_full: {lib.hb_error_string(rc).decode()}")
        if need_seg:
            kern.seg_reduce(gseg.data_ptr(), n_blocks, every, s0, s1 - s0, red.data_ptr(), seg.data_ptr(), None)


def simulate_grad(phantom: Phantom, seq: Sequence, coils, m_init, checkpoint_every, conc=None) -> SimResult:
    """simulate(engine="gpu") with gradients through the full variant (adjoint.py:264-323 with V = 1): ``conc`` =
    (coefficients (N, 6 or 7), mode) of a scanner's concomitant fields, or None."""
    n, pd = len(phantom), phantom.pd
    if coils is None:
        coils = torch.ones(n, 1, dtype=torch.complex64 if pd.dtype == torch.float32 else torch.complex128)
    bound, rf_freq, lineshape, cutoff, memory = None, None, "superlorentzian", None, 300.0
    mb = None
    if phantom.bound is not None:
        mt.check_pool(phantom.bound)
        b = phantom.bound.expand(n, pd)
        bound = torch.stack([b.f, b.k, b.t1, b.t2, b.dw], -1)
        rf_freq = X.rf_frequency(seq)
        lineshape, cutoff, memory = b.lineshape, b.cutoff, float(b.memory)
        mb = m_init[:, 3] if m_init is not None and m_init.shape[1] == 4 else b.f * pd
    if m_init is None:
        cols = [torch.zeros_like(pd), torch.zeros_like(pd), pd]
    else:
        cols = [m_init[:, 0], m_init[:, 1], m_init[:, 2]]
    m0 = torch.stack(cols + ([] if mb is None else [mb.to(pd.dtype)]), -1)
    every = checkpoint_every or default_checkpoint_every(len(seq))
    if every % 64:
        raise ValueError(f"checkpoint_every must be a multiple of 64, got {every}")
    coef, mode = (None, 0) if conc is None else conc
    opts = Opts(list(seq.is_rf), list(seq.adc), list(seq.spoil), None if seq.bound_only is None else
                list(seq.bound_only), every, int(mode), lineshape, cutoff, memory)
    sig, m = FullSimulate.apply(opts, phantom.pos, pd, phantom.t1, phantom.t2, phantom.db0, torch.view_as_real(coils),
                                seq.dt, seq.rf, seq.grad, seq.adc_phase, m0, phantom.b1, seq.b0_offset, coef, bound,
                                rf_freq)
    signal = torch.view_as_complex(sig.contiguous()).T
    if bound is not None:
        return SimResult(signal, m[:, :3], m[:, 3])
    return SimResult(signal, m)


__all__ = ["FullSimulate", "HbFullAdjoint", "N_GRADS", "run_backward", "simulate_grad"]
