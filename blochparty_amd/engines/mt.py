"""The full-variant forward path on the GPU: a rate bound pool (``Phantom.bound``), a scanner's concomitant fields
and bound-only RF, through ``hb_forward_mt`` (``csrc/mt.hip``, ``include/blochparty_amd_mt.h``).

Reference: the V = 1 / ``_full`` variant of the reference's GPU engine (``engines/gpu/driver.py:435-520`` with
``shaders/common.metal:113-208, 563-601``), the physics of ``exchange.py`` and ``fields.py`` (restated in
``blochparty_amd/exchange.py`` and ``blochparty_amd/fields.py``); gradients through ``hb_adjoint_full``
(``engines/full_adjoint.py``). The case keeps the plain records (no runs, merged
runs, reused pulses or NUFFT runs), the sensitivities in rows (2 C, n_pad) and, as present, the bound pool's per-spin
parameters (f, k, R1b, T2b, dw), its Mzb with a low part, every segment's RF frequency and the super-Lorentzian's
ln F table, and the spins' concomitant coefficients (7, n_pad); the signal goes through ``hb_signal_reduce`` as the
plain path's.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from .. import exchange as X
from ..sequence import flag_array
from . import hip

LINE = {"lorentzian": 1, "superlorentzian": 2}
BOUND_ONLY = 262144  # HB_BOUND_ONLY of include/blochparty_amd_mt.h


class HbMt(ctypes.Structure):
    """``hb_mt`` of include/blochparty_amd_mt.h."""

    _fields_ = [("pool", ctypes.c_void_p), ("mb", ctypes.c_void_p), ("sens", ctypes.c_void_p),
                ("rf_freq", ctypes.c_void_p), ("table", ctypes.c_void_p), ("conc", ctypes.c_void_p),
                ("lx0", ctypes.c_float), ("dlx", ctypes.c_float), ("cutoff", ctypes.c_float),
                ("n_table", ctypes.c_int32), ("lineshape", ctypes.c_int32), ("conc_mode", ctypes.c_int32)]


def bind(lib) -> None:
    lib.hb_forward_mt.restype = ctypes.c_int
    lib.hb_forward_mt.argtypes = [ctypes.POINTER(hip.HbCase), ctypes.POINTER(HbMt), ctypes.c_int32, ctypes.c_int32,
                                  ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p]
    lib.hb_adjoint_full.restype = ctypes.c_int
    lib.hb_adjoint_full.argtypes = [ctypes.POINTER(hip.HbCase), ctypes.POINTER(HbMt), ctypes.c_void_p, ctypes.c_int32,
                                    ctypes.c_int32, ctypes.c_void_p]


def check_pool(bound: X.BoundPool) -> None:
    if bound.model != "rate":
        raise NotImplementedError("hip engine: the generalized Bloch bound pool is not implemented")
    if bound.cutoff is None and bound.lineshape == "superlorentzian":
        raise NotImplementedError("hip engine: a super-Lorentzian without its cut (cutoff=None) is not implemented")


def attach(case, phantom, seq, coils, m_init, n: int, n_pad: int, dev, conc=None) -> None:
    """The full variant's device buffers of ``case`` (``case.extra["mt"]``): the sensitivities, the bound pool's when
    the phantom has one, and ``conc`` = (coefficients (N, 6 or 7), mode) of the concomitant fields."""
    C = coils.shape[1]
    sens = torch.zeros(2 * C, n_pad, dtype=torch.float32)
    sens[0::2, :n] = coils.real.T.float().cpu()
    sens[1::2, :n] = coils.imag.T.float().cpu()
    out = dict(pool=None, mb=None, mb0=None, rf_freq=None, table=None, lx0=0.0, dlx=0.0, cutoff=0.0, lineshape=0,
               sens=sens.to(dev).contiguous(), conc=None, conc_mode=0)
    if conc is not None and conc[1]:
        coef = torch.zeros(7, n_pad, dtype=torch.float32)
        coef[6, n:] = 1.0  # padding spins: no transverse field and a nonzero f0 (the exact mode divides by |B| + u)
        coef[: conc[0].shape[1], :n] = conc[0].detach().float().cpu().T
        out.update(conc=coef.to(dev).contiguous(), conc_mode=int(conc[1]))
    if phantom.bound is not None:
        out.update(_pool_buffers(phantom, seq, m_init, n, n_pad, dev))
    case.extra["mt"] = out


def _pool_buffers(phantom, seq, m_init, n: int, n_pad: int, dev) -> dict:
    bound = phantom.bound
    check_pool(bound)
    like = phantom.pd.detach().double().cpu()
    b = bound.expand(n, like)
    pad = lambda v, fill: torch.cat([v.detach().double().cpu(), torch.full((n_pad - n,), fill, dtype=torch.float64)])
    pool = torch.stack([pad(b.f, 0.0), pad(b.k, 0.0), pad(1.0 / b.t1, 1.0), pad(b.t2, 1e-5), pad(b.dw, 0.0)])
    if m_init is not None and m_init.shape[1] == 4:
        mb64 = m_init[:, 3].detach().double().cpu()
    else:
        mb64 = b.f * like
    mb = torch.zeros(2, n_pad, dtype=torch.float32)
    mb[0, :n] = mb64.float()
    mb[1, :n] = (mb64 - mb64.float().double()).float()
    rf_freq = X.rf_frequency(seq).detach().float().cpu()
    lx0, dlx, logf = X.superlorentzian_table()
    return dict(pool=pool.float().to(dev).contiguous(), mb=mb.to(dev).contiguous(), mb0=mb.to(dev).contiguous(),
                rf_freq=rf_freq.to(dev).contiguous(), table=torch.from_numpy(logf.astype(np.float32)).to(dev),
                lx0=lx0, dlx=dlx, cutoff=float(bound.cutoff or 0.0), lineshape=LINE[bound.lineshape])


def struct(mt: dict) -> HbMt:
    ptr = lambda key: None if mt[key] is None else mt[key].data_ptr()
    n_table = 0 if mt["table"] is None else mt["table"].numel()
    return HbMt(ptr("pool"), ptr("mb"), ptr("sens"), ptr("rf_freq"), ptr("table"), ptr("conc"), mt["lx0"], mt["dlx"],
                mt["cutoff"], n_table, mt["lineshape"], mt["conc_mode"])


def forward_dispatch(case, st, s0: int, s1: int, rows: int, row0: int) -> None:
    lib = hip.load()
    bind(lib)
    s = struct(case.extra["mt"])
    rc = lib.hb_forward_mt(ctypes.byref(st), ctypes.byref(s), s0, s1, case.partial.data_ptr(), rows, row0,
                           ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    if rc != 0:
        raise RuntimeError(f"HIP error {rc} in hb_forward_mt: {lib.hb_error_string(rc).decode()}")


def needed(phantom, seq, conc=None) -> bool:
    """Whether a case takes the full variant: a bound pool, concomitant fields or bound-only RF."""
    bound_only = seq.bound_only is not None and any(seq.bound_only)
    return phantom.bound is not None or bound_only or (conc is not None and bool(conc[1]))


def mark_bound_only(rec, seq) -> None:
    """HB_BOUND_ONLY on the records of RF that only the bound pool feels (the water ignores it)."""
    if seq.bound_only is not None:
        rec["flags"][flag_array(seq.bound_only)] |= BOUND_ONLY


__all__ = ["HbMt", "attach", "bind", "check_pool", "forward_dispatch", "mark_bound_only", "needed", "struct"]
