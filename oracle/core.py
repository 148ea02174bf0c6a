"""FP64 CPU oracle of the GPU path: a restatement of the reference's PyTorch core engine, plain Bloch physics only.

TEST INFRASTRUCTURE. Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
or run anything under ``oracle/``, and only as the checker (or the timed CPU baseline), never as the product: the
package ``blochparty_amd`` never imports it, and its GPU engine fails loudly when the HIP library is missing.

What it computes, following ``src/blochparty/engines/core.py`` (file:line of the reference):

- ``relax``            core.py:71-74     exact T1/T2 relaxation over t
- ``rotate``           core.py:97-108    right-handed rotation by the rotation vector w (Rodrigues, sinc forms)
- ``rf_segment``       core.py:124-150   Strang: relax dt/2, rotate about w = -2 pi dt (b1 rf_x, b1 rf_y, bz), relax
- ``sample``           core.py:182-187   (2, C) coil sums of Mxy demodulated by exp(-i psi)
- ``free_run``         core.py:268-319   closed form of a run of RF-free segments, every ADC sample in it blockwise
- ``steps``            core.py:322-334   single RF segments, and free runs ending at a spoiled segment
- ``simulate``         core.py:487-608   m0 (equilibrium or m_init), the run, spoiling, the (C, n_adc) signal

Plain path only: no bound / free pools, diffusion, motion, field model or field maps (the GPU engine refuses them).
Autograd flows through everything, so gradients of this oracle are the reference gradients of the adjoint tests.

How each function is pinned (tests/test_oracle.py):
- ``rotate``: 1e-12 against SciPy's rotation-vector matrix (a textbook routine), and tip angle / phase conventions
  against the closed forms of the reference's own analytic tests (tests/physics/test_analytic.py:16-28);
- ``rf_segment`` with relaxation: a brute-force SciPy DOP853 integration of the Bloch equations on tiny inputs,
  second-order convergence of the Strang error under step halving (ratio 3.5-5);
- ``relax`` / ``free_run`` / ``sample``: the FID, spin-echo, inversion-recovery and spoiled-GRE closed forms of
  tests/physics/test_analytic.py:31-82 and reference.py, merged free runs = segment-by-segment (an invariant), and
  the 3D Cartesian signal equation (a direct DFT) for the k-space encoding;
- ``simulate`` end to end: golden vectors written by the reference's own ``core.simulate`` on reduced cases
  (tests/golden/), which also pin the composition (ADC timing in composite blocks, spoil order, receiver phase).
"""

from __future__ import annotations

import math

import torch
from torch import Tensor

TWO_PI = 2.0 * math.pi
MAX_BLOCK = 1 << 22  # spins x samples per block of a free run's phase matrix


def relax(mx, my, mz, t, pd, r1, r2):
    """Exact relaxation over ``t``: Mxy decays with R2, Mz recovers to PD with R1 (core.py:71-74)."""
    e2 = torch.exp(-t * r2)
    e1 = torch.exp(-t * r1)
    return mx * e2, my * e2, pd + (mz - pd) * e1


def rotate(mx, my, mz, wx, wy, wz):
    """Rotation by the vector w (angle |w|, right-handed), Rodrigues with sin(phi)/phi and (1 - cos phi)/phi^2 as
    sinc forms; 1e-30 under the root keeps the gradient finite at w = 0 (core.py:97-108)."""
    phi = torch.sqrt(wx * wx + wy * wy + wz * wz + 1e-30)
    c = torch.cos(phi)
    s1 = torch.sinc(phi / math.pi)
    s2 = 0.5 * torch.sinc(phi / TWO_PI) ** 2
    dot = wx * mx + wy * my + wz * mz
    return (c * mx + s1 * (wy * mz - wz * my) + s2 * wx * dot,
            c * my + s1 * (wz * mx - wx * mz) + s2 * wy * dot,
            c * mz + s1 * (wx * my - wy * mx) + s2 * wz * dot)


class Spins:
    """The per-spin inputs of the plain path, float64, vectorised over N."""

    def __init__(self, pos, db0, pd, t1, t2, sens_re, sens_im, b1, field=None):
        self.pos, self.db0, self.pd = pos, db0, pd
        self.r1, self.r2 = 1.0 / t1, 1.0 / t2
        self.sens_re, self.sens_im, self.b1 = sens_re, sens_im, b1
        self.field = field  # oracle.fields.Field: the concomitant field (core.py:115-121), or None


def rf_segment(mx, my, mz, dt, rf, g, off, sp: Spins):
    """One RF segment by Strang splitting (core.py:124-150); the concomitant field adds to the offset (:128-129)."""
    bz = sp.pos @ g + sp.db0
    if off is not None:
        bz = bz + off
    if sp.field is not None:
        bz = bz + sp.field.at(g, bz)
    mx, my, mz = relax(mx, my, mz, 0.5 * dt, sp.pd, sp.r1, sp.r2)
    w = -TWO_PI * dt
    wx, wy = w * rf[0], w * rf[1]
    if sp.b1 is not None:
        wx, wy = wx * sp.b1, wy * sp.b1
    mx, my, mz = rotate(mx, my, mz, wx, wy, w * bz)
    return relax(mx, my, mz, 0.5 * dt, sp.pd, sp.r1, sp.r2)


def sample(mx, my, psi, sp: Spins):
    """(2, C): coil sums of Mxy = mx + i my, demodulated by exp(-i psi) (core.py:182-187)."""
    re = mx @ sp.sens_re - my @ sp.sens_im
    im = mx @ sp.sens_im + my @ sp.sens_re
    c, s = torch.cos(psi), torch.sin(psi)
    return torch.stack([re * c + im * s, im * c - re * s])


def free_run(mx, my, mz, dt, g, off, psi, adc_idx, sp: Spins):
    """RF-free segments in closed form (core.py:268-319): with t, k the cumulative time and gradient moment,
    Mxy(t) = Mxy(0) exp(-R2 t) exp(-i 2 pi (k.r + dB0 t + int off)); every ADC sample is one exponential sum over
    spins, evaluated blockwise as matrix products."""
    t = torch.cumsum(dt, 0)
    k = torch.cumsum(g * dt[:, None], 0)
    extra = None if off is None else torch.cumsum(off * dt, 0)
    if sp.field is not None:  # per spin and segment (core.py:190-217): the field of each segment times its length
        lin = sp.pos @ g.T + sp.db0[:, None] + (0.0 if off is None else off[None])
        conc = torch.cumsum(sp.field.at(g, lin) * dt[None], 1)  # (N, n)
        extra = conc if extra is None else conc + extra[None]
    sig = None
    if adc_idx:
        idx = torch.tensor(adc_idx, device=dt.device)
        ta, ka, pa = t[idx], k[idx], psi[idx]
        v_re = sp.sens_re * mx[:, None] - sp.sens_im * my[:, None]
        v_im = sp.sens_re * my[:, None] + sp.sens_im * mx[:, None]
        block = max(1, MAX_BLOCK // max(mx.shape[0], 1))
        parts = []
        for j in range(0, len(adc_idx), block):
            tj, kj = ta[j : j + block], ka[j : j + block]
            cyc = sp.pos @ kj.T + sp.db0[:, None] * tj
            if extra is not None:
                cyc = cyc + extra[..., idx[j : j + block]]
            phase = TWO_PI * cyc
            decay = torch.exp(-sp.r2[:, None] * tj)
            e_re, e_im = decay * torch.cos(phase), -decay * torch.sin(phase)
            parts.append(torch.stack([v_re.T @ e_re - v_im.T @ e_im, v_re.T @ e_im + v_im.T @ e_re]))
        s_re, s_im = torch.cat(parts, dim=-1)
        c, s = torch.cos(pa), torch.sin(pa)
        sig = torch.stack([s_re * c + s_im * s, s_im * c - s_re * s]).permute(2, 0, 1)  # (n, 2, C)
    cyc = sp.pos @ k[-1] + sp.db0 * t[-1]
    if extra is not None:
        cyc = cyc + extra[..., -1]
    phi = -TWO_PI * cyc
    c, s = torch.cos(phi), torch.sin(phi)
    mx, my = c * mx - s * my, s * mx + c * my
    mx, my, mz = relax(mx, my, mz, t[-1], sp.pd, sp.r1, sp.r2)
    return mx, my, mz, sig


def steps(is_rf, spoil, merge: bool) -> list[tuple[int, int]]:
    """[start, end) steps: each RF segment alone; runs of RF-free segments, a run ending at a spoiled segment
    (core.py:322-334). ``merge`` False: one segment per step."""
    out, s, n = [], 0, len(is_rf)
    while s < n:
        e = s + 1
        if merge and not is_rf[s]:
            while e < n and not is_rf[e] and not spoil[e - 1]:
                e += 1
        out.append((s, e))
        s = e
    return out


def simulate(phantom, seq, coils: Tensor | None = None, *, m_init: Tensor | None = None,
             merge_free_runs: bool = True, dtype=torch.float64, field=None, enc_pos=None):
    """The oracle's ``simulate`` (core.py:487-608, plain path): returns (signal (C, n_adc) complex, m (N, 3)).
    ``field``: an ``oracle.fields.Field`` (the concomitant field of a scanner); ``enc_pos``: encoding positions in
    place of the phantom's (a gradient nonlinearity)."""
    pd = phantom.pd.to(dtype)
    dev = pd.device
    conv = lambda t: None if t is None else t.to(dtype=dtype, device=dev)
    dt_all, rf_all, g_all, psi_all = conv(seq.dt), conv(seq.rf), conv(seq.grad), conv(seq.adc_phase)
    off_all = conv(seq.b0_offset)
    use_off = off_all is not None and (off_all.requires_grad or bool(off_all.any()))
    n = pd.shape[0]
    if coils is None:
        sens_re, sens_im = torch.ones_like(pd)[:, None], torch.zeros_like(pd)[:, None]
    else:
        cplx = coils.to(torch.complex128 if dtype == torch.float64 else torch.complex64)
        sens_re, sens_im = cplx.real, cplx.imag
    pos = conv(phantom.pos) if enc_pos is None else conv(enc_pos)
    sp = Spins(pos, conv(phantom.db0), pd, conv(phantom.t1), conv(phantom.t2), sens_re, sens_im, conv(phantom.b1),
               field)
    if m_init is None:
        mx, my, mz = torch.zeros_like(pd), torch.zeros_like(pd), pd
    else:
        mx, my, mz = conv(m_init)[:, :3].unbind(-1)
    samples = []
    for a, b in steps(seq.is_rf, seq.spoil, merge_free_runs):
        if seq.is_rf[a]:
            off = off_all[a] if use_off else None
            mx, my, mz = rf_segment(mx, my, mz, dt_all[a], rf_all[a], g_all[a], off, sp)
            if seq.adc[a]:
                samples.append(sample(mx, my, psi_all[a], sp)[None])
        else:
            adc_idx = [i for i in range(b - a) if seq.adc[a + i]]
            off = off_all[a:b] if use_off else None
            mx, my, mz, sig = free_run(mx, my, mz, dt_all[a:b], g_all[a:b], off, psi_all[a:b], adc_idx, sp)
            if sig is not None:
                samples.append(sig)
        if seq.spoil[b - 1]:
            mx, my = torch.zeros_like(mx), torch.zeros_like(my)
    sig = torch.cat(samples) if samples else pd.new_zeros(0, 2, sens_re.shape[1])
    assert sig.shape[0] == seq.n_adc and mx.shape[0] == n
    return torch.complex(sig[:, 0], sig[:, 1]).T, torch.stack([mx, my, mz], dim=-1)
