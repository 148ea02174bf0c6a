"""Writes the golden fixtures of tests/golden/ from the reference package (run in the build container only; the
reference is not needed, and not present, where the tests run).

    python tests/golden/make_golden.py /root/reference/src

What the fixtures pin (each is inputs + the reference's outputs, small):
- ``sequences.npz``: segment tables (dt, rf, grad, adc_phase, flags) of the reference's builders on reduced cases:
  ideal and realistic GRE (ramps, RF spoiling, gradient spoiling, sinc slab on a 10 us raster) and TSE (hard and
  sinc pulses, orderings, a variable refocusing train). The splitting rules of ``SequenceBuilder.block`` have no
  closed form to check against.
- ``phantoms.npz``: Shepp-Logan maps (2D, 3D), sub-voxel / stratified positions, birdcage sensitivities.
- ``core.npz``: the reference core's float64 ``simulate`` on reduced cases, inputs stored alongside, and the
  gradients of a fixed weighted loss through its autograd: these pin the oracle's composition (ADC timing in
  composite blocks, spoil order, receiver phase, gradient semantics) end to end.
"""

from __future__ import annotations

import math
import sys
from pathlib import Path

import numpy as np
import torch

OUT = Path(__file__).resolve().parent


def seq_arrays(prefix: str, seq) -> dict:
    flags = (np.asarray(seq.is_rf, np.int32) | (np.asarray(seq.adc, np.int32) << 1)
             | (np.asarray(seq.spoil, np.int32) << 2))
    return {
        f"{prefix}/dt": seq.dt.detach().double().numpy(),
        f"{prefix}/rf": seq.rf.detach().double().numpy(),
        f"{prefix}/grad": seq.grad.detach().double().numpy(),
        f"{prefix}/adc_phase": seq.adc_phase.detach().double().numpy(),
        f"{prefix}/flags": flags,
    }


def main(ref_src: str) -> None:
    sys.path.insert(0, ref_src)
    import blochparty as rb
    from blochparty.engines import core as rcore

    torch.manual_seed(0)
    # ---- sequences
    sys10 = rb.System(rf_raster=10e-6)
    sys20 = rb.System(rf_raster=20e-6)
    seqs = {
        "gre2d_16_ideal": rb.gre_2d(16, n_dummy=3, te=5e-3),
        "gre2d_12_rfspoil_ramps": rb.gre_2d(12, n_dummy=4, spoiling="rf", ramps=True),
        "gre2d_8_gradspoil_slice": rb.gre_2d(8, spoiling="gradient", slice_thickness=5e-3, n_dummy=1),
        "gre3d_6_sinc_slab": rb.gre_cartesian((6, 6, 6), flip_deg=10.0, tr=8e-3, slice_thickness=0.22,
                                              rf_duration=2e-3, system=sys10),
        "tse_32_etl4_hard": rb.tse_2d(32, etl=4, tr=0.5),
        "tse_16_etl4_sinc": rb.tse_2d(16, etl=4, tr=0.5, slice_thickness=5e-3, rf_duration=2e-3, system=sys20),
        "tse_16_centric_train": rb.tse_2d(16, etl=4, tr=0.4, ordering="centric",
                                          refocus_flip_deg=[150.0, 160.0, 170.0, 180.0]),
        "tse_16_reverse_gradspoil": rb.tse_2d(16, etl=8, tr=0.4, ordering="reverse_centric", spoiling="gradient"),
    }
    out = {}
    for k, s in seqs.items():
        out.update(seq_arrays(k, s))
    np.savez_compressed(OUT / "sequences.npz", **out)

    # ---- phantoms and coils
    ph = {}
    for n, nd in ((32, 2), (12, 3)):
        m = rb.shepp_logan_maps(n, ndim=nd)
        for key in ("label", "pd", "t1", "t2"):
            ph[f"sl{nd}d_{n}/{key}"] = m[key].numpy()
    m = rb.shepp_logan_maps(8, ndim=3)
    sub = rb.phantom_from_maps(m["pd"], m["t1"], m["t2"], 0.22, spins_per_voxel=(1, 2, 2))
    jit = rb.phantom_from_maps(m["pd"], m["t1"], m["t2"], 0.22, spins_per_voxel=(2, 2, 2), jitter="stratified",
                               seed=5)
    ph["sub122/pos"], ph["sub122/pd"] = sub.pos.numpy(), sub.pd.numpy()
    ph["jit222/pos"], ph["jit222/pd"] = jit.pos.numpy(), jit.pd.numpy()
    c = rb.birdcage_sensitivities(sub.pos, 8)
    ph["coils8/re"], ph["coils8/im"] = c.real.numpy(), c.imag.numpy()
    c = rb.birdcage_sensitivities(sub.pos, 5, radius=0.15, radius_y=0.2, element_radius=0.04, normalize=True)
    ph["coils5e/re"], ph["coils5e/im"] = c.real.numpy(), c.imag.numpy()
    np.savez_compressed(OUT / "phantoms.npz", **ph)

    # ---- core.simulate on reduced cases (inputs stored with the outputs)
    core = {}

    def run(name, phantom, seq, coils=None, m_init=None):
        r = rcore.simulate(phantom, seq, coils, m_init=m_init)
        core.update(seq_arrays(name, seq))
        core[f"{name}/pos"], core[f"{name}/pd"] = phantom.pos.numpy(), phantom.pd.numpy()
        core[f"{name}/t1"], core[f"{name}/t2"], core[f"{name}/db0"] = (phantom.t1.numpy(), phantom.t2.numpy(),
                                                                       phantom.db0.numpy())
        core[f"{name}/b1"] = np.ones(len(phantom)) if phantom.b1 is None else phantom.b1.numpy()
        if coils is not None:
            core[f"{name}/coils_re"], core[f"{name}/coils_im"] = coils.real.numpy(), coils.imag.numpy()
        if m_init is not None:
            core[f"{name}/m_init"] = m_init.numpy()
        if seq.b0_offset is not None:
            core[f"{name}/b0_offset"] = seq.b0_offset.numpy()
        core[f"{name}/signal_re"], core[f"{name}/signal_im"] = r.signal.real.numpy(), r.signal.imag.numpy()
        core[f"{name}/m"] = r.m.numpy()

    p16 = rb.shepp_logan(16)
    run("gre16", p16, rb.gre_2d(16, n_dummy=20, te=5e-3), rb.birdcage_sensitivities(p16.pos, 8))
    g = torch.Generator().manual_seed(7)
    p = rb.Phantom(p16.pos, p16.pd, p16.t1, p16.t2, (torch.rand(len(p16), generator=g, dtype=torch.float64) - 0.5)
                   * 80, b1=0.8 + 0.4 * torch.rand(len(p16), generator=g, dtype=torch.float64))
    s = rb.gre_2d(16, n_dummy=4, spoiling="rf", ramps=True)
    s.b0_offset = 2.0 * torch.cos(torch.arange(len(s), dtype=torch.float64) * 3e-3)
    run("gre16_b1_b0off_rfspoil", p, s, rb.birdcage_sensitivities(p.pos, 4),
        m_init=torch.randn(len(p), 3, generator=g, dtype=torch.float64) * 0.3)
    m = rb.shepp_logan_maps(4, ndim=3)
    p3 = rb.phantom_from_maps(m["pd"], m["t1"], m["t2"], 0.22, spins_per_voxel=(1, 2, 2))
    p3 = rb.Phantom(p3.pos, p3.pd, p3.t1, p3.t2, 60.0 * torch.sin(40.0 * p3.pos[:, 0] + 20.0 * p3.pos[:, 2]))
    run("sinc3d4", p3, rb.gre_cartesian((4, 4, 4), flip_deg=10.0, tr=8e-3, slice_thickness=0.22, rf_duration=2e-3,
                                        system=sys10), rb.birdcage_sensitivities(p3.pos, 8))
    run("tse16", p16, rb.tse_2d(16, etl=4, tr=0.3, slice_thickness=5e-3, rf_duration=1e-3, system=sys20),
        rb.birdcage_sensitivities(p16.pos, 8))

    # gradients of a fixed weighted loss (reference tests/helpers.py:77-92 style) on a tiny case
    pg = rb.shepp_logan(8)
    n = len(pg)
    leaves = {
        "pd": pg.pd.clone().requires_grad_(), "t1": pg.t1.clone().requires_grad_(),
        "t2": pg.t2.clone().requires_grad_(),
        "db0": (20.0 * torch.sin(30 * pg.pos[:, 0])).requires_grad_(),
        "b1": torch.linspace(0.9, 1.1, n, dtype=torch.float64).requires_grad_(),
    }
    flip = torch.tensor(25.0, dtype=torch.float64, requires_grad=True)
    coils = rb.birdcage_sensitivities(pg.pos, 4).clone().requires_grad_()
    sq = rb.gre_2d(8, n_dummy=3, flip_deg=flip, te=4e-3, tr=8e-3)
    sq.b0_offset = torch.full((len(sq),), 1.5, dtype=torch.float64, requires_grad=True)
    phg = rb.Phantom(pg.pos, leaves["pd"], leaves["t1"], leaves["t2"], leaves["db0"], b1=leaves["b1"])
    r = rcore.simulate(phg, sq, coils)
    gw = torch.Generator().manual_seed(1)
    wr = torch.randn(r.signal.shape, generator=gw, dtype=torch.float64)
    wi = torch.randn(r.signal.shape, generator=gw, dtype=torch.float64)
    wm = torch.randn(r.m.shape, generator=gw, dtype=torch.float64)
    loss = (r.signal.real * wr).sum() + (r.signal.imag * wi).sum() + (r.m * wm).sum()
    grads = torch.autograd.grad(loss, [*leaves.values(), flip, coils, sq.b0_offset])
    names = [*leaves, "flip", "coils", "b0_offset"]
    core["grad/pos"] = pg.pos.numpy()
    for k in leaves:
        core[f"grad/in_{k}"] = leaves[k].detach().numpy()
    core["grad/coils_re"], core["grad/coils_im"] = coils.detach().real.numpy(), coils.detach().imag.numpy()
    core["grad/w_re"], core["grad/w_im"], core["grad/w_m"] = wr.numpy(), wi.numpy(), wm.numpy()
    core["grad/loss"] = np.array(float(loss))
    for k, gr in zip(names, grads):
        if gr.is_complex():
            core[f"grad/d_{k}_re"], core[f"grad/d_{k}_im"] = gr.real.numpy(), gr.imag.numpy()
        else:
            core[f"grad/d_{k}"] = gr.numpy()
    np.savez_compressed(OUT / "core.npz", **core)
    for f in ("sequences.npz", "phantoms.npz", "core.npz"):
        print(f, (OUT / f).stat().st_size, "bytes")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src")
