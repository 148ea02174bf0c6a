"""Pulseq import (``blochparty_amd.pulseq``) on ``.seq`` files written by the test writer (tests/seqwriter.py), CPU:
shape decompression, the phase conventions, timing, relaxation and off-resonance against closed forms (after the
reference's tests/sequences/test_pulseq.py:54-176), gradient moments against the events' own integrals, a file GRE
against the same GRE from ``SequenceBuilder``, and the refusals. Simulations run on the FP64 oracle."""

import math

import numpy as np
import pytest
import torch

import blochparty_amd as bp
from blochparty_amd.pulseq import decompress_shape, read_pulseq
from oracle import core
from seqwriter import Writer, compress

f64 = torch.float64


def spins(z=None, n=1, t1=1e9, t2=1e9, db0=0.0, pd=1.0):
    full = lambda v: torch.full((n,), float(v), dtype=f64)
    pos = torch.zeros(n, 3, dtype=f64)
    if z is not None:
        pos[:, 2] = z
    return bp.Phantom(pos=pos, pd=full(pd), t1=full(t1), t2=full(t2), db0=full(db0))


def hard(flip_deg, dur, phase=0.0, freq=0.0, delay=0.0):
    """A block pulse as Pulseq writes one: a time-shaped pulse of two points, at its start and end (so it is one
    constant piece, cut at its centre)."""
    return {"samples": np.ones(2), "times": np.array([0.0, dur]), "amp": flip_deg / 360.0 / dur, "phase": phase,
            "freq": freq, "delay": delay}


def fid(phase=0.0, adc_phase=0.0, flip=90.0, adc_delay=0.0, n=8, dwell=100e-6, gap=1e-3, version=(1, 4, 1)):
    w = Writer(version)
    w.add(0.2e-6, rf=hard(flip, 0.2e-6, phase))
    w.add(gap)
    w.add(adc_delay + n * dwell, adc=(n, dwell, adc_delay, 0.0, adc_phase))
    return w.text()


def test_shape_decompression():
    # v, v, r is v repeated r + 2 times; the samples are the running sum of the derivative so coded
    assert np.allclose(decompress_shape(np.array([0.0, 0.1, 0.1, 3, 0.05]), 7), [0, 0.1, 0.2, 0.3, 0.4, 0.5, 0.55])
    assert np.array_equal(decompress_shape(np.array([1.0, 2.0, 3.0]), 3), [1.0, 2.0, 3.0])  # stored as is
    rng = np.random.default_rng(0)
    for w in (np.ones(40), np.linspace(0, 1, 33), np.concatenate([np.zeros(9), rng.normal(size=5), np.ones(7)])):
        packed = compress(w)
        np.testing.assert_allclose(decompress_shape(np.array(packed), len(w)), w, atol=1e-12)


def test_hard_pulse_phase_conventions():
    # physical: Pulseq phase phi -> Mxy = i exp(-i phi); equal RF and ADC phases cancel; mrzero: the phases as written
    for phi, psi in [(0.0, 0.0), (0.7, 0.0), (-math.pi / 2, 0.0), (0.0, 0.4), (1.3, 1.3)]:
        text = fid(phi, psi, n=1, dwell=1e-6, gap=1e-6)
        s = core.simulate(spins(), read_pulseq(text))[0][0, 0]
        assert abs(s - 1j * np.exp(-1j * (phi - psi))) < 1e-9, (phi, psi, s)
        s = core.simulate(spins(), read_pulseq(text, convention="mrzero"))[0][0, 0]
        assert abs(s - 1j * np.exp(1j * (phi - psi))) < 1e-9, (phi, psi, s)


def test_fid_timing_relaxation_and_off_resonance(tmp_path):
    t2, df, dwell, n, gap, delay = 0.05, 210.0, 100e-6, 16, 1e-3, 30e-6
    path = tmp_path / "fid.seq"
    path.write_text(fid(0.3, 0.1, adc_delay=delay, n=n, dwell=dwell, gap=gap))
    seq = read_pulseq(path)
    s = core.simulate(spins(t2=t2, db0=df), seq)[0][0].numpy()
    t = 0.1e-6 + gap + delay + (np.arange(n) + 0.5) * dwell  # from the pulse centre
    expected = 1j * np.exp(-1j * (0.3 - 0.1)) * np.exp(-t / t2) * np.exp(-2j * math.pi * df * t)
    assert np.abs(s - expected).max() < 1e-4
    assert np.abs(s - expected * np.exp(-1j * math.pi * df * dwell)).max() > 1e-2  # half a dwell off is caught
    t_end = torch.cumsum(seq.dt, 0)[torch.tensor(seq.adc)].numpy()
    np.testing.assert_allclose(t_end, 0.2e-6 + gap + delay + (np.arange(n) + 0.5) * dwell, atol=1e-12)
    assert seq.meta["adc_per_readout"] == [n]


def test_adc_frequency_offset_follows_off_resonant_spins():
    df, n, dwell = 400.0, 32, 50e-6
    for sign in (1.0, -1.0):
        w = Writer()
        w.add(20e-6, rf=hard(90.0, 20e-6))
        w.add(n * dwell, adc=(n, dwell, 0.0, sign * df, 0.0))
        s = core.simulate(spins(db0=df), read_pulseq(w.text()))[0][0].numpy()
        if sign > 0:  # demodulated at the spins' own offset: constant
            assert np.abs(s - s[0]).max() < 1e-9 and abs(s[0] - 1j) < 0.05
        else:  # the opposite offset winds the phase by 2 x 2 pi df per second
            assert abs(np.angle(s[1] / s[0]) + 4 * math.pi * df * dwell) < 1e-9


def test_ppm_offsets_and_adc_phase_modulation():
    """Format 1.5: ppm offsets weighted by gamma x B0 from the definitions; an ADC phase-modulation shape adds per
    sample phase."""
    defs = {"B0": 3.0, "Gamma": 42.576e6}
    f0 = 3.0 * 42.576e6 * 1e-6

    def read(rf_kw, adc_ppm, freq, phase, mod=None):
        w = Writer((1, 5, 0), defs=defs)
        w.add(100e-6, rf={**hard(30.0, 100e-6), **rf_kw})
        w.add(6 * 20e-6, adc=(6, 20e-6, 0.0, freq, phase, mod, adc_ppm))
        return read_pulseq(w.text())

    a = read({"ppm": (-3.4, 0.08), "centre": 50e-6}, (1.2, -0.3), 0.0, 0.0)
    b = read({"freq": -3.4 * f0, "phase": 0.08 * f0, "centre": 50e-6}, (0.0, 0.0), 1.2 * f0, -0.3 * f0)
    assert torch.allclose(a.rf, b.rf, atol=1e-9) and torch.allclose(a.adc_phase, b.adc_phase, atol=1e-12)
    assert a.rf[:, 1].abs().max() > 1.0
    mod = np.linspace(0, 1, 6)
    c = read({"centre": 50e-6}, (0.0, 0.0), 0.0, 0.0, mod)
    assert torch.allclose(c.adc_phase[torch.tensor(c.adc)], torch.as_tensor(-mod))


def sinc_slice(z0, convention="physical"):
    """A 30 deg windowed sinc (2 ms, TBW 4, 10 us raster samples written on a 0.1 us file raster as a time shape)
    with a slice-select trapezoid, freq offset G z0 and the phase offset that zeroes the phase at the centre, then
    the rephaser."""
    dur, tbw, thk, raster = 2e-3, 4.0, 5e-3, 10e-6
    k = round(dur / raster)
    t = (np.arange(k) + 0.5) * raster
    tc = t - dur / 2
    env = (0.5 + 0.5 * np.cos(2 * np.pi * tc / dur)) * np.sinc(tbw / dur * tc)
    amp_area = math.radians(30) / (2 * math.pi)
    amp = amp_area / (env.sum() * raster)
    g = tbw / dur / thk
    ramp = 100e-6
    freq = g * z0
    w = Writer(rasters={"RadiofrequencyRasterTime": 1e-7})
    rf = {"samples": env / env.max(), "amp": amp * env.max(), "times": t, "delay": ramp, "freq": freq,
          "phase": -2 * math.pi * freq * (dur / 2)}
    w.add(dur + 2 * ramp, rf=rf, gz=("trap", g, ramp, dur, ramp, 0.0))
    area = g * (dur / 2 + ramp / 2)
    w.add(1e-3, gz=("trap", -area / (0.9e-3), ramp, 0.8e-3, ramp, 0.0))
    z = torch.linspace(-0.03, 0.03, 241, dtype=f64)
    m = core.simulate(spins(z=z, n=len(z)), read_pulseq(w.text(), convention=convention))[1]
    return z, torch.complex(m[:, 0], m[:, 1])


def test_rf_frequency_offset_selects_the_slice_at_plus_f_over_g():
    z, centred = sinc_slice(0.0)
    _, shifted = sinc_slice(0.01)
    dz = float(z[1] - z[0])
    assert abs(float(z[shifted.abs().argmax()]) - 0.01) <= dz / 2
    i0, i1 = int(torch.argmin(z.abs())), int(torch.argmin((z - 0.01).abs()))
    span = 30
    assert (shifted[i1 - span : i1 + span + 1] - centred[i0 - span : i0 + span + 1]).abs().max() < 2e-3
    assert abs(centred[i0] - 1j * math.sin(math.radians(30))) < 2e-3  # rephased, along +y
    _, mirrored = sinc_slice(0.01, convention="mrzero")
    assert abs(float(z[mirrored.abs().argmax()]) + 0.01) <= dz / 2


def piecewise_integral(t, v, a, b):
    """Integral over [a, b] of the piecewise-linear waveform through (t, v), zero outside (fine trapezoid rule on
    the union of the corners and the ends: exact for piecewise-linear functions)."""
    pts = np.unique(np.clip(np.concatenate([t, [a, b]]), a, b))
    vals = np.interp(pts, t, v, left=0.0, right=0.0)
    return float(np.sum(0.5 * (vals[1:] + vals[:-1]) * np.diff(pts)))


def test_gradient_moments_at_the_samples_equal_the_events_integrals():
    """Trapezoids, an arbitrary gradient (raster samples with first / last) and an extended trapezoid (a time
    shape): the segment table's moment at every ADC sample equals the integral of the file's own waveforms."""
    w = Writer()
    w.add(40e-6, rf=hard(10.0, 40e-6))
    wave = 2e5 * np.sin(np.linspace(0, 3 * math.pi, 57)) ** 2
    w.add(30e-6 + 57 * 1e-5 + 20e-6, gx=("arb", 2e5, wave, 0.0, float(wave[-1]), 30e-6),
          adc=(40, 13e-6, 17e-6, 0.0, 0.0))
    times, amps = np.array([0, 1e-4, 3e-4, 3.5e-4]), np.array([0, 3e5, -1e5, 0])
    w.add(4e-4, gx=("trap", -1.5e5, 50e-6, 200e-6, 50e-6, 10e-6), gy=("ext", 3e5, times, amps, 0.0),
          adc=(25, 14e-6, 0.0, 0.0, 0.0))
    seq = read_pulseq(w.text())
    t_end = torch.cumsum(seq.dt, 0).numpy()
    k_end = torch.cumsum(seq.grad * seq.dt[:, None], 0).numpy()
    starts = np.cumsum([0.0, 40e-6, 30e-6 + 57e-5 + 20e-6])
    # the waveforms on the global clock
    tc = (np.arange(57) + 0.5) * 1e-5
    arb_t = starts[1] + 30e-6 + np.concatenate([[0.0], tc, [57e-5]])
    arb_v = np.concatenate([[0.0], wave, [wave[-1]]])
    trap_t = starts[2] + 10e-6 + np.cumsum([0.0, 50e-6, 200e-6, 50e-6])
    trap_v = -1.5e5 * np.array([0.0, 1.0, 1.0, 0.0])
    ext_t, ext_v = starts[2] + times, amps
    adc = np.flatnonzero(seq.adc)
    for i in adc:
        b = t_end[i]
        kx = piecewise_integral(arb_t, arb_v, 0.0, b) + piecewise_integral(trap_t, trap_v, 0.0, b)
        ky = piecewise_integral(ext_t, ext_v, 0.0, b)
        assert abs(k_end[i, 0] - kx) < 1e-9 * 1e3 and abs(k_end[i, 1] - ky) < 1e-9 * 1e3, (i, k_end[i], kx, ky)
    assert seq.meta["adc_per_readout"] == [40, 25]


def test_file_gre_equals_builder_gre():
    """An ideal GRE (block gradients written as constant arbitrary gradients) read from a file against the same GRE
    from SequenceBuilder (Pulseq phase phi -> -phi): the same oracle signal."""
    n, fov, dwell, t_pre, tr, rf_dur, flip = 8, 0.2, 10e-6, 400e-6, 3e-3, 20e-6, 30.0
    dk = 1 / fov
    w = Writer()
    b = bp.SequenceBuilder()
    const = lambda amp, dur: ("arb", amp if amp else 1.0, np.full(round(dur / 1e-5), amp), amp, amp, 0.0)
    t_spoil, t_fill = 500e-6, tr - rf_dur - t_pre - n * dwell - 500e-6
    for i in range(n):
        phase = 0.5 * 2 * math.pi * i * (i + 1) / 2 / 3
        gpre = (-(n / 2 + 0.5) * dk / t_pre, (i - n // 2) * dk / t_pre)
        w.add(rf_dur, rf=hard(flip, rf_dur, phase))
        w.add(t_pre, gx=const(gpre[0], t_pre), gy=const(gpre[1], t_pre))
        w.add(n * dwell, gx=const(dk / dwell, n * dwell), adc=(n, dwell, 0.0, 0.0, phase))
        w.add(t_spoil, gx=const(3 * n * dk / t_spoil, t_spoil))
        w.add(t_fill)
        b.rf(rf_dur, flip_deg=flip, phase=-phase, n_steps=2)  # the RF centre is a knot of the file's pulse
        b.delay(t_pre, grad=(*gpre, 0.0))
        b.readout(n, dwell, grad=(dk / dwell, 0.0, 0.0), phase=-phase)
        b.delay(t_spoil, grad=(3 * n * dk / t_spoil, 0.0, 0.0))
        b.delay(t_fill)
    phantom = bp.shepp_logan(n, fov)
    phantom.t2 = phantom.t2 * 0 + 0.02
    s_file = core.simulate(phantom, read_pulseq(w.text()))[0]
    s_b = core.simulate(phantom, b.build())[0]
    assert (s_file - s_b).abs().max() < 1e-10 * s_b.abs().max()


def test_extensions_and_format_refusals():
    w = Writer()
    w.add(0.2e-6, rf=hard(90.0, 0.2e-6))
    w.add(1e-4, ext=["LABELSET"])  # labels are ignored
    w.add(8 * 1e-5, adc=(8, 1e-5, 0.0, 0.0, 0.0), ext=["TRIGGERS", "LABELINC"])
    assert read_pulseq(w.text()).n_adc == 8
    w.add(1e-4, ext=["ROTATIONS"])
    with pytest.raises(NotImplementedError, match="ROTATIONS"):
        read_pulseq(w.text())
    old = fid().replace("minor 4", "minor 3")
    with pytest.raises(NotImplementedError, match="1.3"):
        read_pulseq(old)
    with pytest.raises(ValueError, match="convention"):
        read_pulseq(fid(), convention="mri")
