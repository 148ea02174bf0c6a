"""The bound-pool oracle (oracle/exchange.py) pinned by independent checks, and the package's lineshape table against
it (CPU): SciPy's matrix exponential, the steady state, the absorption lineshapes' normalisation, a DOP853
integration of the coupled equations, and the plain path at f = 0."""

import math

import numpy as np
import pytest
import torch
from scipy.integrate import quad, solve_ivp
from scipy.linalg import expm

import blochparty_amd as bp
from blochparty_amd import exchange as X
from oracle import core
from oracle import exchange as E

f64 = torch.float64


def test_z_exchange_is_the_matrix_exponential():
    rng = np.random.default_rng(0)
    for _ in range(20):
        pd, r1, f, k, r1b, w = rng.uniform(0.5, 1), rng.uniform(0.3, 3), rng.uniform(0.05, 0.3), \
            rng.uniform(2, 60), rng.uniform(0.5, 2), rng.uniform(0, 50)
        t = rng.uniform(1e-6, 0.3)
        mz, mb = rng.uniform(-1, 1), rng.uniform(-0.3, 0.3)
        A = np.array([[-(r1 + k * f), k, r1 * pd], [k * f, -(r1b + k + w), r1b * f * pd], [0, 0, 0]])
        want = expm(A * t) @ np.array([mz, mb, 1.0])
        T = lambda v: torch.tensor(v, dtype=f64)
        got = E.z_exchange(T(mz), T(mb), T(t), T(pd), T(r1), T(f), T(k), T(r1b), T(w))
        np.testing.assert_allclose([float(got[0]), float(got[1])], want[:2], rtol=0, atol=1e-12)
    # long times: the steady state -A^-1 c (no saturation: both pools at equilibrium)
    T = lambda v: torch.tensor(v, dtype=f64)
    zs = E.z_exchange(T(-0.7), T(0.0), T(50.0), T(0.9), T(1.0), T(0.15), T(20.0), T(1.0), None)
    assert abs(float(zs[0]) - 0.9) < 1e-12 and abs(float(zs[1]) - 0.15 * 0.9) < 1e-12


def test_lineshapes_are_normalised_absorption_lines():
    """int G(w) dw = pi over all w (G = pi g, g of unit area): int_0^inf F(x) dx = pi / 2 for the super-Lorentzian,
    and the Lorentzian T2b / (1 + x^2) likewise."""
    # F on a geometric grid (one vectorised quadrature), integrated as int F x d(ln x); below 1e-9 adds < 1e-7
    lx = np.linspace(math.log(1e-9), math.log(12.0), 3001)
    x = np.exp(lx)
    half = float(np.trapezoid(E.sl_F(x) * x, lx))
    assert abs(half - math.pi / 2) < 1e-5, half
    lor = quad(lambda x: 1 / (1 + x * x), 0, np.inf)[0]
    assert abs(lor - math.pi / 2) < 1e-12


def test_package_lineshape_table_against_the_oracle():
    xs = np.geomspace(1e-3, 25.0, 37)
    # (quad_vec controls the error of the whole vector: its entries of ~1e-30 and below are loose)
    np.testing.assert_allclose(X.superlorentzian_integral(xs), E.sl_F(xs), rtol=1e-9, atol=1e-25)
    pool = X.BoundPool(f=0.1, k=20.0, t1=1.0, t2=12e-6)
    delta = torch.tensor([0.0, 300.0, 1200.0, -5e3, 2e4, 1e5], dtype=f64)
    got = X.lineshape(pool, delta)
    want = E.lineshape_G("superlorentzian", delta, torch.tensor(12e-6, dtype=f64), 1e3)
    # the table interpolates ln F linearly in ln x: ~h^2 x^2 / 2 relative, 1e-4 where ln F falls steeply (x ~ 7)
    np.testing.assert_allclose(got.numpy(), want.numpy(), rtol=2e-4)
    at_cut = float(X.lineshape(pool, torch.tensor([1e3], dtype=f64))[0])
    assert float(got[0]) == float(got[1]) == at_cut  # held at the cut inside it


def test_rf_segment_bound_against_the_coupled_ode():
    """One 1 ms off-resonant RF pulse with saturation on two spins: the Strang steps against DOP853 on
    (mx, my, mz, mzb); the error falls ~4x per halving of the step."""
    pd, t1, t2, db0 = 1.0, 0.8, 0.05, 40.0
    f, k, t1b, t2b, frf = 0.12, 25.0, 1.0, 12e-6, 1500.0
    amp, T = 400.0, 1e-3
    r1, r2, r1b = 1 / t1, 1 / t2, 1 / t1b
    pool = X.BoundPool(f=f, k=k, t1=t1b, t2=t2b, lineshape="lorentzian")
    G = float(E.lineshape_G("lorentzian", torch.tensor(frf - db0, dtype=f64), torch.tensor(t2b, dtype=f64), None))
    W = (2 * math.pi * amp) ** 2 * G

    def rhs(_, m):
        mx, my, mz, mb = m
        wx, wz = -2 * math.pi * amp, -2 * math.pi * db0  # rotation vector rate (right-handed about w)
        rot = np.cross([wx, 0.0, wz], [mx, my, mz])
        return [rot[0] - r2 * mx, rot[1] - r2 * my, rot[2] + r1 * (pd - mz) - k * f * mz + k * mb,
                r1b * (f * pd - mb) + k * f * mz - k * mb - W * mb]

    ref = solve_ivp(rhs, (0, T), [0, 0, pd, f * pd], method="DOP853", rtol=1e-12, atol=1e-13).y[:, -1]
    sp = core.Spins(torch.zeros(1, 3, dtype=f64), torch.full((1,), db0, dtype=f64), torch.full((1,), pd, dtype=f64),
                    torch.full((1,), t1, dtype=f64), torch.full((1,), t2, dtype=f64), None, None, None)
    P = E.Pool(pool, 1, f64)
    errs = []
    for n in (20, 40, 80):
        m = [torch.zeros(1, dtype=f64), torch.zeros(1, dtype=f64), torch.full((1,), pd, dtype=f64),
             torch.full((1,), f * pd, dtype=f64)]
        dt = torch.tensor(T / n, dtype=f64)
        rf = torch.tensor([amp, 0.0], dtype=f64)
        for _ in range(n):
            m = list(E.rf_segment_bound(*m, dt, rf, torch.zeros(3, dtype=f64), None, torch.tensor(frf, dtype=f64),
                                        sp, P))
        errs.append(max(abs(float(a[0]) - b) for a, b in zip(m, ref)))
    assert errs[-1] < 1e-6, errs
    assert 3.0 < errs[0] / errs[1] < 5.0 and 3.0 < errs[1] / errs[2] < 5.0, errs


def test_no_pool_is_the_plain_path():
    ph = bp.shepp_logan(8)
    seq = bp.gre_2d(8, n_dummy=3, spoiling="rf", ramps=True, slice_thickness=5e-3)
    coils = bp.birdcage_sensitivities(ph.pos, 2)
    ph0 = bp.Phantom(ph.pos, ph.pd, ph.t1, ph.t2, ph.db0,
                     bound=X.BoundPool(f=0.0, k=0.0, t1=1.0, t2=12e-6, lineshape="lorentzian"))
    s, m = E.simulate_bound(ph0, seq, coils)
    s_ref, m_ref = core.simulate(ph, seq, coils)
    assert (s - s_ref).abs().max() < 1e-12 * s_ref.abs().max() and (m[:, :3] - m_ref).abs().max() < 1e-12


def test_rf_frequency_closed_form_and_gradient():
    """A pulse turning as A exp(-i 2 pi f t) on a dt raster has frequency f at every RF segment (neighbours on one
    side at its ends), 0 where no neighbour carries RF; the oracle's rule (atan2) equals the package's (the angle of
    the complex product); its autograd gradient in rf and dt equals central differences."""
    f, dt, k = 730.0, 10e-6, 12
    t = torch.arange(k, dtype=f64) * dt
    amp = 200.0 * (1.5 + torch.sin(torch.arange(k, dtype=f64)))
    rf = torch.zeros(k + 3, 2, dtype=f64)
    rf[:k, 0], rf[:k, 1] = amp * torch.cos(2 * math.pi * f * t), -amp * torch.sin(2 * math.pi * f * t)
    rf[k + 1] = torch.tensor([50.0, 10.0], dtype=f64)  # an isolated RF segment
    is_rf = [True] * k + [False, True, False]
    dts = torch.full((k + 3,), dt, dtype=f64)
    got = E._rf_frequency(is_rf, rf, dts)
    assert torch.allclose(got[:k], torch.full((k,), f, dtype=f64), rtol=1e-12)
    assert float(got[k:].abs().max()) == 0.0
    seq = bp.Sequence(dts, rf, torch.zeros(k + 3, 3, dtype=f64), torch.zeros(k + 3, dtype=f64), is_rf,
                      [False] * (k + 3), [False] * (k + 3))
    assert torch.allclose(X.rf_frequency(seq), got, rtol=1e-12, atol=1e-9)
    # gradients: L = w . f_rf against central differences in rf and dt
    w = torch.linspace(-1.0, 2.0, k + 3, dtype=f64)
    rf_l, dt_l = rf.clone().requires_grad_(), dts.clone().requires_grad_()
    g_rf, g_dt = torch.autograd.grad((w * E._rf_frequency(is_rf, rf_l, dt_l)).sum(), [rf_l, dt_l])
    for idx in [(3, 0), (3, 1), (k - 1, 1), (k + 1, 0)]:
        h = 1e-3
        p, m = rf.clone(), rf.clone()
        p[idx] += h
        m[idx] -= h
        fd = float((w * (E._rf_frequency(is_rf, p, dts) - E._rf_frequency(is_rf, m, dts))).sum()) / (2 * h)
        assert abs(fd - float(g_rf[idx])) < 1e-6 * max(1.0, abs(fd)), (idx, fd, float(g_rf[idx]))
    h = 1e-12
    p, m = dts.clone(), dts.clone()
    p[4] += h
    m[4] -= h
    fd = float((w * (E._rf_frequency(is_rf, rf, p) - E._rf_frequency(is_rf, rf, m))).sum()) / (2 * h)
    assert abs(fd - float(g_dt[4])) < 1e-5 * abs(fd), (fd, float(g_dt[4]))
