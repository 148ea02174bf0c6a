"""The reverse sweep's algebra at a reused pulse's occurrence (blochparty_amd/csrc/adjoint.hip: pulse_reverse), CPU
only, restated in float64 against the forward's application of the map (bloch_device.h: apply_map,
m' = m + Rz(psi) (D Rz(-psi) m + v)):
- lambda's step is the transpose of the forward's: <lambda', M m> = <M^T lambda', m>;
- the seeds t u^T and t (t = Rz(-psi) lambda', u = Rz(-psi) m) are the derivatives of <lambda', m'> in D's entries
  and v's;
- the RF-phase gradient (lambda' x m')_z - (lambda x m)_z is the derivative of <lambda', m'> in psi.
Each is checked by central differences, so a transposed map, a wrong frame turn or a sign error fails here; the GPU
tests check the compiled kernel end to end against the oracle."""

import numpy as np


def rz(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])


def apply_map(D, v, psi, m):
    """The forward's occurrence: m + Rz(psi) (D Rz(-psi) m + v)."""
    return m + rz(psi) @ (D @ (rz(-psi) @ m) + v)


def pulse_reverse(D, v, psi, m, lam):
    """The kernel's reverse, restated: (lambda before, seeds on D, seeds on v, dL/dpsi)."""
    c, s = np.cos(psi), np.sin(psi)
    t = np.array([c * lam[0] + s * lam[1], c * lam[1] - s * lam[0], lam[2]])
    u = np.array([c * m[0] + s * m[1], c * m[1] - s * m[0], m[2]])
    b = t + D.T @ t
    lam_b = np.array([c * b[0] - s * b[1], s * b[0] + c * b[1], b[2]])
    ma = apply_map(D, v, psi, m)
    dpsi = (lam[1] * ma[0] - lam[0] * ma[1]) - (lam_b[1] * m[0] - lam_b[0] * m[1])
    return lam_b, np.outer(t, u), t, dpsi


def test_pulse_reverse_is_the_transpose_and_its_seeds_are_the_derivatives():
    rng = np.random.default_rng(2)
    for _ in range(20):
        D = rng.normal(scale=0.3, size=(3, 3))
        v = rng.normal(size=3)
        psi = rng.uniform(-np.pi, np.pi)
        m, lam = rng.normal(size=3), rng.normal(size=3)
        lam_b, sD, sv, dpsi = pulse_reverse(D, v, psi, m, lam)
        # the linear part's transpose: <lam, M m> = <lam_b, m> with M m = apply_map(m) - apply_map(0)
        lin = apply_map(D, v, psi, m) - apply_map(D, v, psi, np.zeros(3))
        assert abs(lam @ lin - lam_b @ m) < 1e-12
        L = lambda D_, v_, p_: lam @ apply_map(D_, v_, p_, m)
        h = 1e-6
        for i in range(3):
            for k in range(3):
                E = np.zeros((3, 3))
                E[i, k] = h
                fd = (L(D + E, v, psi) - L(D - E, v, psi)) / (2 * h)
                assert abs(fd - sD[i, k]) < 1e-8
            e = np.zeros(3)
            e[i] = h
            assert abs((L(D, v + e, psi) - L(D, v - e, psi)) / (2 * h) - sv[i]) < 1e-8
        fd_psi = (L(D, v, psi + h) - L(D, v, psi - h)) / (2 * h)
        assert abs(fd_psi - dpsi) < 1e-7 * max(1.0, abs(dpsi))


def test_a_wrong_frame_turn_is_caught():
    """Planted mistake: seeds formed in the lab frame instead of the pulse's frame differ from the derivatives."""
    rng = np.random.default_rng(3)
    D, v, psi = rng.normal(scale=0.3, size=(3, 3)), rng.normal(size=3), 0.7
    m, lam = rng.normal(size=3), rng.normal(size=3)
    _, sD, _, _ = pulse_reverse(D, v, psi, m, lam)
    assert np.abs(np.outer(lam, m) - sD).max() > 1e-2


def test_adjoint_width_and_coil_tiles():
    """Coils up to 16 are padded to a register width; more take whole tiles of 16 (one sweep per tile)."""
    from blochparty_amd.engines import adjoint

    assert [adjoint.adjoint_width(c) for c in (1, 2, 3, 5, 8, 9, 16)] == [1, 2, 4, 8, 8, 16, 16]
    assert [adjoint.adjoint_width(c) for c in (17, 20, 32, 33)] == [32, 32, 32, 48]
