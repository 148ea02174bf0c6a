"""FP64 CPU oracle of the scanner field model (reference ``src/blochparty/fields.py``, used by ``engines/core.py``).

TEST INFRASTRUCTURE. Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU leg may import or run
anything under ``oracle/``, and only as the checker: the package ``blochparty_amd`` never imports it.

Each function takes a different route from the package's (``blochparty_amd/fields.py``), so that the two check each
other:

- ``transverse_field``  fields.py:11-15  (Bx, By) = (gx z - a gz x, gy z - (1 - a) gz y) [Hz] per spin and gradient
- ``concomitant``       fields.py:16-20, core.py:115-121  the field a spin adds in a segment: leading order
                        (Bx^2 + By^2) / (2 f0) from the field itself (the package expands it in six monomial
                        coefficients); exact |B| - f0 - lin = sqrt((f0 + lin)^2 + Bx^2 + By^2) - f0 - lin directly
- ``displacement``      fields.py:36-46, 89-99  the nonlinearity's d(r) from SciPy's associated Legendre functions
                        in spherical coordinates (the package uses a Cartesian recurrence), the Condon-Shortley phase
                        taken out
- ``encoding``          fields.py:101-103  r + d(r)

How each is pinned (tests/test_fields.py): the leading order against Bernstein's published form for the symmetric
coil, [(gx^2 + gy^2) z^2 + gz^2 (x^2 + y^2) / 4 - gx gz x z - gy gz y z] / (2 f0), and against the exact mode at
high field (the difference falls as 1 / f0^2); the exact mode against an extended-precision |B| (numpy longdouble);
the displacement against hand-written n <= 2 terms (n = 1, m = 1: sqrt(3)/2 x A_11; n = 1, m = 0: z A_10; n = 2,
m = 0: (2 z^2 - x^2 - y^2) / (2 R0) A_20; n = 2, m = 1: sqrt(5/12) 3 y z / R0 B_21); the oracle's simulate with a
field against Bernstein's phase of a bipolar gradient pair, -2 pi f_c(g) 2 delta, in free runs and stepped.
"""

from __future__ import annotations

import math

import numpy as np
import torch

GAMMA_BAR = 42.577478518e6


def transverse_field(pos: torch.Tensor, g: torch.Tensor, a: float) -> tuple[torch.Tensor, torch.Tensor]:
    """(Bx, By) [Hz] of spins at true positions pos (N, 3) under gradient g (3,) or (n, 3): (N,) or (N, n)."""
    x, y, z = (pos[:, i : i + 1] if g.dim() == 2 else pos[:, i] for i in range(3))
    gx, gy, gz = (g[:, i][None] if g.dim() == 2 else g[i] for i in range(3))
    return gx * z - a * gz * x, gy * z - (1 - a) * gz * y


def concomitant(mode: int, pos_true, g, lin, f0: float, a: float):
    """The concomitant field [Hz] (N,) or (N, n): mode 1 leading order, 2 exact (lin: the rest of the offset)."""
    bx, by = transverse_field(pos_true, g, a)
    if mode == 1:
        return (bx * bx + by * by) / (2 * f0)
    u = f0 + lin
    return torch.sqrt(u * u + bx * bx + by * by) - u


def displacement(pos: torch.Tensor, alpha: np.ndarray, beta: np.ndarray, r0: float) -> torch.Tensor:
    """(N, 3) d(r) = R0 sum (r/R0)^n N_nm P_n^m(cos theta) (A cos m phi + B sin m phi) per coil, m <= n."""
    from scipy.special import lpmv

    p = pos.detach().cpu().numpy().astype(np.float64)
    r = np.linalg.norm(p, axis=1)
    cos_t = np.where(r > 0, p[:, 2] / np.where(r > 0, r, 1.0), 1.0)
    phi = np.arctan2(p[:, 1], p[:, 0])
    L = alpha.shape[-1]
    out = np.zeros_like(p)
    for n in range(L):
        for m in range(n + 1):
            norm = 1.0 if m == 0 else math.sqrt((2 * n + 1) * math.factorial(n - m) / (2 * math.factorial(n + m)))
            legendre = (-1) ** m * lpmv(m, n, cos_t)  # without the Condon-Shortley phase
            radial = r0 * (r / r0) ** n * norm * legendre
            for axis in range(3):
                out[:, axis] += radial * (alpha[axis, n, m] * np.cos(m * phi) + beta[axis, n, m] * np.sin(m * phi))
    return torch.as_tensor(out, dtype=pos.dtype)


def encoding(pos: torch.Tensor, alpha, beta, r0: float) -> torch.Tensor:
    return pos + displacement(pos, np.asarray(alpha, dtype=np.float64), np.asarray(beta, dtype=np.float64), r0)


class Field:
    """What the oracle's simulate needs of a scanner: true positions, f0 [Hz], z asymmetry, mode (1 leading,
    2 exact)."""

    def __init__(self, pos_true: torch.Tensor, f0: float, a: float, mode: int):
        self.pos, self.f0, self.a, self.mode = pos_true, f0, a, mode

    def at(self, g, lin):
        return concomitant(self.mode, self.pos, g, lin, self.f0, self.a)
