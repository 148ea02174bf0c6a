"""The scanner field model: what a spin sees beyond the ideal B0 + G . r (reference ``src/blochparty/fields.py``,
whose public names, argument meanings and conventions this module keeps: ``simulate(..., scanner=Scanner(...))``).

Concomitant (Maxwell) fields. A gradient coil set cannot produce a pure z field: with the cylindrical coils, and the
z coil's transverse field split a : 1 - a between x and y (a = 1/2 for the symmetric coil), the field is [Hz]

    Bx = gx z - a gz x,   By = gy z - (1 - a) gz y,   Bz = f0 + g . r

and a spin precesses at |B| - f0. ``"leading"`` keeps the first term in 1/f0, (Bx^2 + By^2) / (2 f0): a quadratic
form in g whose six coefficients depend on the spin's true position (``concomitant_coefficients``), so a segment's
phase is the coefficients times the segment's moments of g_i g_j. ``"exact"`` evaluates |B| - f0 - lin per segment,
lin the rest of the spin's offset (``exact_excess``). The coefficients use the true positions and the linear coil
model, also with a nonlinearity.

Gradient nonlinearity. Coil a's field per unit gradient is r_a + d_a(r), d_a a solid-harmonic expansion in the
Siemens convention of gradunwarp's coefficient files,

    d_a(r) = R0 sum_{n, m <= n} (r / R0)^n N_nm P_n^m(cos theta) (A_nm cos(m phi) + B_nm sin(m phi)),

N_nm = sqrt((2n + 1) (n - m)! / (2 (n + m)!)) for m > 0 and 1 for m = 0, P_n^m without the Condon-Shortley phase.
The spins are encoded at f(r) = r + d(r): the kernels take the encoding positions in place of the true ones, so the
coefficients get gradients through the positions' gradients (autograd on the host).

The basis is built in Cartesian form, (x + i y)^m times a polynomial in z and r^2 (``solid_harmonics``), smooth
everywhere and zero wherever the coefficients are.
"""

from __future__ import annotations

import math
import re
from dataclasses import dataclass
from pathlib import Path
from typing import NamedTuple

import torch
from torch import Tensor

GAMMA_BAR = 42.577478518e6  # 1H gyromagnetic ratio / 2 pi [Hz/T]
CONCOMITANT_MODES = ("leading", "exact")


def _norm(n: int, m: int) -> float:
    if m == 0:
        return 1.0
    return math.sqrt((2 * n + 1) * math.factorial(n - m) / (2 * math.factorial(n + m)))


def solid_harmonics(pos: Tensor, n_max: int, r0: float) -> tuple[Tensor, Tensor]:
    """(N, L, L) cosine and sine basis terms R0 (r / R0)^n N_nm P_n^m(cos theta) (cos(m phi), sin(m phi)), L =
    n_max + 1, index [n, m], zero for m > n.

    With s = r / R0: s^n P_n^m(cos theta) (cos + i sin)(m phi) = (sx + i sy)^m q_n^m(sz, |s|^2), where q_m^m is
    (2m - 1)!!, q_{m+1}^m = (2m + 1) sz q_m^m, and (n - m) q_n^m = (2n - 1) sz q_{n-1}^m - (n + m - 1) |s|^2 q_{n-2}^m
    (the associated Legendre recurrence multiplied through by s^n / rho^m)."""
    s = pos / r0
    sx, sy, sz = s[..., 0], s[..., 1], s[..., 2]
    s2 = (s * s).sum(-1)
    L = n_max + 1
    zero = torch.zeros_like(sx)
    cos = [[zero] * L for _ in range(L)]
    sin = [[zero] * L for _ in range(L)]
    re, im = torch.ones_like(sx), torch.zeros_like(sx)  # (sx + i sy)^m
    double_fact = 1.0
    for m in range(L):
        if m > 0:
            re, im = re * sx - im * sy, re * sy + im * sx
            double_fact *= 2 * m - 1
        q_prev, q = None, torch.full_like(sx, double_fact)
        for n in range(m, L):
            if n == m + 1:
                q_prev, q = q, (2 * m + 1) * sz * q
            elif n > m + 1:
                q_prev, q = q, ((2 * n - 1) * sz * q - (n + m - 1) * s2 * q_prev) / (n - m)
            w = r0 * _norm(n, m) * q
            cos[n][m], sin[n][m] = w * re, w * im
    stack = lambda rows: torch.stack([torch.stack(r, -1) for r in rows], -2)
    return stack(cos), stack(sin)


@dataclass
class GradientNonlinearity:
    """Each gradient coil's deviation from linearity as solid-harmonic coefficients.

    alpha, beta: (3, L, L) A_nm, B_nm of the x, y and z coils (entries with m > n are ignored); tensors that require
        grad get gradients.
    r0: the reference radius [m]."""

    alpha: Tensor
    beta: Tensor
    r0: float

    @property
    def n_max(self) -> int:
        return self.alpha.shape[-1] - 1

    def displacement(self, pos: Tensor) -> Tensor:
        """(N, 3) d(r): where a spin is encoded minus where it is [m]."""
        c, s = solid_harmonics(pos, self.n_max, self.r0)
        keep = torch.tril(torch.ones(self.n_max + 1, self.n_max + 1, dtype=torch.bool, device=pos.device))
        a = torch.where(keep, self.alpha.to(pos), torch.zeros((), dtype=pos.dtype, device=pos.device))
        b = torch.where(keep, self.beta.to(pos), torch.zeros((), dtype=pos.dtype, device=pos.device))
        return torch.einsum("pnm,anm->pa", c, a) + torch.einsum("pnm,anm->pa", s, b)

    def encoding_positions(self, pos: Tensor) -> Tensor:
        """(N, 3) f(r) = r + d(r)."""
        return pos + self.displacement(pos)

    @classmethod
    def from_file(cls, path, r0: float | None = None, dtype=torch.float64) -> "GradientNonlinearity":
        """A gradunwarp-style coefficient file: Siemens ``.grad`` / ``.coef`` lines ``<no> A( n, m) <value> <axis>``
        (or ``B``) with the radius on a line ``<r0> m = R0`` (or ``R0 = <r0> [m|mm]``), or lines ``alpha_x n m <value>``
        / ``beta_z n m <value>`` taken as written. ``r0`` [m] overrides or supplies the radius."""
        number = r"[-+]?[\d.]+(?:[eE][-+]?\d+)?"
        siemens = re.compile(rf"^\s*\d+\s+([AaBb])\s*\(\s*(\d+)\s*,\s*(\d+)\s*\)\s+({number})\s+([xyzXYZ])\s*$")
        named = re.compile(rf"^\s*(alpha|beta)_([xyz])\s+(\d+)\s+(\d+)\s+({number})\s*;?\s*$")
        radius_a = re.compile(r"([\d.]+(?:[eE][-+]?\d+)?)\s*m\s*=\s*R0")
        radius_b = re.compile(r"R0\s*=\s*([\d.]+(?:[eE][-+]?\d+)?)\s*(mm|m)?\b")
        found, radius = [], None
        for line in Path(path).read_text().splitlines():
            if line.lstrip().startswith("#"):
                continue
            if hit := siemens.match(line):
                kind, n, m, value, axis = hit.groups()
                found.append(("alpha" if kind in "Aa" else "beta", axis.lower(), int(n), int(m), float(value)))
            elif hit := named.match(line):
                kind, axis, n, m, value = hit.groups()
                found.append((kind, axis, int(n), int(m), float(value)))
            elif hit := radius_a.search(line):
                radius = float(hit.group(1))
            elif hit := radius_b.search(line):
                radius = float(hit.group(1)) * (1e-3 if hit.group(2) == "mm" else 1.0)
        radius = r0 if r0 is not None else radius
        if radius is None:
            raise ValueError(f"{path}: no reference radius (R0) in the file; pass r0")
        if not found:
            raise ValueError(f"{path}: no coefficients found")
        size = 1 + max(max(n, m) for _, _, n, m, _ in found)
        coef = {k: torch.zeros(3, size, size, dtype=dtype) for k in ("alpha", "beta")}
        for kind, axis, n, m, value in found:
            if m > n:
                raise ValueError(f"{path}: coefficient ({n}, {m}) has m > n")
            coef[kind]["xyz".index(axis), n, m] = value
        return cls(coef["alpha"], coef["beta"], radius)


@dataclass
class Scanner:
    """The hardware field model.

    field_strength: B0 [T] (float or tensor, differentiable); concomitant fields are modelled when it is given.
    concomitant: "leading" (the default), "exact" or None (off).
    z_asymmetry: the share a of the z coil's transverse field along x (1 - a along y), 1/2 for the symmetric coil.
    nonlinearity: the coils' deviation from linearity (``GradientNonlinearity``) or None."""

    field_strength: float | Tensor | None = None
    concomitant: str | None = "leading"
    z_asymmetry: float | Tensor = 0.5
    nonlinearity: GradientNonlinearity | None = None

    def __post_init__(self):
        if self.concomitant is not None and self.concomitant not in CONCOMITANT_MODES:
            raise ValueError(f"concomitant must be one of {CONCOMITANT_MODES} or None, got {self.concomitant!r}")

    @property
    def larmor(self) -> float | Tensor:
        """f0 = gamma_bar B0 [Hz]."""
        return GAMMA_BAR * self.field_strength


class SpinFields(NamedTuple):
    """The field model's per-spin inputs. pos: (N, 3) encoding positions; conc: None, (N, 6) leading-order
    coefficients (mode 1) or (N, 7) transverse coefficients and f0 (mode 2); mode: 0 none, 1 leading, 2 exact."""

    pos: Tensor
    conc: Tensor | None
    mode: int


def transverse_coefficients(pos: Tensor, z_asymmetry=0.5) -> Tensor:
    """(N, 6) rows ax, ay with the transverse field (Bx, By) = (ax . g, ay . g) [Hz] at true positions ``pos``."""
    x, y, z = pos[..., 0], pos[..., 1], pos[..., 2]
    a = torch.as_tensor(z_asymmetry, dtype=pos.dtype, device=pos.device)
    zero = torch.zeros_like(x)
    return torch.stack([z, zero, -a * x, zero, z, -(1 - a) * y], -1)


def gradient_monomials(g: Tensor) -> Tensor:
    """(..., 6) gx^2, gy^2, gz^2, gx gy, gx gz, gy gz of gradients (..., 3)."""
    gx, gy, gz = g[..., 0], g[..., 1], g[..., 2]
    return torch.stack([gx * gx, gy * gy, gz * gz, gx * gy, gx * gz, gy * gz], -1)


def concomitant_coefficients(pos: Tensor, f0, z_asymmetry=0.5) -> Tensor:
    """(N, 6) c with the leading-order concomitant field f_c = c . gradient_monomials(g) [Hz] (g in Hz/m, f0 in Hz):
    (ax . g)^2 + (ay . g)^2 over 2 f0, expanded in the monomials."""
    t = transverse_coefficients(pos, z_asymmetry)
    ax, ay = t[..., :3], t[..., 3:]
    square = ax * ax + ay * ay
    cross = lambda i, k: 2 * (ax[..., i] * ax[..., k] + ay[..., i] * ay[..., k])
    c = torch.cat([square, torch.stack([cross(0, 1), cross(0, 2), cross(1, 2)], -1)], -1)
    return c / (2 * f0)


def exact_excess(lin: Tensor, tx: Tensor, ty: Tensor, f0) -> Tensor:
    """|B| - f0 - lin for Bz = f0 + lin and transverse field (tx, ty) [Hz], as t^2 / (|B| + u), u = f0 + lin, which
    does not cancel."""
    t2 = tx * tx + ty * ty
    u = f0 + lin
    return t2 / (torch.sqrt(u * u + t2) + u)


def spin_fields(scanner: Scanner | None, pos: Tensor) -> SpinFields:
    """The field model's inputs for ``scanner`` at true positions ``pos``."""
    if scanner is None:
        return SpinFields(pos, None, 0)
    enc = pos if scanner.nonlinearity is None else scanner.nonlinearity.encoding_positions(pos)
    if scanner.concomitant is None or scanner.field_strength is None:
        return SpinFields(enc, None, 0)
    f0 = torch.as_tensor(scanner.larmor, dtype=pos.dtype, device=pos.device)
    if scanner.concomitant == "leading":
        return SpinFields(enc, concomitant_coefficients(pos, f0, scanner.z_asymmetry), 1)
    t = transverse_coefficients(pos, scanner.z_asymmetry)
    return SpinFields(enc, torch.cat([t, f0.expand(pos.shape[0], 1)], -1), 2)


__all__ = ["CONCOMITANT_MODES", "GAMMA_BAR", "GradientNonlinearity", "Scanner", "SpinFields",
           "concomitant_coefficients", "exact_excess", "gradient_monomials", "solid_harmonics", "spin_fields",
           "transverse_coefficients"]
