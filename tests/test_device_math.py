"""The device's own sin/cos and expm1 (blochparty_amd/csrc/bloch_device.h: sincospi, expm1x), CPU only: their
coefficients are read from the header and evaluated in float32 with single-rounding FMAs, as the GPU evaluates them,
against float64 numpy. A wrong coefficient, a dropped term, or a wrong quadrant or threshold fails here; the GPU
parity tests cover the compiled code end to end."""

import re
from pathlib import Path

import numpy as np

HEADER = Path(__file__).resolve().parent.parent / "blochparty_amd" / "csrc" / "bloch_device.h"
f32 = np.float32


def body(name: str) -> str:
    src = HEADER.read_text()
    start = src.index(f"__device__ __forceinline__ {'void' if name == 'sincospi' else 'float'} {name}(")
    return src[start : src.index("\n}\n", start)]


def floats(text: str) -> list[np.float32]:
    """The float literals of ``text`` in order (``1.0f / 720.0f`` as one value)."""
    out = []
    for m in re.finditer(r"(-?\d+\.\d+f)(?:\s*/\s*(\d+\.\d+f))?", text):
        v = float(m.group(1)[:-1]) / (float(m.group(2)[:-1]) if m.group(2) else 1.0)
        out.append(f32(v))
    return out


def fma(a, b, c):
    return (np.float64(a) * np.float64(b) + np.float64(c)).astype(np.float32)


def constants() -> dict[str, np.float32]:
    """The header's ``constexpr float kName = value;`` constants."""
    src = HEADER.read_text()
    return {m.group(1): f32(float(m.group(2))) for m in re.finditer(r"(k\w+) = (-?[\d.]+(?:e[+-]?\d+)?)f", src)}


def sincospi(x: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    k_ = constants()
    x = x.astype(np.float32)
    k = np.rint(f32(2) * x).astype(np.float32)
    r = fma(f32(-0.5), k, x)
    r2 = (r * r).astype(np.float32)
    ps = fma(k_["kS9"], r2, k_["kS7"])
    for c in ("kS5", "kS3", "kPiLo"):
        ps = fma(ps, r2, k_[c])
    s = fma(r, k_["kPiHi"], (r * ps).astype(np.float32))
    pc = fma(k_["kC10"], r2, k_["kC8"])
    for c in ("kC6", "kC4", "kC2Lo"):
        pc = fma(pc, r2, k_[c])
    c_ = (fma(r2, k_["kC2Hi"], (r2 * pc).astype(np.float32)) + f32(1)).astype(np.float32)
    q = k.astype(np.int64)
    odd = (q & 1) == 1
    s0, c0 = np.where(odd, c_, s), np.where(odd, s, c_)
    return np.where(q & 2, -s0, s0), np.where((q + 1) & 2, -c0, c0)


def expm1x(x: np.ndarray) -> np.ndarray:
    b = body("expm1x")
    coef = floats("\n".join(l for l in b.splitlines() if "float p =" in l or "fmaf(p, x," in l))
    thr = floats(re.search(r"fabsf\(x\) < ([\d.]+f)", b).group(0))[0]
    l2e = floats(re.search(r"x \* ([\d.]+f)", b).group(0))[0]
    assert len(coef) == 6
    x = x.astype(np.float32)
    p = coef[0]
    for c in coef[1:]:
        p = fma(p, x, c)
    small = fma(p, (x * x).astype(np.float32), x)
    big = (np.exp2((x * l2e).astype(np.float32).astype(np.float64)).astype(np.float32) - f32(1)).astype(np.float32)
    return np.where(np.abs(x) < thr, small, big)


def test_sincospi_is_within_one_ulp_and_keeps_the_norm():
    rng = np.random.default_rng(0)
    x = np.concatenate([np.linspace(-1, 1, 400001), rng.uniform(-1, 1, 200000), rng.uniform(-300, 300, 100000),
                        [0.25, 0.5, 0.75, -0.25, -0.5, -0.75, 1.0, -1.0]]).astype(np.float32)
    s, c = sincospi(x)
    xd = x.astype(np.float64)
    assert np.abs(s - np.sin(np.pi * xd)).max() < 2.0**-24
    assert np.abs(c - np.cos(np.pi * xd)).max() < 2.0**-24
    # the rotation's magnitude: no bias for a precession to accumulate over thousands of steps (float32 pi alone
    # would give +1.6e-8 per step)
    norm = s.astype(np.float64) ** 2 + c.astype(np.float64) ** 2 - 1
    assert abs(norm[: 600001].mean()) < 2e-9
    small = np.logspace(-30, -2, 500).astype(np.float32)  # relative accuracy where sin(pi x) is tiny
    ss, cs = sincospi(small)
    assert (np.abs(ss / np.sin(np.pi * small.astype(np.float64)) - 1) < 2.5e-7).all()
    assert (cs <= 1).all()
    # exact values at the quadrant points
    s, c = sincospi(np.array([0.5, 1.0, -0.5, 1.5], dtype=np.float32))
    assert list(s) == [1, 0, -1, -1] and list(np.abs(c)) == [0, 1, 0, 0]


def test_expm1x_keeps_relative_accuracy_for_short_segments():
    rng = np.random.default_rng(1)
    x = -np.concatenate([np.logspace(-12, 2, 400000), rng.uniform(0, 2, 200000)]).astype(np.float32)
    e, t = expm1x(x), np.expm1(x.astype(np.float64))
    rel = np.abs(e - t) / np.abs(t)
    assert rel.max() < 4 * 2.0**-24  # 4 ulp at worst (beside the switch at |x| = 1/4)
    short = np.abs(x) < 0.2
    assert rel[short].max() < 2 * 2.0**-24  # decay factors of short segments: their low bits matter
    assert np.abs(e - t).max() < 1e-7
    assert expm1x(np.array([0.0, -1e-30, -200.0], dtype=np.float32)).tolist() == [0.0, np.float32(-1e-30), -1.0]
