"""The learned RF surrogate (``simulate(..., pulses="surrogate")``): a network trained on the GPU, by this package's
own kernels, on exact spinor products of a reused pulse, that maps a spin's offset during the pulse f = df + g . r
and its B1 scale times the pulse's peak b to the pulse's block rotations, from which the spin's pulse map is composed
with the relaxation put back exactly (the kernels' description is in ``csrc/surrogate.hip``). The network sees
features of (f, b): both scaled into the box, and sin, cos of 2 pi f q S for q = 1 .. ``FOURIER``, S the mean block
length over 2 ``FOURIER`` (to first order in the RF a block's rotation is a sum of such terms over its substeps). It
learns each block's quaternion less its first-order (small-tip) part, which is linear in b and, per unit b, a
function of f alone, tabulated per network over its box's f range (``FIRST_GRID`` points, ``hb_sur_first_order``):
what is left is two orders of magnitude smaller in x and y. The relaxation put back around each block gets the
split's correction (first order in the rates, second in the RF) from the blocks' time moments, tabulated beside
(``csrc/surrogate.hip``).

It has no counterpart in the reference. It stands where the reference composes a reused pulse's substeps per spin
(``src/blochparty/engines/gpu/cuda/pulse.cu.in:3-32``, grouped by ``driver.py:109-164``): the reused pulses, their
occurrences and the application of the maps are the ``pulses="reuse"`` path's. The exact composition is both the
fallback and the judge:

- spins whose inputs lie outside the box the network was trained on get the exact maps (``hb_pulse_maps`` with a
  mask), and so do pulses the surrogate does not cover (a gradient that changes during the pulse, fewer substeps
  than blocks);
- on every run the exact maps of a sample of spins (every ``JUDGE_STRIDE``-th) are composed too and compared with
  the network's: ``SimResult.surrogate`` reports the largest difference of a map entry (D or v, magnetization
  units), and if it exceeds ``JUDGE_TOL`` every spin falls back to the exact maps.

Amplitude: a network learns a pulse's shape, its RF divided by its peak |RF| (``PulseShape.scale``), and takes as
its second input the spin's B1 scale times that peak, the spin's peak RF in Hz. Pulses that differ only in amplitude
(a flip-angle change, a train of refocusing flips) share one network: each pulse uses it with the B1 column of the
input normalisation divided by its own peak, which is exact since RF enters the Bloch equations only as B1 x RF.

Training: ``N_TRAIN`` inputs drawn from the box (the range of f and B1 x peak over the phantom's spins and every
pulse of the shape, plus a margin; B1 x peak at least +-``MIN_REL_B1`` of its centre, so a later flip within that
margin finds the network): uniformly, but for a fraction ``FOCUS`` whose f follows the pulse's exact flip angle over
f (where a selective pulse acts, and where its signal comes from); their exact block quaternions from ``hb_spinor_blocks`` (float64), then
``STEPS`` Adam steps on batches of ``BATCH``, the training set reshuffled every epoch (``hb_mlp_train_steps``: a fused
forward + backward, the weight gradients in fixed-order chunk sums, Adam; the step loop runs in C), the step size on
a cosine schedule. Trained networks are kept for the life of the process and reused for any later case of the same
shape whose spins' inputs (not the padded box) lie in their box, when that box is at most ``REUSE_SLACK`` times as
wide as the case's own would be on each axis (a network trained on a much wider box is less accurate).
"""

from __future__ import annotations

import ctypes
import hashlib
import math
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import hip

HIDDEN = 64
BLOCKS = 10
N_OUT = 4 * BLOCKS
FOURIER = 4  # HB_MLP_FOURIER
N_IN = 2 + 2 * FOURIER  # HB_MLP_INPUTS
NORM_HEAD = 6  # HB_MLP_NORM_HEAD
FIRST_GRID = 2049  # HB_MLP_FIRST_GRID
N_PARAMS = HIDDEN * N_IN + HIDDEN + HIDDEN * HIDDEN + HIDDEN + N_OUT * HIDDEN + N_OUT
N_TRAIN = 1 << 20
STEPS = 20000
BATCH = 8192
LR = (1e-2, 1e-5)  # cosine from the first to the second
FOCUS = 0.25  # fraction of the training inputs drawn with density proportional to the pulse's flip angle over f
FOCUS_GRID = 4097  # offsets at which that flip angle is taken
MARGIN = 0.05  # the box extends the phantom's range by this fraction of its width on each side
MIN_HALF = (1.0, 1e-6)  # smallest half-widths of the box: f [Hz], b1 x peak [Hz]
MIN_REL_B1 = 0.01  # the box's B1 x peak half-width is at least this fraction of its centre
REUSE_SLACK = 1.25  # a cached network serves a case whose padded box is at least 1 / this of its own, per axis
JUDGE_STRIDE = 257
JUDGE_TOL = 1e-3

_CACHE: dict[tuple, "Net"] = {}


class HbMlp(ctypes.Structure):
    """``hb_mlp`` of the header."""

    _fields_ = [("params", ctypes.c_void_p), ("adam_m", ctypes.c_void_p), ("adam_v", ctypes.c_void_p),
                ("norm", ctypes.c_void_p), ("hidden", ctypes.c_int32), ("n_out", ctypes.c_int32),
                ("first", ctypes.c_void_p), ("n_grid", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class HbSurPulse(ctypes.Structure):
    """``hb_sur_pulse`` of the header."""

    _fields_ = [("blocks", ctypes.c_void_p), ("g", ctypes.c_double * 3), ("pid", ctypes.c_int32),
                ("pad", ctypes.c_int32)]


@dataclass
class Net:
    params: torch.Tensor  # (N_PARAMS,) float32
    adam_m: torch.Tensor
    adam_v: torch.Tensor
    norm: torch.Tensor  # (NORM_HEAD + 2 N_OUT,) float32: c_f, h_f, c_b, h_b, S, s_b, output means, scales
    box: tuple  # ((centre f, b1 x peak), (half-widths))
    train_s: float = 0.0
    final_loss: float = float("nan")
    first: torch.Tensor | None = None  # (BLOCKS, FIRST_GRID, 8) float32 F_j, H_j per Hz of b, G_j per Hz^2, 0

    def struct(self, norm: torch.Tensor | None = None) -> HbMlp:
        """The ``hb_mlp`` of this network (with ``norm`` in place of its own: a pulse's, ``pulse_norm``)."""
        norm = self.norm if norm is None else norm
        first = None if self.first is None else self.first.data_ptr()
        return HbMlp(self.params.data_ptr(), self.adam_m.data_ptr(), self.adam_v.data_ptr(), norm.data_ptr(),
                     HIDDEN, N_OUT, first, FIRST_GRID if first else 0, 0)


@dataclass
class PulseShape:
    """One reused pulse as the surrogate sees it."""

    table: np.ndarray  # (K, 8) float64 rows of its first occurrence (pulses.pulse_table's columns), RF / scale
    bounds: np.ndarray  # (BLOCKS + 1,) int32 block boundaries (substep indices)
    blocks: np.ndarray  # (BLOCKS, 2) float64 tau_j, b0_offset area B_j
    g: np.ndarray  # (3,) float64 the constant gradient [Hz/m]
    scale: float = 1.0  # the pulse's peak |RF| [Hz]

    @property
    def fourier_base(self) -> float:
        """S [s]: the features' phases are 2 pi f q S, q = 1 .. FOURIER, up to half the mean block length."""
        return float(self.blocks[:, 0].mean()) / (2 * FOURIER)

    def key(self) -> str:
        """The shape's identity: every column quantised to 2^-32 of its largest value (pulses equal up to rounding
        of the division by their peak share a key)."""
        peak = np.abs(self.table).max(0)
        q = np.round(self.table / np.where(peak > 0, peak, 1.0) * 2.0**32).astype(np.int64)
        head = np.array([float(f"{v:.9e}") for v in peak])
        return hashlib.sha1(q.tobytes() + head.tobytes()).hexdigest()


def block_bounds(k: int) -> np.ndarray:
    """BLOCKS + 1 substep indices cutting K substeps into near-equal blocks."""
    return np.round(np.linspace(0, k, BLOCKS + 1)).astype(np.int32)


def shape_of(table: np.ndarray) -> PulseShape | None:
    """The surrogate's view of a pulse's rows, or None when it does not cover the pulse (a gradient that changes
    during the pulse, or fewer substeps than blocks)."""
    k = len(table)
    g = table[:, 3:6]
    scale = float(np.hypot(table[:, 1], table[:, 2]).max()) if k else 0.0
    if k < BLOCKS or scale == 0.0 or np.abs(g - g[0]).max() > 1e-9 * max(np.abs(g).max(), 1e-300):
        return None
    shape_rows = np.array(table, dtype=np.float64)
    shape_rows[:, 1:3] /= scale
    bounds = block_bounds(k)
    tau = np.add.reduceat(table[:, 0], bounds[:-1])
    area = np.add.reduceat(table[:, 0] * table[:, 6], bounds[:-1])
    return PulseShape(np.ascontiguousarray(shape_rows), bounds, np.ascontiguousarray(np.stack([tau, area], 1)),
                      g[0].copy(), scale)


def input_range(spins: torch.Tensor, pos_lo: torch.Tensor, n: int, g: np.ndarray, scale: float) -> np.ndarray:
    """(2, 2) low and high of the spins' inputs (f = df + g . r, b1 x the pulse's peak)."""
    pos = spins[:3, :n].double() + pos_lo[:, :n].double()
    u = (torch.as_tensor(g, dtype=torch.float64, device=pos.device)[:, None] * pos).sum(0)
    cols = (spins[6, :n].double() + u, spins[7, :n].double() * scale)
    return np.array([[float(c.min()) for c in cols], [float(c.max()) for c in cols]])


def padded_box(rng: np.ndarray) -> tuple:
    """(centre, half-widths) of the training box around an input range: the margin on each side, and the minimum
    half-widths (B1 x peak: ``MIN_REL_B1`` of its centre)."""
    lo, hi = rng
    centre = 0.5 * (lo + hi)
    floor = np.array(MIN_HALF, dtype=np.float64)
    floor[1] = max(floor[1], MIN_REL_B1 * abs(centre[1]))
    half = np.maximum(0.5 * (hi - lo) * (1 + 2 * MARGIN), floor)
    return tuple(centre), tuple(half)


def pulse_norm(net: "Net", shape: PulseShape) -> torch.Tensor:
    """``net``'s normalisation for a pulse of peak ``shape.scale``: the B1 column's centre and half-width divided by
    the peak and s_b = the peak, so the kernels' per-spin B1 scale enters as B1 x peak."""
    norm = net.norm.clone()
    norm[2] /= shape.scale
    norm[3] /= shape.scale
    norm[5] = shape.scale
    return norm


def _check(rc: int, lib, name: str) -> None:
    if rc != 0:
        raise RuntimeError(f"HIP error {rc} in {name}: {lib.hb_error_string(rc).decode()}")


def stream() -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def spinor_blocks(shape: PulseShape, x: torch.Tensor) -> torch.Tensor:
    """(n, N_OUT) exact interaction-frame block quaternions at inputs x (n, 2) (hb_spinor_blocks)."""
    lib = hip.load()
    dev = x.device
    table = torch.from_numpy(shape.table).to(dev)
    bounds = torch.from_numpy(shape.bounds).to(dev)
    y = torch.empty(len(x), N_OUT, dtype=torch.float32, device=dev)
    _check(lib.hb_spinor_blocks(table.data_ptr(), bounds.data_ptr(), x.data_ptr(), len(x), y.data_ptr(), stream()),
           lib, "hb_spinor_blocks")
    return y


def pulse_flip(shape: PulseShape, f: torch.Tensor, b: float) -> np.ndarray:
    """The exact flip angle [rad] (the tilt of the z axis) of the whole pulse at offsets ``f`` (float32, on the
    device) and B1 x peak ``b``: the block quaternions composed in the lab frame (each block's Rz(theta/2) before and
    after, theta = -2 pi (f tau + B), as the maps kernel composes them)."""
    x = torch.stack([f, torch.full_like(f, b)], 1).contiguous()
    q = spinor_blocks(shape, x).double().cpu().numpy().reshape(len(f), BLOCKS, 4)
    fd = f.double().cpu().numpy()
    w, xx, yy, zz = np.ones_like(fd), np.zeros_like(fd), np.zeros_like(fd), np.zeros_like(fd)

    def mul(a, c):
        return (a[0] * c[0] - a[1] * c[1] - a[2] * c[2] - a[3] * c[3], a[0] * c[1] + a[1] * c[0] + a[2] * c[3] - a[3] * c[2],
                a[0] * c[2] - a[1] * c[3] + a[2] * c[0] + a[3] * c[1], a[0] * c[3] + a[1] * c[2] - a[2] * c[1] + a[3] * c[0])

    for j in range(BLOCKS):
        theta = -2 * math.pi * (fd * shape.blocks[j, 0] + shape.blocks[j, 1])
        h = (np.cos(0.25 * theta), 0.0 * fd, 0.0 * fd, np.sin(0.25 * theta))
        qj = q[:, j] / np.linalg.norm(q[:, j], axis=1, keepdims=True)
        w, xx, yy, zz = mul(mul(h, mul(tuple(qj.T), h)), (w, xx, yy, zz))
    return 2 * np.arctan2(np.hypot(xx, yy), np.hypot(w, zz))


def training_inputs(shape: PulseShape, box: tuple, n: int, gen: torch.Generator, dev) -> torch.Tensor:
    """(n, 2) float32 training inputs (f, b) on the device: uniform over the box, except a fraction ``FOCUS`` whose
    offsets follow the density of the pulse's flip angle over the box's f range (at its centre B1): where a
    selective pulse acts, its maps change fastest, and there lie the spins that carry its signal."""
    c, h = torch.tensor(box[0], dtype=torch.float64), torch.tensor(box[1], dtype=torch.float64)
    x = (torch.rand(n, 2, generator=gen, dtype=torch.float64) * 2 - 1) * h + c
    grid = torch.linspace(float(c[0] - h[0]), float(c[0] + h[0]), FOCUS_GRID, dtype=torch.float64)
    flip = pulse_flip(shape, grid.float().to(dev), float(c[1]))
    nf = int(FOCUS * n)
    if nf and np.isfinite(flip).all() and flip.sum() > 0:
        cdf = np.concatenate([[0.0], np.cumsum(0.5 * (flip[1:] + flip[:-1]))])
        u = torch.rand(nf, generator=gen, dtype=torch.float64).numpy() * cdf[-1]
        x[:nf, 0] = torch.from_numpy(np.interp(u, cdf, grid.numpy()))  # piecewise-linear density between grid points
    return x.float().to(dev)


def init_params(seed: int) -> np.ndarray:
    """Seeded Glorot-uniform weights, zero biases, in the parameter layout."""
    rng = np.random.default_rng(seed)
    out = []
    for fan_in, fan_out in ((N_IN, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, N_OUT)):
        lim = math.sqrt(6.0 / (fan_in + fan_out))
        out += [rng.uniform(-lim, lim, fan_out * fan_in), np.zeros(fan_out)]
    p = np.concatenate(out).astype(np.float32)
    assert len(p) == N_PARAMS
    return p


def new_net(box: tuple, base: float, mu: torch.Tensor, sd: torch.Tensor, seed: int, dev,
            first: torch.Tensor | None = None) -> Net:
    """An untrained network on ``box`` with features' base time ``base`` [s], output means / scales and first-order
    table ``first`` (None: none)."""
    (cf, cb), (hf, hb) = box
    head = torch.tensor([cf, hf, cb, hb, base, 1.0], dtype=torch.float32, device=dev)
    norm = torch.cat([head, mu.float().to(dev), sd.float().to(dev)])
    params = torch.from_numpy(init_params(seed)).to(dev)
    return Net(params, torch.zeros_like(params), torch.zeros_like(params), norm.contiguous(), box, first=first)


def first_order(shape: PulseShape, box: tuple, dev) -> torch.Tensor:
    """(BLOCKS, FIRST_GRID, 8) the pulse's first-order block rotation vectors F_j, their time moments H_j (per Hz of
    b) and G_j (per Hz^2 of b) over the box's f range (hb_sur_first_order)."""
    lib = hip.load()
    (cf, _), (hf, _) = box
    lo, step = float(np.float32(cf) - np.float32(hf)), float(2 * np.float32(hf) / np.float32(FIRST_GRID - 1))
    table = torch.from_numpy(shape.table).to(dev)
    bounds = torch.from_numpy(shape.bounds).to(dev)
    out = torch.empty(BLOCKS, FIRST_GRID, 8, dtype=torch.float32, device=dev)
    _check(lib.hb_sur_first_order(table.data_ptr(), bounds.data_ptr(), lo, step, FIRST_GRID, out.data_ptr(), stream()),
           lib, "hb_sur_first_order")
    return out


def first_quats(net: Net, x: torch.Tensor) -> torch.Tensor:
    """(n, N_OUT) the first-order block quaternions at inputs x (n, 2) (hb_sur_first_quats)."""
    lib = hip.load()
    q = torch.empty(len(x), N_OUT, dtype=torch.float32, device=x.device)
    _check(lib.hb_sur_first_quats(ctypes.byref(net.struct()), x.data_ptr(), len(x), q.data_ptr(), stream()), lib,
           "hb_sur_first_quats")
    return q


def train(shape: PulseShape, box: tuple, *, n_train: int = N_TRAIN, steps: int = STEPS, batch: int = BATCH,
          lr: tuple = LR, seed: int = 0, dev=None) -> Net:
    """A network for ``shape`` on ``box``, trained on the GPU (module docstring)."""
    lib = hip.load()
    dev = dev or hip.HIP.device
    t0 = time.perf_counter()
    gen = torch.Generator(device="cpu").manual_seed(seed)
    c, h = torch.tensor(box[0], dtype=torch.float64), torch.tensor(box[1], dtype=torch.float64)
    x = training_inputs(shape, box, n_train, gen, dev)
    net = new_net(box, shape.fourier_base, torch.zeros(N_OUT), torch.ones(N_OUT), seed, dev,
                  first_order(shape, box, dev))
    y = spinor_blocks(shape, x) - first_quats(net, x)  # the network's targets: what the first order leaves
    mu, sd = y.mean(0), y.std(0) + 1e-3
    net.norm[NORM_HEAD:] = torch.cat([mu, sd])
    st = net.struct()
    work = torch.empty(lib.hb_mlp_workspace(batch), dtype=torch.float32, device=dev)
    gdev = torch.Generator(device=dev).manual_seed(seed)
    per_epoch = n_train // batch
    done = 0
    while done < steps:  # epochs: the training set in a fresh random order, consumed in consecutive batches
        perm = torch.randperm(n_train, generator=gdev, device=dev)
        xs, ys = x[perm].contiguous(), y[perm].contiguous()
        k = min(per_epoch, steps - done)
        _check(lib.hb_mlp_train_steps(ctypes.byref(st), xs.data_ptr(), ys.data_ptr(), batch, k, work.data_ptr(), done,
                                      steps, lr[0], lr[1], stream()), lib, "hb_mlp_train_steps")
        done += k
    # final loss on a held-out draw, in normalised units
    xt = ((torch.rand(65536, 2, generator=gen, dtype=torch.float64) * 2 - 1) * h + c).float().to(dev)
    yt = spinor_blocks(shape, xt) - first_quats(net, xt)
    net.final_loss = float((((evaluate(net, xt) - yt) / sd) ** 2).mean())
    torch.cuda.synchronize()
    net.train_s = time.perf_counter() - t0
    return net


def evaluate(net: Net, x: torch.Tensor) -> torch.Tensor:
    """(n, N_OUT) de-normalised network outputs (hb_mlp_eval), without the first-order part."""
    lib = hip.load()
    y = torch.empty(len(x), N_OUT, dtype=torch.float32, device=x.device)
    _check(lib.hb_mlp_eval(ctypes.byref(net.struct()), x.data_ptr(), len(x), y.data_ptr(), stream()), lib,
           "hb_mlp_eval")
    return y


def block_quats(net: Net, x: torch.Tensor) -> torch.Tensor:
    """(n, N_OUT) the block quaternions the network stands for (first-order part + outputs), not normalised."""
    return first_quats(net, x) + evaluate(net, x)


def gradient(net: Net, x: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
    """(N_PARAMS,) the loss gradient on one batch (hb_mlp_grad)."""
    lib = hip.load()
    work = torch.empty(lib.hb_mlp_workspace(len(x)), dtype=torch.float32, device=x.device)
    g = torch.empty(N_PARAMS, dtype=torch.float32, device=x.device)
    _check(lib.hb_mlp_grad(ctypes.byref(net.struct()), x.data_ptr(), y.data_ptr(), len(x), work.data_ptr(),
                           g.data_ptr(), stream()), lib, "hb_mlp_grad")
    return g


def prepare(case, seq) -> None:
    """Train (or fetch) a network for every reused pulse of ``case`` the surrogate covers, one per pulse shape (the
    pulses of one shape, whatever their amplitudes, share it); sets ``case.extra["surrogate"]`` = [(pid, shape,
    net)]."""
    table = case.pulse_table.cpu().numpy()
    index = case.pulse_index.cpu().numpy()
    by_key: dict[str, list] = {}
    for pid, (off, k) in enumerate(index[: case.n_pulses]):
        shape = shape_of(table[off : off + k])
        if shape is not None:
            by_key.setdefault(shape.key(), []).append((pid, shape))
    nets = []
    for key, members in by_key.items():
        rngs = [input_range(case.spins, case.pos_lo, case.n, shape.g, shape.scale) for _, shape in members]
        rng = np.stack([np.min([r[0] for r in rngs], 0), np.max([r[1] for r in rngs], 0)])
        net = cached(key, rng)
        if net is None:
            box = padded_box(rng)
            net = train(members[0][1], box, dev=case.spins.device)
            _CACHE[(key, box)] = net
        nets += [(pid, shape, net) for pid, shape in members]
    case.extra["surrogate"] = sorted(nets, key=lambda t: t[0])


def cached(key: str, rng: np.ndarray) -> Net | None:
    """A network trained on this pulse shape whose box holds the input range ``rng`` and is not much wider than the
    box ``rng`` would be trained on (``REUSE_SLACK`` per axis: a network spreads its accuracy over its whole box), the
    smallest such; or None."""
    want = np.array(padded_box(rng)[1])
    best = None
    for (k, b), net in _CACHE.items():
        lo, hi = np.array(b[0]) - np.array(b[1]), np.array(b[0]) + np.array(b[1])
        inside = bool(np.all(rng[0] >= lo) and np.all(rng[1] <= hi))
        snug = bool(np.all(np.array(b[1]) <= REUSE_SLACK * want))
        if k == key and inside and snug and (best is None or np.prod(b[1]) < np.prod(best.box[1])):
            best = net
    return best


def compose_maps(case, lib) -> None:
    """The pulse maps of ``case`` from the networks (``prepare``), the exact fallback, and the judge; the report goes
    to ``case.extra["surrogate_report"]`` (device tensors until ``report``)."""
    nets = case.extra["surrogate"]
    st = case.struct()
    dev = case.spins.device
    if len(nets) < case.n_pulses:  # pulses the surrogate does not cover
        _check(lib.hb_pulse_maps(ctypes.byref(st), case.pulse_table.data_ptr(), case.pulse_index.data_ptr(),
                                 case.n_pulses, None, 0, case.maps.data_ptr(), stream()), lib, "hb_pulse_maps")
    outside = torch.zeros(case.n_pad, dtype=torch.int32, device=dev)
    judge = case.extra.get("judge_list")
    if judge is None:  # every JUDGE_STRIDE-th spin, composed exactly on every run
        judge = torch.arange(0, case.n, JUDGE_STRIDE, dtype=torch.int32, device=dev)
        case.extra["judge_list"] = judge
        case.extra["judge_maps"] = torch.empty_like(case.maps)
    jm = case.extra["judge_maps"]
    # the judge: exact maps of the sampled spins, every pulse (one lane per judged spin)
    _check(lib.hb_pulse_maps(ctypes.byref(st), case.pulse_table.data_ptr(), case.pulse_index.data_ptr(),
                             case.n_pulses, judge.data_ptr(), len(judge), jm.data_ptr(), stream()), lib,
           "hb_pulse_maps")
    sel = judge.long()
    report = []
    for pid, shape, net in nets:
        blocks = torch.from_numpy(shape.blocks).to(dev)
        sp = HbSurPulse(blocks.data_ptr(), (ctypes.c_double * 3)(*shape.g), pid, 0)
        norm = pulse_norm(net, shape)
        mlp = net.struct(norm)
        _check(lib.hb_surrogate_maps(ctypes.byref(st), ctypes.byref(mlp), ctypes.byref(sp), case.maps.data_ptr(),
                                     outside.data_ptr(), stream()), lib, "hb_surrogate_maps")
        mine = case.maps.view(case.n_pulses, 12, case.n_pad)[pid][:, sel]
        exact = jm.view(case.n_pulses, 12, case.n_pad)[pid][:, sel]
        err = (mine - exact).abs().max()
        # the fallback: spins outside the box, composed exactly (listed on the device; nothing to do for none)
        out_list = torch.nonzero(outside).flatten().to(torch.int32)
        if len(out_list):
            _check(lib.hb_pulse_maps(ctypes.byref(st), case.pulse_table.data_ptr(), case.pulse_index.data_ptr(),
                                     case.n_pulses, out_list.data_ptr(), len(out_list), case.maps.data_ptr(),
                                     stream()), lib, "hb_pulse_maps")
        report.append((pid, err, len(out_list), net))
        del blocks
    case.extra["surrogate_report"] = report
    fallback = [r for r in report if float(r[1]) > JUDGE_TOL]
    if fallback:  # the judge rejects a network: exact maps for every spin of every pulse
        _check(lib.hb_pulse_maps(ctypes.byref(st), case.pulse_table.data_ptr(), case.pulse_index.data_ptr(),
                                 case.n_pulses, None, 0, case.maps.data_ptr(), stream()), lib, "hb_pulse_maps")
        case.extra["surrogate_rejected"] = True


def report(case) -> dict:
    """The judge's report of the last run: per pulse the largest map difference on the sample, the spins outside
    the box, the training time and held-out loss; and whether the exact maps replaced the network."""
    rows = []
    for pid, err, n_out, net in case.extra.get("surrogate_report", []):
        rows.append({"pulse": pid, "judge_max_map_err": float(err), "judge_spins": len(case.extra["judge_list"]),
                     "outside_box": int(n_out), "train_s": net.train_s, "held_out_loss": net.final_loss})
    return {"pulses": rows, "covered": len(rows), "reused_pulses": case.n_pulses,
            "rejected": bool(case.extra.get("surrogate_rejected", False)), "tolerance": JUDGE_TOL}
