"""The learned RF surrogate (``simulate(..., pulses="surrogate")``): a network trained on the GPU, by this package's
own kernels, on exact spinor products of a reused pulse, that maps a spin's (df, u = g . r, b1) to the pulse's block
rotations, from which the spin's pulse map is composed with the relaxation put back exactly (the kernels'
description is in ``csrc/surrogate.hip``).

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
its third input the spin's B1 scale times that peak, the spin's peak RF in Hz. Pulses that differ only in amplitude
(a flip-angle change, a train of refocusing flips) share one network: each pulse uses it with the B1 column of the
input normalisation divided by its own peak, which is exact since RF enters the Bloch equations only as B1 x RF.

Training: ``N_TRAIN`` inputs drawn uniformly from the box (the range of df, u and B1 x peak over the phantom's spins
and every pulse of the shape, plus a margin; B1 x peak at least +-``MIN_REL_B1`` of its centre, so a later flip
within that margin finds the network), their exact block quaternions from ``hb_spinor_blocks`` (float64), then
``STEPS`` Adam steps on batches of ``BATCH``, the training set reshuffled every epoch (``hb_mlp_train_steps``: a fused
forward + backward, the weight gradients in fixed-order chunk sums, Adam; the step loop runs in C), the step size on
a cosine schedule. Trained networks are kept for the life of the process and reused for any later case of the same
shape whose spins' inputs (not the padded box) lie in their box.
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
N_PARAMS = HIDDEN * 3 + HIDDEN + HIDDEN * HIDDEN + HIDDEN + N_OUT * HIDDEN + N_OUT
N_TRAIN = 1 << 20
STEPS = 20000
BATCH = 8192
LR = (3e-3, 1e-5)  # cosine from the first to the second
MARGIN = 0.05  # the box extends the phantom's range by this fraction of its width on each side
MIN_HALF = (1.0, 1.0, 1e-6)  # smallest half-widths of the box: df [Hz], u [Hz], b1 x peak [Hz]
MIN_REL_B1 = 0.01  # the box's B1 x peak half-width is at least this fraction of its centre
JUDGE_STRIDE = 257
JUDGE_TOL = 1e-3

_CACHE: dict[tuple, "Net"] = {}


class HbMlp(ctypes.Structure):
    """``hb_mlp`` of the header."""

    _fields_ = [("params", ctypes.c_void_p), ("adam_m", ctypes.c_void_p), ("adam_v", ctypes.c_void_p),
                ("norm", ctypes.c_void_p), ("hidden", ctypes.c_int32), ("n_out", ctypes.c_int32)]


class HbSurPulse(ctypes.Structure):
    """``hb_sur_pulse`` of the header."""

    _fields_ = [("blocks", ctypes.c_void_p), ("g", ctypes.c_double * 3), ("pid", ctypes.c_int32),
                ("pad", ctypes.c_int32)]


@dataclass
class Net:
    params: torch.Tensor  # (N_PARAMS,) float32
    adam_m: torch.Tensor
    adam_v: torch.Tensor
    norm: torch.Tensor  # (6 + 2 N_OUT,) float32: box centre, half-widths, output means, scales
    box: tuple  # ((centre df, u, b1), (half-widths))
    train_s: float = 0.0
    final_loss: float = float("nan")

    def struct(self) -> HbMlp:
        return HbMlp(self.params.data_ptr(), self.adam_m.data_ptr(), self.adam_v.data_ptr(), self.norm.data_ptr(),
                     HIDDEN, N_OUT)


@dataclass
class PulseShape:
    """One reused pulse as the surrogate sees it."""

    table: np.ndarray  # (K, 8) float64 rows of its first occurrence (pulses.pulse_table's columns), RF / scale
    bounds: np.ndarray  # (BLOCKS + 1,) int32 block boundaries (substep indices)
    blocks: np.ndarray  # (BLOCKS, 2) float64 tau_j, b0_offset area B_j
    g: np.ndarray  # (3,) float64 the constant gradient [Hz/m]
    scale: float = 1.0  # the pulse's peak |RF| [Hz]

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
    """(2, 3) low and high of the spins' inputs (df, u, b1 x the pulse's peak)."""
    pos = spins[:3, :n].double() + pos_lo[:, :n].double()
    u = (torch.as_tensor(g, dtype=torch.float64, device=pos.device)[:, None] * pos).sum(0)
    cols = (spins[6, :n].double(), u, spins[7, :n].double() * scale)
    return np.array([[float(c.min()) for c in cols], [float(c.max()) for c in cols]])


def padded_box(rng: np.ndarray) -> tuple:
    """(centre, half-widths) of the training box around an input range: the margin on each side, and the minimum
    half-widths (B1 x peak: ``MIN_REL_B1`` of its centre)."""
    lo, hi = rng
    centre = 0.5 * (lo + hi)
    floor = np.array(MIN_HALF, dtype=np.float64)
    floor[2] = max(floor[2], MIN_REL_B1 * abs(centre[2]))
    half = np.maximum(0.5 * (hi - lo) * (1 + 2 * MARGIN), floor)
    return tuple(centre), tuple(half)


def pulse_norm(net: "Net", shape: PulseShape) -> torch.Tensor:
    """``net``'s normalisation for a pulse of peak ``shape.scale``: the B1 column's centre and half-width divided by
    the peak, so the kernels' per-spin B1 scale enters as B1 x peak."""
    norm = net.norm.clone()
    norm[2] /= shape.scale
    norm[5] /= shape.scale
    return norm


def _check(rc: int, lib, name: str) -> None:
    if rc != 0:
        raise RuntimeError(f"HIP error {rc} in {name}: {lib.hb_error_string(rc).decode()}")


def stream() -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def spinor_blocks(shape: PulseShape, x: torch.Tensor) -> torch.Tensor:
    """(n, N_OUT) exact interaction-frame block quaternions at inputs x (n, 3) (hb_spinor_blocks)."""
    lib = hip.load()
    dev = x.device
    table = torch.from_numpy(shape.table).to(dev)
    bounds = torch.from_numpy(shape.bounds).to(dev)
    y = torch.empty(len(x), N_OUT, dtype=torch.float32, device=dev)
    _check(lib.hb_spinor_blocks(table.data_ptr(), bounds.data_ptr(), x.data_ptr(), len(x), y.data_ptr(), stream()),
           lib, "hb_spinor_blocks")
    return y


def init_params(seed: int) -> np.ndarray:
    """Seeded Glorot-uniform weights, zero biases, in the parameter layout."""
    rng = np.random.default_rng(seed)
    out = []
    for fan_in, fan_out in ((3, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, N_OUT)):
        lim = math.sqrt(6.0 / (fan_in + fan_out))
        out += [rng.uniform(-lim, lim, fan_out * fan_in), np.zeros(fan_out)]
    p = np.concatenate(out).astype(np.float32)
    assert len(p) == N_PARAMS
    return p


def new_net(box: tuple, mu: torch.Tensor, sd: torch.Tensor, seed: int, dev) -> Net:
    norm = torch.cat([torch.tensor(box[0] + box[1], dtype=torch.float32, device=dev), mu.float(), sd.float()])
    params = torch.from_numpy(init_params(seed)).to(dev)
    return Net(params, torch.zeros_like(params), torch.zeros_like(params), norm.contiguous(), box)


def train(shape: PulseShape, box: tuple, *, n_train: int = N_TRAIN, steps: int = STEPS, batch: int = BATCH,
          lr: tuple = LR, seed: int = 0, dev=None) -> Net:
    """A network for ``shape`` on ``box``, trained on the GPU (module docstring)."""
    lib = hip.load()
    dev = dev or hip.HIP.device
    t0 = time.perf_counter()
    gen = torch.Generator(device="cpu").manual_seed(seed)
    c, h = torch.tensor(box[0], dtype=torch.float64), torch.tensor(box[1], dtype=torch.float64)
    x = ((torch.rand(n_train, 3, generator=gen, dtype=torch.float64) * 2 - 1) * h + c).float().to(dev)
    y = spinor_blocks(shape, x)
    mu, sd = y.mean(0), y.std(0) + 1e-3
    net = new_net(box, mu, sd, seed, dev)
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
    xt = ((torch.rand(65536, 3, generator=gen, dtype=torch.float64) * 2 - 1) * h + c).float().to(dev)
    yt = spinor_blocks(shape, xt)
    net.final_loss = float((((evaluate(net, xt) - yt) / sd) ** 2).mean())
    torch.cuda.synchronize()
    net.train_s = time.perf_counter() - t0
    return net


def evaluate(net: Net, x: torch.Tensor) -> torch.Tensor:
    """(n, N_OUT) de-normalised network outputs (hb_mlp_eval)."""
    lib = hip.load()
    y = torch.empty(len(x), N_OUT, dtype=torch.float32, device=x.device)
    _check(lib.hb_mlp_eval(ctypes.byref(net.struct()), x.data_ptr(), len(x), y.data_ptr(), stream()), lib,
           "hb_mlp_eval")
    return y


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
    """A network trained on this pulse shape whose box holds the input range ``rng`` (the smallest such), or
    None."""
    best = None
    for (k, b), net in _CACHE.items():
        lo, hi = np.array(b[0]) - np.array(b[1]), np.array(b[0]) + np.array(b[1])
        inside = bool(np.all(rng[0] >= lo) and np.all(rng[1] <= hi))
        if k == key and inside and (best is None or np.prod(b[1]) < np.prod(best.box[1])):
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
        mlp = HbMlp(net.params.data_ptr(), net.adam_m.data_ptr(), net.adam_v.data_ptr(), norm.data_ptr(), HIDDEN,
                    N_OUT)
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
