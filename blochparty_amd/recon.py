"""Downstream of ``SimResult.signal``: Cartesian k-space gridding, centred FFTs, root-sum-of-squares and coil
combination, the k-space trajectory of a sequence, and the raw data as ISMRMRD acquisitions.

Reference interface: ``src/blochparty/recon.py:13-60`` (``kspace_grid``, ``ifftnc``, ``reconstruct``, ``rss``,
``coil_combine``, ``fft2c``), ``sequence.py:173-217`` (``flip_angle``, ``kspace_trajectory``) and
``io/ismrmrd.py`` (one acquisition per readout, its trajectory and encode indices). On the GPU, complex64 k-space with
power-of-two axes takes the HIP kernels (``csrc/recon.hip``: one centred inverse DFT pass per axis, and the RSS);
other shapes and dtypes, and CPU tensors, take torch's FFT as the reference does. Writing an ISMRMRD HDF5 file needs
``h5py`` (absent here, as in the reference without its ``io`` extra: ``ImportError``); the acquisitions are built
without it.
"""

from __future__ import annotations

import json
import math

import numpy as np
import torch
from torch import Tensor

from .sequence import Sequence


def kspace_grid(signal: Tensor, seq: Sequence) -> Tensor:
    """(C, n_adc) samples on the sequence's Cartesian grid, k = 0 at index n // 2 of each axis: a GRE's
    ``kspace_shape`` (samples in grid order), or a TSE's ``matrix`` with ``kspace_index`` (each sample's flat grid
    index, in acquisition order)."""
    meta = seq.meta
    if "kspace_shape" in meta:
        return signal.reshape(signal.shape[0], *meta["kspace_shape"])
    if "matrix" in meta and "kspace_index" in meta:
        grid = torch.zeros(signal.shape[0], int(np.prod(meta["matrix"])), dtype=signal.dtype, device=signal.device)
        grid[:, torch.as_tensor(meta["kspace_index"], device=signal.device)] = signal
        return grid.reshape(signal.shape[0], *meta["matrix"])
    raise ValueError("the sequence's meta names no Cartesian grid (kspace_shape, or matrix and kspace_index)")


def _on_hip(t: Tensor, dtype: torch.dtype) -> bool:
    return t.is_cuda and t.dtype == dtype


def _pow2(n: int) -> bool:
    return 2 <= n <= 4096 and n & (n - 1) == 0


def ifftnc(k: Tensor, ndim: int) -> Tensor:
    """Centred inverse DFT over the last ``ndim`` axes, backward normalisation: a spin of magnetization m at a
    voxel centre gives that voxel the value m. complex64 on the GPU with power-of-two axes (2..4096): one HIP pass
    per axis (``hb_ifftc_axis``), the first out of place; otherwise torch's FFT on the tensor's device."""
    axes = tuple(range(-ndim, 0))
    if _on_hip(k, torch.complex64) and 1 <= ndim <= k.ndim and all(_pow2(k.shape[a]) for a in axes):
        from .engines import hip

        lib = hip.HIP.lib()
        src = k.contiguous()
        out = torch.empty_like(src)
        shape = list(src.shape)
        for i, a in enumerate(range(src.ndim - ndim, src.ndim)):
            lib.ifftc_axis((src if i == 0 else out).data_ptr(), out.data_ptr(), math.prod(shape[:a]), shape[a],
                           math.prod(shape[a + 1 :]))
        return out
    return torch.fft.fftshift(torch.fft.ifftn(torch.fft.ifftshift(k, dim=axes), dim=axes), dim=axes)


def ifft2c(k: Tensor) -> Tensor:
    return ifftnc(k, 2)


def ifft3c(k: Tensor) -> Tensor:
    return ifftnc(k, 3)


def fft2c(x: Tensor) -> Tensor:
    """Centred orthonormal 2D FFT over the last two axes (``ifft2c_ortho`` is its inverse and adjoint)."""
    axes = (-2, -1)
    return torch.fft.fftshift(torch.fft.fft2(torch.fft.ifftshift(x, dim=axes), norm="ortho"), dim=axes)


def ifft2c_ortho(k: Tensor) -> Tensor:
    axes = (-2, -1)
    return torch.fft.fftshift(torch.fft.ifft2(torch.fft.ifftshift(k, dim=axes), norm="ortho"), dim=axes)


def reconstruct(signal: Tensor, seq: Sequence) -> Tensor:
    """(C, n_adc) Cartesian signal -> (C, *grid) coil images."""
    k = kspace_grid(signal, seq)
    return ifftnc(k, k.ndim - 1)


def rss(images: Tensor, dim: int = 0) -> Tensor:
    """Root-sum-of-squares over the coil axis (complex64 on the GPU over the leading axis: ``hb_rss``, one read of
    the images)."""
    if _on_hip(images, torch.complex64) and images.ndim >= 1 and dim in (0, -images.ndim) and images.numel():
        from .engines import hip

        src = images.contiguous()
        out = torch.empty(src.shape[1:], dtype=torch.float32, device=src.device)
        hip.HIP.lib().rss(src.data_ptr(), src.shape[0], out.numel(), out.data_ptr())
        return out
    return (images.abs() ** 2).sum(dim=dim).sqrt()


def coil_combine(images: Tensor, sens: Tensor) -> Tensor:
    """Sensitivity-weighted combination sum_c conj(S_c) I_c / sum_c |S_c|^2 of (C, ...) coil images (0 where no coil
    sees the voxel)."""
    num = (sens.conj() * images).sum(dim=0)
    den = (sens.abs() ** 2).sum(dim=0)
    return torch.where(den > 0, num / den.clamp_min(1e-30), torch.zeros_like(num))


def flip_angle(signal: np.ndarray, dt) -> float:
    """Flip angle [deg] of a pulse (complex samples [Hz], steps ``dt`` [s]) for a spin at the pulse's own frequency:
    the frequency is the mean phase step between samples weighted by the samples' magnitudes (so the pi steps at a
    sinc's zero crossings, between near-zero samples, barely count), and the angle is that of the composed rotation
    of the segments for a spin at that offset, acos of the final Mz from equilibrium."""
    s = np.asarray(signal, dtype=complex)
    if s.size == 0:
        return 0.0
    dt = np.broadcast_to(np.asarray(dt, dtype=np.float64), s.shape)
    f = 0.0
    if s.size > 1:
        pair = s[1:] * np.conj(s[:-1])
        w = np.abs(pair)
        if w.sum() > 0:
            f = float(np.sum(w * np.angle(pair) / (np.pi * (dt[1:] + dt[:-1]))) / w.sum())
    m = np.array([0.0, 0.0, 1.0])
    for z, h in zip(s, dt):  # rotation about -2 pi h (Re z, Im z, f)
        v = -2 * np.pi * h * np.array([z.real, z.imag, f])
        phi = float(np.linalg.norm(v))
        if phi == 0.0:
            continue
        a = v / phi
        m = m * math.cos(phi) + np.cross(a, m) * math.sin(phi) + a * np.dot(a, m) * (1 - math.cos(phi))
    return math.degrees(math.acos(max(-1.0, min(1.0, m[2]))))


def kspace_trajectory(seq: Sequence) -> Tensor:
    """(n_adc, 3) k-space [1/m] at every ADC sample: the gradient moment since the last excitation's centre, negated
    at each refocusing pulse's centre (a pulse is a run of RF segments centred at its middle; an excitation flips
    less than 90.01 deg by ``flip_angle``); NaN before the first excitation."""
    dt = seq.dt.detach().cpu().double().numpy()
    tb = np.concatenate([[0.0], np.cumsum(dt)])
    mom = np.concatenate([np.zeros((1, 3)), np.cumsum(seq.grad.detach().cpu().double().numpy() * dt[:, None], 0)])
    moment_at = lambda t: np.stack([np.interp(t, tb, mom[:, a]) for a in range(3)], -1)
    on = np.asarray(seq.is_rf, dtype=np.int8)
    edges = np.flatnonzero(np.diff(np.concatenate([[0], on, [0]])))
    first, last = edges[::2], edges[1::2]
    rf = seq.rf.detach().cpu().double().numpy()
    b1 = rf[:, 0] + 1j * rf[:, 1]
    centre = 0.5 * (tb[first] + tb[last])
    k_centre = np.full((len(first), 3), np.nan)  # k just after each pulse's centre
    k = None
    for p, (a, b) in enumerate(zip(first, last)):
        if flip_angle(b1[a:b], dt[a:b]) < 90.01:
            k = np.zeros(3)
        elif k is not None:
            k = -(k + moment_at(centre[p]) - moment_at(centre[p - 1]))
        if k is not None:
            k_centre[p] = k
    t = tb[1:][np.asarray(seq.adc, dtype=bool)]
    p = np.searchsorted(centre, t, side="right") - 1
    out = np.full((len(t), 3), np.nan)
    ok = p >= 0
    out[ok] = k_centre[p[ok]] + moment_at(t[ok]) - moment_at(centre[p[ok]])
    return torch.from_numpy(out)


# ---- ISMRMRD raw data


def acquisitions(signal: Tensor, seq: Sequence) -> list[dict]:
    """One ISMRMRD acquisition per run of consecutive ADC segments: ``data`` (C, samples) complex64, ``traj``
    (samples, 3) k-space [1/m], ``sample_time_us``, ``scan_counter``, and the Cartesian encode indices
    (``kspace_encode_step_1`` / ``_2``: the line's phase-encode indices) where the grid is known."""
    adc = np.asarray(seq.adc, dtype=bool)
    edges = np.flatnonzero(np.diff(np.concatenate([[0], adc.astype(np.int8), [0]])))
    runs = list(zip(edges[::2], edges[1::2]))
    row = np.cumsum(adc) - 1
    traj = kspace_trajectory(seq).numpy()
    dt = seq.dt.detach().cpu().double().numpy()
    sig = signal.detach().cpu().to(torch.complex64).numpy()
    shape = seq.meta.get("kspace_shape")
    index = seq.meta.get("kspace_index")
    out = []
    for n, (a, b) in enumerate(runs):
        rows = row[a:b]
        acq = {"scan_counter": n, "number_of_samples": int(b - a), "active_channels": int(sig.shape[0]),
               "sample_time_us": float(np.median(dt[a + 1 : b])) * 1e6 if b - a > 1 else float(dt[a]) * 1e6,
               "data": sig[:, rows], "traj": traj[rows]}
        flat = None
        if shape is not None:
            flat = int(rows[0])
            dims = tuple(shape)
        elif index is not None:
            flat = int(np.asarray(index)[rows[0]])
            dims = tuple(seq.meta["matrix"])
        if flat is not None:
            line = np.unravel_index(flat, dims)[:-1]  # the phase-encode indices of the line's first sample
            acq["kspace_encode_step_1"] = int(line[-1])
            acq["kspace_encode_step_2"] = int(line[-2]) if len(line) > 1 else 0
        out.append(acq)
    return out


def ismrmrd_header(seq: Sequence, n_coils: int, field_strength: float = 3.0) -> str:
    """A minimal ISMRMRD XML header: field strength, encoded matrix and field of view, TR / TE, channels."""
    matrix = seq.meta.get("kspace_shape") or seq.meta.get("matrix") or (seq.n_adc,)
    m = list(matrix)[::-1] + [1] * (3 - len(matrix))
    fov_mm = 1e3 * float(seq.meta.get("fov", 0.0))
    tr = 1e3 * float(seq.meta.get("tr", 0.0)) if not isinstance(seq.meta.get("tr"), Tensor) else 1e3 * float(
        seq.meta["tr"].detach())
    te = 1e3 * float(seq.meta.get("te", 0.0))
    return (f'<?xml version="1.0"?><ismrmrdHeader xmlns="http://www.ismrm.org/ISMRMRD">'
            f"<acquisitionSystemInformation><systemFieldStrength_T>{field_strength}</systemFieldStrength_T>"
            f"<receiverChannels>{n_coils}</receiverChannels></acquisitionSystemInformation>"
            f"<encoding><encodedSpace><matrixSize><x>{m[0]}</x><y>{m[1]}</y><z>{m[2]}</z></matrixSize>"
            f"<fieldOfView_mm><x>{fov_mm}</x><y>{fov_mm}</y><z>{fov_mm}</z></fieldOfView_mm></encodedSpace>"
            f"<trajectory>cartesian</trajectory></encoding>"
            f"<sequenceParameters><TR>{tr}</TR><TE>{te}</TE></sequenceParameters></ismrmrdHeader>")


def write_dataset(path, signal: Tensor, seq: Sequence, field_strength: float = 3.0) -> None:
    """The acquisitions and header as an ISMRMRD HDF5 file (group ``/dataset``: ``xml``, and per acquisition its
    data, trajectory and header fields as JSON). Needs ``h5py``."""
    try:
        import h5py
    except ImportError as e:
        raise ImportError("write_dataset needs h5py (not installed here)") from e
    acqs = acquisitions(signal, seq)
    with h5py.File(path, "w") as f:
        g = f.create_group("dataset")
        g.create_dataset("xml", data=ismrmrd_header(seq, signal.shape[0], field_strength))
        for a in acqs:
            sub = g.create_group(f"acquisition_{a['scan_counter']:06d}")
            sub.create_dataset("data", data=a["data"])
            sub.create_dataset("traj", data=a["traj"])
            sub.attrs["head"] = json.dumps({k: v for k, v in a.items() if k not in ("data", "traj")})


__all__ = ["acquisitions", "coil_combine", "fft2c", "flip_angle", "ifft2c", "ifft2c_ortho", "ifft3c", "ifftnc",
           "ismrmrd_header", "kspace_grid", "kspace_trajectory", "reconstruct", "rss", "write_dataset"]
