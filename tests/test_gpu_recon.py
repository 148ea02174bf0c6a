"""The HIP reconstruction kernels (``csrc/recon.hip`` through ``blochparty_amd.recon``) against numpy in float64:
the centred inverse DFT over 1, 2 and 3 axes (the reference's recon.py:17-30 ``ifftnc``: fftshift(ifftn(ifftshift)))
on shapes with partial tiles in both layouts, the root-sum-of-squares, ``reconstruct`` of a GRE's signal, and the
C ABI's refusals. Float32 transforms of up to 2^21 points: the bound is 1e-5 of the largest value."""

import math

import numpy as np
import pytest
import torch

import blochparty_amd as bp
from blochparty_amd import recon

pytestmark = pytest.mark.gpu

TOL = 1e-5


def numpy_ifftnc(k: np.ndarray, ndim: int) -> np.ndarray:
    axes = tuple(range(k.ndim - ndim, k.ndim))
    return np.fft.fftshift(np.fft.ifftn(np.fft.ifftshift(k, axes=axes), axes=axes), axes=axes)


def seeded(shape, seed=0) -> torch.Tensor:
    g = torch.Generator().manual_seed(seed)
    return torch.randn(*shape, dtype=torch.complex64, generator=g)


@pytest.mark.parametrize("shape, ndim", [
    ((3, 8, 16, 32), 3),     # 3D, rows of 32 and strided axes
    ((2, 64, 64), 2),        # the 2D GRE's grid
    ((4, 128), 1),
    ((7, 1024), 1),          # rows of 1024: four lines per tile, the last tile partial
    ((2, 128, 64), 2),       # the first axis 128 long, 64 inner: two tiles of 32 lines per outer index
    ((1, 2, 4096), 1),       # the longest axis: one line per tile
    ((5, 2, 2, 2), 3),       # the shortest axes
    ((16, 4, 8, 16), 3),
])
def test_ifftnc_matches_numpy(gpu, shape, ndim):
    k = seeded(shape, seed=sum(shape))
    got = recon.ifftnc(k.cuda(), ndim).cpu()
    ref = numpy_ifftnc(k.numpy().astype(np.complex128), ndim)
    assert got.dtype == torch.complex64 and got.shape == k.shape
    err = np.abs(got.numpy().astype(np.complex128) - ref).max() / np.abs(ref).max()
    assert err < TOL, err


def test_strided_partial_tiles_and_in_place_passes(gpu):
    """(outer, L, inner) = (3, 128, 48) through the C ABI directly: tiles of 32 lines, every second one 16 lines
    full."""
    x = seeded((3, 128, 48), seed=5)
    xd = x.cuda()
    out = torch.empty_like(xd)
    gpu.lib().ifftc_axis(xd.data_ptr(), out.data_ptr(), 3, 128, 48)
    ref = np.fft.fftshift(np.fft.ifft(np.fft.ifftshift(x.numpy().astype(np.complex128), axes=1), axis=1), axes=1)
    err = np.abs(out.cpu().numpy() - ref).max() / np.abs(ref).max()
    assert err < TOL, err
    # in place gives the same bits
    gpu.lib().ifftc_axis(xd.data_ptr(), xd.data_ptr(), 3, 128, 48)
    assert torch.equal(xd.cpu(), out.cpu())


def test_other_shapes_and_dtypes_take_torch(gpu):
    k = seeded((2, 6, 8), seed=3).cuda()  # 6 is no power of two
    ref = torch.fft.fftshift(torch.fft.ifftn(torch.fft.ifftshift(k, dim=(-2, -1)), dim=(-2, -1)), dim=(-2, -1))
    assert torch.equal(recon.ifftnc(k, 2), ref)
    k128 = k.to(torch.complex128)
    assert recon.ifftnc(k128, 2).dtype == torch.complex128


def test_rss_matches_numpy(gpu):
    x = seeded((8, 16, 32, 4), seed=9)
    got = recon.rss(x.cuda()).cpu()
    ref = np.sqrt((np.abs(x.numpy().astype(np.complex128)) ** 2).sum(axis=0))
    assert got.dtype == torch.float32 and got.shape == x.shape[1:]
    assert np.abs(got.numpy() - ref).max() / ref.max() < 1e-6
    # another axis: torch's
    assert torch.allclose(recon.rss(x.cuda(), dim=1).cpu(), recon.rss(x, dim=1), rtol=1e-6, atol=1e-6)


def test_reconstruct_gre_signal_on_the_gpu(gpu):
    """A 32^2 spoiled GRE simulated on the GPU, reconstructed on the GPU: its coil-combined image against the
    steady-state closed form, and the RSS against torch's float64 reconstruction of the same signal on the CPU."""
    n, fov, flip, tr = 32, 0.22, 20.0, 10e-3
    g = torch.Generator().manual_seed(4)
    u = lambda lo, hi: lo + (hi - lo) * torch.rand(n, n, generator=g, dtype=torch.float64)
    pd, t1, t2 = u(0.5, 1.0), u(0.15, 0.4), u(0.05, 0.2)
    ph = bp.phantom_from_maps(pd, t1, t2, fov)
    seq = bp.gre_2d(n, fov, flip_deg=flip, tr=tr, te=2e-3, dwell=1e-6, n_dummy=300)
    coils = bp.birdcage_sensitivities(ph.pos, 4)
    sig = bp.simulate(ph, seq, coils, engine="gpu").signal
    images = recon.reconstruct(sig.cuda(), seq)
    assert images.is_cuda and images.dtype == torch.complex64
    sens = coils.T.reshape(4, n, n).to(torch.complex64).cuda()
    img = recon.coil_combine(images, sens).cpu().to(torch.complex128)
    e1 = torch.exp(-tr / t1)
    a = math.radians(flip)
    expected = 1j * pd * math.sin(a) * (1 - e1) / (1 - math.cos(a) * e1) * torch.exp(-float(seq.meta["te"]) / t2)
    assert float((img - expected).abs().max() / expected.abs().max()) < 5e-4
    rss_gpu = recon.rss(images).cpu().double()
    rss_cpu = recon.rss(recon.reconstruct(sig.to(torch.complex128), seq))
    assert float((rss_gpu - rss_cpu).abs().max() / rss_cpu.abs().max()) < TOL


def test_refusals(gpu):
    x = torch.zeros(4, 6, dtype=torch.complex64, device="cuda")
    for L in (1, 3, 6, 8192):
        with pytest.raises(RuntimeError):
            gpu.lib().ifftc_axis(x.data_ptr(), x.data_ptr(), 1, L, 1)
    with pytest.raises(RuntimeError):
        gpu.lib().ifftc_axis(x.data_ptr(), x.data_ptr(), 1, 4, 0)
    with pytest.raises(RuntimeError):
        gpu.lib().rss(x.data_ptr(), 0, 6, x.data_ptr())
    assert math.isfinite(float(recon.rss(x).sum()))
