"""hb_signal_reduce_noalloc (the entry point that allocates nothing) against hb_signal_reduce with a workspace, called
through the C ABI: bit for bit on two calls in a row, the second smaller (the fixed device array keeps the first
call's contents, and the second must not read them), and sizes beyond the array refused (GPU)."""

import ctypes

import pytest
import torch

from blochparty_amd.engines import hip

pytestmark = pytest.mark.gpu
NOALLOC_DOUBLES = 1 << 21  # HB_NOALLOC_SCRATCH_DOUBLES of the header


def stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def both(lib, n_blocks, rows, n_tiles, n_coils, row0, n_adc, seed):
    g = torch.Generator(device="cpu").manual_seed(seed)
    partial = torch.randn(n_blocks * rows * 16 * n_tiles, generator=g).cuda()
    psi = (torch.rand(n_adc, generator=g, dtype=torch.float64) * 6.3).cuda()
    work = torch.empty(lib.hb_signal_reduce_workspace(n_blocks, rows, n_tiles), dtype=torch.float64, device="cuda")
    with_ws = torch.full((n_adc, n_coils, 2), float("nan"), dtype=torch.float64, device="cuda")
    no_ws = with_ws.clone()
    args = (partial.data_ptr(), n_blocks, rows, n_tiles, n_coils, psi.data_ptr(), row0, None)
    assert lib.hb_signal_reduce(*args, work.data_ptr(), with_ws.data_ptr(), stream()) == 0
    assert lib.hb_signal_reduce_noalloc(*args, no_ws.data_ptr(), stream()) == 0
    torch.cuda.synchronize()
    return with_ws.cpu(), no_ws.cpu()


def test_noalloc_matches_the_workspace_entry_point_twice(gpu):
    lib = hip.load()
    print("noalloc call 1: 4096 blocks x 512 rows x 1 tile, 8 coils", flush=True)
    a, b = both(lib, 4096, 512, 1, 8, 0, 512, 0)
    assert torch.equal(a, b) and torch.isfinite(a).all()
    print("noalloc call 2: 300 blocks x 37 rows x 2 tiles, 11 coils, row0 5", flush=True)
    a, b = both(lib, 300, 37, 2, 11, 5, 50, 1)
    assert torch.equal(a[5:42], b[5:42]) and torch.isfinite(a[5:42]).all()
    assert torch.isnan(b[:5]).all() and torch.isnan(b[42:]).all()  # rows outside [row0, row0 + rows) untouched


def test_noalloc_refuses_sizes_beyond_its_array(gpu):
    lib = hip.load()
    rows = NOALLOC_DOUBLES // (16 * 16) + 1  # 16 chunks of 256 blocks, one tile: one row past the array
    assert lib.hb_signal_reduce_workspace(4096, rows, 1) > NOALLOC_DOUBLES
    rc = lib.hb_signal_reduce_noalloc(None, 4096, rows, 1, 8, None, 0, None, None, stream())
    assert rc != 0 and "invalid" in lib.hb_error_string(rc).decode().lower()
