"""The full variant's float64 reverse mode (blochparty_amd/csrc/full_device.h: the bound pool's exchange, the
lineshapes, concomitant fields, the rotation, bound-only RF), compiled for the host from the same source as the
kernel, against Richardson-extrapolated central differences of its own forward (tests/full_device_fd.cpp): every
spin parameter, the initial state and every segment parameter, over 12 configurations. A dropped or mis-signed term
shows at 1e-2 or more; differences in these double-precision sweeps agree to a few 1e-5."""

import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs a host C++ compiler")
def test_full_variant_reverse_mode_against_finite_differences(tmp_path):
    exe = tmp_path / "full_device_fd"
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", str(exe), str(HERE / "full_device_fd.cpp")], check=True)
    res = subprocess.run([str(exe), "5e-4"], capture_output=True, text=True, timeout=120)
    print(res.stdout)
    assert res.returncode == 0, res.stdout
    worst = float(res.stdout.strip().splitlines()[-1].split()[1])
    assert worst < 5e-4
