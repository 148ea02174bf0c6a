"""Builds the HIP library in-tree: every ``csrc/*.hip`` into ``blochparty_amd/libblochparty_amd.so`` for gfx950.

``python -m blochparty_amd.build`` (or ``__graft_entry__.build()``). hipcc cross-compiles without a GPU.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "libblochparty_amd.so"
ARCH = os.environ.get("BLOCHPARTY_AMD_ARCH", "gfx950")

FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    f"--offload-arch={ARCH}",
    # no packed FP32 instructions (v_pk_*_f32): keep every float op a plain VALU op
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is needed to build blochparty_amd)")


OBJ = HERE.parent / "_build"


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")


def build(verbose: bool = False) -> Path:
    """Compile the library (always: every source to an object, in parallel, then one link) and return its path."""
    sources = sorted(CSRC.glob("*.hip"))
    OBJ.mkdir(exist_ok=True)
    compile_flags = [f for f in FLAGS if f != "-shared"]
    objects = [OBJ / (src.stem + ".o") for src in sources]
    jobs = max(1, min(len(sources), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(jobs) as pool:
        for f in [pool.submit(_run, [hipcc(), *compile_flags, "-c", "-I", str(HERE.parent / "include"), str(src),
                                     "-o", str(obj)], verbose) for src, obj in zip(sources, objects)]:
            f.result()
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc(), *FLAGS, *map(str, objects), "-o", str(tmp)], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
