"""Ahead-of-time build of the HIP library ``liblbm_hip.so`` for gfx950.

Compiles every ``csrc/*.hip`` file with hipcc (one process per file, in
parallel) and links them into ``lbm/liblbm_hip.so``. Flags:

* ``-ffp-contract=off``: no fused multiply-adds, so every node update rounds
  exactly like the NumPy oracle's separate multiplies and adds;
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32 ops;
* ``-fvisibility=hidden``: only the ``LBM_API`` entry points are exported.

Usage: ``python -m lbm._build`` (with the package directory on ``sys.path``).
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent
CSRC = PACKAGE.parent / "csrc"
INCLUDE = PACKAGE.parent.parent / "include"
LIBRARY = PACKAGE / "liblbm_hip.so"
ARCH = "gfx950"

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fvisibility=hidden",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]  # fmt: skip


def hipcc() -> str:
    for candidate in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if candidate and Path(candidate).exists():
            return candidate
    raise RuntimeError("hipcc not found; install ROCm or set HIPCC")


def _compile(source: Path, objdir: Path) -> Path:
    obj = objdir / (source.stem + ".o")
    cmd = [hipcc(), *FLAGS, f"-I{INCLUDE}", "-c", str(source), "-o", str(obj)]
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode != 0:
        raise RuntimeError(f"hipcc failed on {source.name}:\n{result.stderr}")
    return obj


def build(verbose: bool = True) -> Path:
    """Compile and link the library (always rebuilds). Returns its path."""
    sources = sorted(CSRC.glob("*.hip"))
    objdir = CSRC / "_build"
    objdir.mkdir(exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(sources))) as pool:
        objects = list(pool.map(lambda s: _compile(s, objdir), sources))
    tmp = LIBRARY.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objects), "-o", str(tmp)]
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode != 0:
        raise RuntimeError(f"link failed:\n{result.stderr}")
    os.replace(tmp, LIBRARY)
    if verbose:
        print(f"built {LIBRARY} from {len(sources)} sources")
    return LIBRARY


if __name__ == "__main__":
    build()
    sys.exit(0)
