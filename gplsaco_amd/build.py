"""Build libfaco_hip.so in-tree for gfx950 with hipcc (no JIT, no site-packages)."""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent


def build(jobs: int | None = None, force: bool = True) -> Path:
    """Compile every HIP/C++ source of the engine; returns the library path."""
    jobs = jobs or min(8, os.cpu_count() or 1)
    if force:
        subprocess.run(["make", "-s", "-C", str(PACKAGE), "clean"], check=True)
    subprocess.run(["make", "-s", "-C", str(PACKAGE), f"-j{jobs}", "all"], check=True)
    lib = PACKAGE / "libfaco_hip.so"
    if not lib.exists():
        raise RuntimeError("libfaco_hip.so was not produced")
    return lib


if __name__ == "__main__":
    print(build())
