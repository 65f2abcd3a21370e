"""In-tree build of the gfx950 library, the pybind11 module and the CPU
oracle (make; hipcc for HIP sources, g++/gcc for host-only sources)."""

from __future__ import annotations

import os
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)


def build_all(force: bool = True, jobs: int | None = None) -> None:
    jobs = jobs or min(16, os.cpu_count() or 4)
    flag = ["-B"] if force else []
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")] + flag, check=True)
    subprocess.run(["make", "-C", os.path.join(PKG, "csrc"), f"-j{jobs}"] + flag, check=True)


if __name__ == "__main__":
    build_all()
