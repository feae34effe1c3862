"""Build the native libraries in-tree (hipcc for gfx950, g++ for the host)."""
from __future__ import annotations

import os
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)


def _make(directory: str, jobs: int = 8) -> None:
    subprocess.run(["make", "-C", directory, f"-j{jobs}"], check=True)


def build_native(jobs: int = 8) -> None:
    """Compile libmxc_gpu.so, libmxc_h2.so and lorasp into mxc_amd/lib."""
    _make(os.path.join(PKG, "csrc"), jobs)


def build_oracle(jobs: int = 8) -> None:
    """Compile the CPU oracle (test infrastructure) into oracle/_build."""
    _make(os.path.join(ROOT, "oracle"), jobs)
