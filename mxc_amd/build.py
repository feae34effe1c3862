"""Build the native libraries in-tree (hipcc for gfx950, g++ for the host)."""
from __future__ import annotations

import os
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)


def _make(directory: str, jobs: int = 8, force: bool = False) -> None:
    cmd = ["make", "-C", directory, f"-j{jobs}"] + (["-B"] if force else [])
    subprocess.run(cmd, check=True)


def build_native(jobs: int = 8, force: bool = False) -> None:
    """Compile libmxc_gpu.so, libmxc_h2.so and lorasp into mxc_amd/lib
    (everything again when force is set, whatever is already built)."""
    _make(os.path.join(PKG, "csrc"), jobs, force)


def build_oracle(jobs: int = 8, force: bool = False) -> None:
    """Compile the CPU oracle (test infrastructure) into oracle/_build."""
    _make(os.path.join(ROOT, "oracle"), jobs, force)
