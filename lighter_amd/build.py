"""Build helpers: the product library (hipcc, gfx950) and the CPU checkers."""
from __future__ import annotations

import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "lighter_amd", "csrc")
ORACLE = os.path.join(ROOT, "oracle")


def _make(path, *args):
    jobs = str(min(16, os.cpu_count() or 1))
    cmd = ["make", "-C", path, "-j" + jobs, *args]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"{' '.join(cmd)} failed:\n{res.stdout[-4000:]}\n{res.stderr[-4000:]}")


def build_library(force: bool = True) -> str:
    """Compile liblighter_amd.so for gfx950 in-tree."""
    _make(CSRC, *(["-B"] if force else []))
    return os.path.join(ROOT, "lighter_amd", "liblighter_amd.so")


def build_checkers(force: bool = True) -> None:
    """Compile the C oracle and, where the reference checkout exists, the
    reference build used as the second checker."""
    _make(ORACLE, *(["-B"] if force else []), "all")
