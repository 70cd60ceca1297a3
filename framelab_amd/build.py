"""Build step for the native library (libframelab_hip.so, gfx950)."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def build(verbose: bool = False) -> str:
    """Compile every HIP source of the package into framelab_amd/libframelab_hip.so.

    Always recompiles (``make -B``) so a stale library is never picked up; the
    sources compile in parallel, one object each."""
    jobs = str(max(1, min(8, os.cpu_count() or 1)))
    cmd = ["make", "-B", "-j", jobs, "-C", os.path.join(HERE, "csrc")]
    if not verbose:
        cmd.insert(1, "-s")
    subprocess.run(cmd, check=True)
    return os.path.join(HERE, "libframelab_hip.so")
