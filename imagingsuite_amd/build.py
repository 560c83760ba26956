"""In-tree build of the native libraries (gfx950).

``build()`` always recompiles:
  * imagingsuite_amd/lib/libmuhrec_hip.so  -- HIP kernels + C ABI (hipcc, gfx950)
  * imagingsuite_amd/lib/*.so module plugins (when their sources exist)
"""
from __future__ import annotations

import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))


def _jobs() -> str:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 4
    return str(max(1, min(16, n)))


def build(verbose: bool = False) -> None:
    csrc = os.path.join(_HERE, "csrc")
    out = None if verbose else subprocess.DEVNULL
    subprocess.run(["make", "-C", csrc, "-B", "-j", _jobs()], check=True, stdout=out)
    mods = os.path.join(csrc, "modules")
    if os.path.exists(os.path.join(mods, "Makefile")):
        subprocess.run(["make", "-C", mods, "-B", "-j", _jobs()], check=True, stdout=out)


if __name__ == "__main__":
    build(verbose=True)
