"""Builds the HIP library libfideslib_amd.so in-tree (hipcc, gfx950).

Usage: python -m fideslib_amd.build [-j N]
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")


def build(jobs: int = 4, force: bool = False) -> str:
    cmd = ["make", "-C", CSRC, "-j%d" % jobs]
    if force:
        cmd.append("-B")
    subprocess.run(cmd, check=True)
    return os.path.join(HERE, "libfideslib_amd.so")


if __name__ == "__main__":
    j = 4
    if "-j" in sys.argv:
        j = int(sys.argv[sys.argv.index("-j") + 1])
    print(build(j, force="--force" in sys.argv))
