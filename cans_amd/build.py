"""Builds libcans_hip.so for gfx950 in-tree (hipcc via cans_amd/csrc/Makefile)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def build(jobs=4, force=True):
    csrc = os.path.join(HERE, "csrc")
    if force:
        subprocess.check_call(["make", "-s", "-C", csrc, "clean"])
    subprocess.check_call(["make", "-s", "-j%d" % jobs, "-C", csrc])
    return os.path.join(HERE, "libcans_hip.so")


if __name__ == "__main__":
    print(build())
