"""Build libpyqcm_ed.so in-tree with hipcc for gfx950 (recipe: csrc/Makefile)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def build(force=True, jobs=8):
    args = ["make", "-C", os.path.join(HERE, "csrc"), "-j%d" % jobs]
    if force:
        args.insert(1, "-B")
    subprocess.check_call(args)
    return os.path.join(HERE, "libpyqcm_ed.so")


if __name__ == "__main__":
    print(build())
