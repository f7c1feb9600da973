"""Build recipe of the native parts: libcf3hip.so (gfx950 HIP) and the C++ host-layer test programs."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def build_native(jobs=8):
    """Compile every HIP translation unit for gfx950 and link libcf3hip.so (always rebuilds)."""
    csrc = os.path.join(HERE, "csrc")
    subprocess.check_call(["make", "-C", csrc, "clean"], stdout=subprocess.DEVNULL)
    subprocess.check_call(["make", "-C", csrc, "-j%d" % jobs])
    return os.path.join(HERE, "libcf3hip.so")


def build_oracle():
    subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle"), "-B"])
