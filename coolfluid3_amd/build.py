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


HOST_TEST = os.path.join(HERE, "host", "test_lss_api")


def build_host_test():
    """The C++ host-layer test program (plain g++ over the C ABI; links libcf3hip.so and the oracle)."""
    src = os.path.join(ROOT, "tests", "cpp", "test_lss_api.cpp")
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-Wall", src, "-o", HOST_TEST,
                           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(HERE, "host"),
                           "-I" + os.path.join(ROOT, "oracle"),
                           "-L" + HERE, "-lcf3hip", "-L" + os.path.join(ROOT, "oracle", "build"), "-loracle",
                           "-Wl,-rpath,$ORIGIN/..:$ORIGIN/../../oracle/build"])
    return HOST_TEST
