"""The C++ host layer (coolfluid3_amd/host/cf3hip_lss.hpp) driven by a C++ program written like the
reference's LSS unit tests (tests/cpp/test_lss_api.cpp)."""
import os
import subprocess

import pytest

from coolfluid3_amd import build


def host_test_binary(oracle):
    if not os.path.exists(build.HOST_TEST):
        build.build_host_test()
    return build.HOST_TEST


def test_host_layer_exception_mapping_without_device(oracle):
    exe = host_test_binary(oracle)
    env = dict(os.environ, HIP_VISIBLE_DEVICES="")
    out = subprocess.run([exe, "--no-device"], capture_output=True, text=True, timeout=120, env=env)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr


@pytest.mark.gpu
def test_host_layer_on_device(oracle):
    exe = host_test_binary(oracle)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
