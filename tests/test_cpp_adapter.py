"""The header-only C++ adapter (include/xyz_hip.hpp) compiles against the C ABI
and its host-only part runs; the GPU part runs on a device."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "adapter_example.cpp")
BIN = os.path.join(ROOT, "xyz_amd", "adapter_example")


def compile_example(out):
    subprocess.check_call([
        "g++", "-std=c++17", "-O1", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include"),
        "-I", "/opt/rocm/include", SRC, "-o", out, "-L", os.path.join(ROOT, "xyz_amd"),
        "-lxyz_hip", "-L", "/opt/rocm/lib", "-lamdhip64",
        "-Wl,-rpath," + os.path.join(ROOT, "xyz_amd"), "-Wl,-rpath,/opt/rocm/lib"])


def test_adapter_compiles_and_builds_tree(capi_mod, tmp_path):
    out = str(tmp_path / "adapter_example")
    compile_example(out)
    r = subprocess.run([out, "host"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "refs" in r.stdout


@pytest.mark.gpu
def test_adapter_on_device(capi_mod, device):
    if not os.path.exists(BIN):
        pytest.skip("adapter example not built (run __graft_entry__.build())")
    r = subprocess.run([BIN, "gpu"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "film sum" in r.stdout
