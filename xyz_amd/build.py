"""Compiles the gfx950 library xyz_amd/libxyz_hip.so in-tree with hipcc."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libxyz_hip.so")

SOURCES = ["xyz_hip.hip", "render.hip", "kdtree_gpu.hip", "kdtree_build.cpp", "mqo_loader.cpp", "importon_map.hip"]

# IEEE binary32 as the reference's source semantics: no FMA contraction,
# correctly rounded division/sqrt (hipcc's default), denormals kept.  Packed
# FP32 instructions are disabled for all device code.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math", "-fno-gpu-flush-denormals-to-zero",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.sep not in c or os.path.exists(c)):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose=False, out=None, defines=(), tmp_dir=None):
    """Compiles SOURCES into `out` (default xyz_amd/libxyz_hip.so)."""
    objs = []
    tmp_dir = tmp_dir or os.path.join(HERE, "_build")
    os.makedirs(tmp_dir, exist_ok=True)
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(tmp_dir, src + ".o")
        cmd = ([hipcc()] + [f for f in HIPCC_FLAGS if f != "-shared"] + list(defines)
               + ["-c", path, "-o", obj])
        if src.endswith(".cpp"):
            cmd = [hipcc(), "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
                   "-Wall", "-x", "c++", "-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
        objs.append(obj)
    target = out or LIB
    tmp = target + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, target)
    return target


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
