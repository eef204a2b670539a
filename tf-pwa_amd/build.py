"""Build libtfpwa_amd.so in-tree for gfx950 (hipcc: the device code and C ABI in tpa_capi.hip,
the host serving runtime in tpa_serve.cpp, the single-fit session in tpa_fcn.cpp)."""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = [os.path.join(HERE, "csrc", n) for n in ("tpa_capi.hip", "tpa_serve.cpp", "tpa_fcn.cpp")]
OUT = os.path.join(HERE, "lib", "libtfpwa_amd.so")

FLAGS = [
    "-O3", "--offload-arch=gfx950", "-fPIC", "-shared", "-std=c++17", "-pthread",
    # packed FP32 instructions stay off in every HIP object of this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(HERE, "csrc"),
]


def build(verbose=False):
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = OUT + ".tmp"
    cmd = [hipcc] + FLAGS + ["-o", tmp] + SRC
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
