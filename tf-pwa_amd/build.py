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
LIBS = ["-lhiprtc", "-Wl,-rpath,/opt/rocm/lib"]


# Device headers the fused pass's runtime compile (hipRTC) includes, embedded as strings.
RTC_HEADERS = ("tpa_model.h", "tpa_math.h")
RTC_INC = os.path.join(HERE, "csrc", "tpa_rtc_headers.inc")


def write_rtc_headers():
    parts = []
    for name in RTC_HEADERS:
        text = open(os.path.join(HERE, "csrc", name)).read()
        assert ")TPA_RTC\"" not in text
        parts.append('R"TPA_RTC(' + text + ')TPA_RTC"')
    body = ("// written by build.py from " + ", ".join(RTC_HEADERS) + "\n"
            "static const int kRtcHeaderCount = {};\n".format(len(RTC_HEADERS))
            + "static const char* kRtcHeaderNames[] = {"
            + ", ".join('"{}"'.format(n) for n in RTC_HEADERS) + "};\n"
            + "static const char* kRtcHeaders[] = {\n" + ",\n".join(parts) + "};\n")
    old = open(RTC_INC).read() if os.path.exists(RTC_INC) else None
    if old != body:
        with open(RTC_INC, "w") as f:
            f.write(body)


def build(verbose=False):
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    write_rtc_headers()
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = OUT + ".tmp"
    cmd = [hipcc] + FLAGS + ["-o", tmp] + SRC + LIBS
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
