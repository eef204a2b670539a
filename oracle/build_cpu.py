"""Build the oracle's C restatement (cpu_eval.c -> oracle/lib/libtfpwa_cpu.so) with gcc and
OpenMP. x86-64-v3 (AVX2, FMA) so the library runs on any current EPYC host."""

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "lib", "libtfpwa_cpu.so")


def build(verbose=False):
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    # a per-process temporary, so concurrent builds (parallel test workers) never share one
    tmp = "{}.tmp{:x}".format(OUT, os.getpid())
    cmd = ["gcc", "-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared", "-std=c11",
           "-o", tmp, os.path.join(HERE, "cpu_eval.c"), "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


def load():
    """ctypes handle of the built library (cpu_eval, cpu_eval_threads)."""
    import ctypes

    if not os.path.exists(OUT):
        raise RuntimeError("oracle C library not built; run __graft_entry__.build()")
    L = ctypes.CDLL(OUT)
    dp = ctypes.POINTER(ctypes.c_double)
    L.cpu_eval.argtypes = [dp, dp, ctypes.c_int64, ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, dp]
    L.cpu_eval.restype = ctypes.c_int
    L.cpu_eval_threads.restype = ctypes.c_int
    return L


if __name__ == "__main__":
    print(build(verbose=True))
