"""Build libvpfft.so (HIP kernels + C ABI) for gfx950 with hipcc.

Usage: python -m vpfft_amd.build [--verbose]
The library is written next to this file so it travels with the source tree.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libvpfft.so")
RUNNER = os.path.join(HERE, "vpfft_run")
SOURCES = ["capi.hip", "fft_green.hip", "voxel_kernels.hip", "microstructure.cpp", "deck_io.cpp", "serve.cpp"]
ARCH = os.environ.get("VPFFT_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    return "hipcc"


def build(verbose: bool = False) -> str:
    objs = []
    common = [
        "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
        # no packed-FP32 instructions (v_pk_*_f32) anywhere in the device code
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-ffp-contract=fast-honor-pragmas", "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
    ]
    outdir = os.path.join(HERE, "_build")
    os.makedirs(outdir, exist_ok=True)
    # the translation units are independent: compile them side by side
    procs = []
    for src in SOURCES:
        obj = os.path.join(outdir, src + ".o")
        cmd = [hipcc()] + common + ["-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd = [hipcc(), "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-x", "c++", "-c",
                   os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd)))
        objs.append(obj)
    failed = [cmd for cmd, p in procs if p.wait() != 0]
    if failed:
        raise subprocess.CalledProcessError(1, failed[0])
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    # the deck-driven command-line runner, linked against the library next to it
    cmd = [hipcc(), "-O2", "-std=c++17", "-Wall", "-x", "c++", os.path.join(CSRC, "vpfft_run.cpp"),
           "-o", RUNNER + ".tmp", "-L", HERE, "-lvpfft", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(RUNNER + ".tmp", RUNNER)
    # the C++ facades' test drivers (tests/cpp), linked the same way
    for name in ("facade_check", "solvers_check"):
        check_src = os.path.join(os.path.dirname(HERE), "tests", "cpp", name + ".cpp")
        if not os.path.exists(check_src):
            continue
        out = os.path.join(os.path.dirname(check_src), name)
        cmd = [hipcc(), "-O2", "-std=c++17", "-Wall", "-x", "c++", check_src, "-o", out + ".tmp",
               "-I", os.path.join(os.path.dirname(HERE), "include"), "-L", HERE, "-lvpfft",
               "-Wl,-rpath,$ORIGIN/../../vpfft_amd"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(out + ".tmp", out)
    return LIB


if __name__ == "__main__":
    print(build(verbose="--verbose" in sys.argv))
