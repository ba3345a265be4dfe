"""Build libpagani.so (HIP, gfx950) and the command-line tools, in-tree.

    python pagani-toolkit_amd/build.py [--lib-only]

Outputs:
    pagani-toolkit_amd/lib/libpagani.so
    pagani-toolkit_amd/bin/{CUCorMat,CUBC,Cp,Degree,CUBFW_Lp,CUBFS_Lp,CUmodularity,CUEC,PC_CPU,ConvertNII,Maslov,SmallWorldProperty}
"""
from __future__ import annotations

import argparse
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ARCH = os.environ.get("PAGANI_OFFLOAD_ARCH", "gfx950")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Wno-unused-result",
    # no packed FP32 (v_pk_*_f32) instructions in any kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

TOOLS = ["CUCorMat", "CUBC", "Cp", "Degree", "CUBFW_Lp", "CUBFS_Lp", "CUmodularity", "CUEC", "PC_CPU", "ConvertNII", "Maslov", "SmallWorldProperty"]


def run(cmd):
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out)
    if r.returncode != 0:
        raise SystemExit(f"build failed: {cmd[0]} exited {r.returncode}")


def build_lib() -> str:
    os.makedirs(os.path.join(HERE, "lib"), exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip"))) + \
        sorted(glob.glob(os.path.join(HERE, "csrc", "*.cpp")))
    out = os.path.join(HERE, "lib", "libpagani.so")
    tmp = out + ".tmp"
    run([HIPCC, *HIP_FLAGS, "-shared", "-I", os.path.join(ROOT, "include"),
         "-I", os.path.join(HERE, "csrc"), *srcs, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_tools() -> None:
    hostdir = os.path.join(HERE, "host")
    srcs = sorted(glob.glob(os.path.join(hostdir, "tool_*.cpp")))
    if not srcs:
        return
    bindir = os.path.join(HERE, "bin")
    os.makedirs(bindir, exist_ok=True)
    common = [os.path.join(hostdir, "io.cpp")]
    libdir = os.path.join(HERE, "lib")
    for s in srcs:
        name = os.path.basename(s)[len("tool_"):-len(".cpp")]
        run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-I", hostdir,
             s, *common, "-L", libdir, "-lpagani", f"-Wl,-rpath,$ORIGIN/../lib",
             "-o", os.path.join(bindir, name)])


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib-only", action="store_true")
    a = ap.parse_args(argv)
    build_lib()
    if not a.lib_only:
        build_tools()
    return 0


if __name__ == "__main__":
    sys.exit(main())
