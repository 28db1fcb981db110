"""Build libpllcart3d_amd.so (HIP, gfx950) and the run3_amd driver in-tree.

hipcc compiles every HIP source for gfx950 only; the plain FP64 path needs
-ffp-contract=off (no fused multiply-add, the reference's arithmetic).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INC = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libpllcart3d_amd.so")
DRIVER = os.path.join(HERE, "run3_amd")
OBJDIR = os.path.join(HERE, "_build")

# (source, extra flags): the FP64 kernels and the host code never fuse
# multiply-adds; the FP32 kernels of the mixed path may
HIP_SOURCES = [
    ("pllc_kernels_f64.hip", ["-ffp-contract=off"]),
    ("pllc_kernels_f32.hip", ["-ffp-contract=on", "-fno-hip-fp32-correctly-rounded-divide-sqrt"]),
    ("pllc_runtime.hip", ["-ffp-contract=off"]),
    ("pllc_abi.hip", ["-ffp-contract=off"]),
    ("pllc_closure.hip", ["-ffp-contract=on"]),
]
HIPFLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-I" + INC, "-I" + CSRC,
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build step failed: " + " ".join(cmd[:3]) + " ...")
    return out


def build(verbose=False):
    os.makedirs(OBJDIR, exist_ok=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objs = []
    for src, extra in HIP_SOURCES:
        obj = os.path.join(OBJDIR, src.replace(".hip", ".o"))
        out = _run([hipcc] + HIPFLAGS + extra + ["-c", os.path.join(CSRC, src), "-o", obj])
        if verbose and out.strip():
            print(out)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([hipcc, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    _run(["gcc", "-O2", "-Wall", "-I" + INC, os.path.join(CSRC, "driver.c"), "-o", DRIVER,
          "-L" + HERE, "-lpllcart3d_amd", "-Wl,-rpath,$ORIGIN", "-lm"])
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
