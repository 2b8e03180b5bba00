"""Builds the gfx950 HIP library apirl_amd/libapirl_hip.so in-tree with hipcc.

    python -m apirl_amd.build

The library is self-contained (HIP runtime only); its C-ABI is include/apirl_hip.h."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libapirl_hip.so")
BUILD = os.path.join(ROOT, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("APIRL_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: no fused multiply-add anywhere, so every floor() at a voxel face
#   and every weight rounds as in the reference's float code.
# correctly rounded fp32 division / sqrt (the reference runs on IEEE x86 SSE).
# packed-fp32 instructions disabled (v_pk_* fp32 forms are avoided on this pool).
COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
          "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
DEVICE = ["--offload-arch=" + ARCH, "-fhip-fp32-correctly-rounded-divide-sqrt",
          "-fno-gpu-flush-denormals-to-zero", "-munsafe-fp-atomics",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

SOURCES = [("siddon_kernels.hip", True), ("gather_kernels.hip", True), ("geometry.cpp", False),
           ("capi.cpp", False), ("batch.cpp", False)]


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for name, dev in SOURCES:
        src = os.path.join(CSRC, name)
        if not os.path.exists(src):
            continue
        obj = os.path.join(BUILD, name + ".o")
        cmd = [HIPCC] + COMMON + (DEVICE if dev else ["-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"]) + \
            ["-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd))
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
