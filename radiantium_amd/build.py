"""Builds libradiantium_hip.so in-tree with hipcc for gfx950.

Every translation unit is compiled with -ffp-contract=off and correctly
rounded f32 division / sqrt so that float expressions round like the
reference's (.NET never contracts to FMA); device code is also built without
packed-FP32 instructions.  Usage: python -m radiantium_amd.build
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libradiantium_hip.so")
ARCH = os.environ.get("RD_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
          "-Wno-unused-variable"]
# hipcc compiles the .cpp host sources as HIP too: they get the same target
# (without it hipcc adds a code object for its default architecture)
TARGET = [f"--offload-arch={ARCH}", "-fhip-fp32-correctly-rounded-divide-sqrt",
          "-fno-gpu-flush-denormals-to-zero", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
DEVICE = ["-x", "hip"] + TARGET
HOST_SOURCES = ["rd_api.cpp", "rd_camera.cpp", "rd_bvh_build.cpp", "rd_envmap.cpp", "rd_json.cpp", "rd_image.cpp", "rd_octree.cpp"]
DEVICE_SOURCES = ["rd_kernels.hip", "rd_volpath.hip", "rd_sort.hip", "rd_scratch.hip", "rd_bvh_gpu.hip", "rd_light.hip", "rd_bdpt.hip"]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    cc = hipcc()
    objs = []
    jobs = []
    for src in HOST_SOURCES + DEVICE_SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        cmd = [cc] + COMMON + (DEVICE if src in DEVICE_SOURCES else TARGET) + ["-c", os.path.join(CSRC, src), "-o", obj]
        jobs.append((cmd, src))
        objs.append(obj)
    procs = [(subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT), src) for cmd, src in jobs]
    failed = False
    for p, src in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        # hipcc's host pass does not know the packed-fp32 target feature; that warning is expected
        lines = [l for l in text.splitlines() if "packed-fp32-ops" not in l and l.strip()]
        if p.returncode != 0:
            failed = True
            sys.stderr.write(f"--- {src} ---\n" + "\n".join(lines) + "\n")
        elif verbose and lines:
            sys.stderr.write(f"--- {src} ---\n" + "\n".join(lines) + "\n")
    if failed:
        raise RuntimeError("hipcc compilation failed")
    tmp = LIB + ".tmp"
    subprocess.check_call([cc, "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-lz"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
