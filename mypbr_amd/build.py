"""Builds libmpt.so, the MI355X path-tracing core, in-tree.

hipcc compiles the HIP stages for gfx950 and the host C++ (scene parsing,
flattening, BVH build, tables) into one shared library next to this file.
Geometry is IEEE double with FP contraction off everywhere, so that device
and host arithmetic match the reference renderer's operation order bit for
bit; packed-FP32 instructions are disabled for the device.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.environ.get("MPT_LIB_OUT", os.path.join(HERE, "libmpt.so"))

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MPT_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra",
          "-Wno-unused-parameter", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"] + os.environ.get("MPT_DEVICE_FLAGS", "").split()

HOST_SOURCES = ["scene_host.cpp", "scene_api.cpp", "gltf_loader.cpp", "image_io.cpp", "image_decode.cpp"]
DEVICE_SOURCES = ["kernels.hip", "denoise.hip", "learn.hip", "nrc.hip", "display.hip"]


def run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(verbose=True):
    build_dir = os.environ.get("MPT_BUILD_DIR", BUILD)
    os.makedirs(build_dir, exist_ok=True)
    jobs = []
    for src in HOST_SOURCES:
        obj = os.path.join(build_dir, src + ".o")
        jobs.append((obj, [HIPCC, "-x", "c++", *COMMON, "-c", os.path.join(CSRC, src), "-o", obj]))
    for src in DEVICE_SOURCES:
        obj = os.path.join(build_dir, src + ".o")
        jobs.append((obj, [HIPCC, "-x", "hip", *COMMON, *DEVICE, "-c", os.path.join(CSRC, src), "-o", obj]))
    # the translation units compile independently; the largest (the stage
    # kernels) sets the wall time
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(len(jobs), os.cpu_count() or 1)) as pool:
        for f in [pool.submit(run, cmd) for _, cmd in jobs]:
            f.result()
    objs = [obj for obj, _ in jobs]
    tmp = LIB + ".tmp"
    run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, *objs, "-o", tmp, "-lpthread", "-lz"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    try:
        build()
    except subprocess.CalledProcessError as e:
        sys.exit(e.returncode)
