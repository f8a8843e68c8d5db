"""Builds the package's native libraries in place:
    libvtgpu.so    the HIP kernels and the C-ABI (include/vtgpu.h), gfx950 only
    libvtsynth.so  the synthetic-frame generator (include/vtsynth.h), plain C
Each HIP translation unit compiles to its own object under _build/ in parallel, then one link."""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(REPO, "_build")
INC = os.path.join(REPO, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -packed-fp32-ops: no v_pk_*_f32 from the compiler (their results have been unreliable on these boxes).
HIP_FLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-I" + INC, "-I" + CSRC,
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-result"]
HIP_SOURCES = ["ctx.cpp", "pyramid.hip", "lk.hip", "ncc.hip", "filters.hip", "still.hip", "stereo.hip", "tune.cpp", "serve.cpp", "door.cpp",
               "track.cpp"]

LIB_GPU = os.path.join(HERE, "libvtgpu.so")
LIB_SYNTH = os.path.join(HERE, "libvtsynth.so")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("build failed: " + os.path.basename(cmd[-1]))
    return r


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]

    def one(src):
        obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else ["-x", "hip"]
        _run([HIPCC] + HIP_FLAGS + lang + ["-c", "-o", obj, os.path.join(CSRC, src)])
        return obj

    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(one, srcs))
    tmp = LIB_GPU + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs)
    os.replace(tmp, LIB_GPU)
    tmp = LIB_SYNTH + ".tmp"
    _run(["gcc", "-std=gnu11", "-O2", "-fPIC", "-shared", "-pthread", "-I" + INC, "-o", tmp,
          os.path.join(CSRC, "synth.c"), "-lm"])
    os.replace(tmp, LIB_SYNTH)
    if verbose:
        print("built", LIB_GPU, LIB_SYNTH)


if __name__ == "__main__":
    build(verbose=True)
