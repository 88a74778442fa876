"""Builds libppm (HIP, gfx950), the table generator and the drop-in programs
in-tree with explicit hipcc / gcc invocations.

    python -m gpu_ppm_amd.build            # everything
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BIN = os.path.join(PKG, "_bin")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libppm.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PPM_OFFLOAD_ARCH", "gfx950")
# packed FP32 (v_pk_*_f32) is disabled on purpose: only ordinary FP32 VALU ops
HIP_FLAGS = ["-std=c++20", "-O3", "-fPIC", "--offload-arch=" + ARCH,
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-result", "-I" + os.path.join(ROOT, "include")]
HOST_FLAGS = ["-std=c++20", "-O2", "-fPIC", "-Wall", "-Wno-unused-result",
              "-I" + os.path.join(ROOT, "include")]

LIB_SOURCES = ["ppm_kernels.hip", "ppm_batch_kernels.hip", "ppm_api.cpp", "ppm_tables.cpp",
               "ppm_batch.cpp", "ppm_serve.cpp", "ppm_response_gen.hip", "ppm_lensing_gen.hip"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n  %s\n%s" % (" ".join(cmd), r.stdout))
    return r.stdout


def build_synth():
    os.makedirs(BIN, exist_ok=True)
    _run(["gcc", "-O2", "-fopenmp", "-Wall", "-o", os.path.join(BIN, "ppm_synth"),
          os.path.join(PKG, "synth", "ppm_synth.c"), "-lm"])


def build_lib():
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in LIB_SOURCES if os.path.exists(os.path.join(CSRC, s))]

    def one(src):
        obj = os.path.join(OBJ, src + ".o")
        cmd = [HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        if not src.endswith(".hip"):
            cmd = [HIPCC] + HIP_FLAGS + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
        _run(cmd)
        return obj

    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(one, srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "--offload-arch=" + ARCH, "-o", tmp] + objs + ["-lpthread"])
    os.replace(tmp, LIB)


def build_apps():
    os.makedirs(BIN, exist_ok=True)
    apps = os.path.join(PKG, "apps")
    # host-only tool (no device code): the response builder
    _run(["g++"] + HOST_FLAGS + ["-O2", "-o", os.path.join(BIN, "ppm_response"), os.path.join(apps, "ppm_response.cpp")])
    for name in ("os1_gpu", "devarshi24st_gpu", "t_interp_gpu"):
        src = os.path.join(apps, name + ".cpp")
        if not os.path.exists(src):
            continue
        _run(["g++"] + HOST_FLAGS + ["-o", os.path.join(BIN, name), src, "-L" + PKG, "-lppm",
                                     "-Wl,-rpath,$ORIGIN/..", "-lpthread"])


def build_oracle():
    """The FP64 oracle used by the tests (and the reference CPU programs when
    their sources are present)."""
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-C", odir, "all"])
    if os.path.isdir("/root/reference/cpu") and shutil.which("clang++"):
        _run(["make", "-s", "-C", odir, "ref"])


def build_all():
    build_synth()
    build_lib()
    build_apps()
    build_oracle()


if __name__ == "__main__":
    build_all()
    print("built", LIB)
    sys.exit(0)
