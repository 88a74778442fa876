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
PHASE_LIB = os.path.join(PKG, "libppm_phase.so")
PHASE_SOURCE = "ppm_batch_kernels.hip"  # the one object that differs in libppm_phase.so

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
    """libppm.so, and libppm_phase.so: the same library whose batched
    integral carries phase timers (ppm_phase_timers; PPM_LIBRARY=phase
    loads it)."""
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in LIB_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    jobs = [(s, ()) for s in srcs] + [(PHASE_SOURCE, ("-DPPM_PHASE_TIMERS=1",))]

    def one(job):
        src, defs = job
        obj = os.path.join(OBJ, src + (".phase" if defs else "") + ".o")
        lang = [] if src.endswith(".hip") else ["-x", "hip"]
        _run([HIPCC] + HIP_FLAGS + list(defs) + lang + ["-c", os.path.join(CSRC, src), "-o", obj])
        return obj

    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        objs = list(ex.map(one, jobs))
    plain, phase = objs[:len(srcs)], [o if s != PHASE_SOURCE else objs[-1] for o, s in zip(objs, srcs)]
    for out, ob in ((LIB, plain), (PHASE_LIB, phase)):
        tmp = out + ".tmp"
        _run([HIPCC, "-shared", "--offload-arch=" + ARCH, "-o", tmp] + ob + ["-lpthread"])
        os.replace(tmp, out)


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
