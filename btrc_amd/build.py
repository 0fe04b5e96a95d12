"""Build recipe of libbtrc_hip.so (gfx950) and of the CPU oracle.

    python -m btrc_amd.build            # library + oracle (+ oracle/_ref when the reference is present)

Device code is compiled for gfx950 only, with FP contraction off so that
host and device arithmetic agree bit for bit, and with packed-FP32 code
generation disabled.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "btrc_amd")
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(PKG, "libbtrc_hip.so")

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-I" + os.path.join(ROOT, "include")]

# per-source extra flags: the cache's kernels keep MFMA results in VGPRs (no
# accumulator-register read-backs between layers)
EXTRA_FLAGS = {
    "nrc.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
}

HIP_SOURCES = [
    os.path.join(CSRC, "kernels", "wavefront.hip"),
    os.path.join(CSRC, "kernels", "nrc.hip"),
    os.path.join(CSRC, "kernels", "bvh_gpu.hip"),
    os.path.join(CSRC, "host", "context.hip"),
]
HOST_SOURCES = [
    os.path.join(CSRC, "host", "bvh_build.cpp"),
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops' is not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build_library(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        jobs.append([HIPCC, *HIP_FLAGS, *EXTRA_FLAGS.get(os.path.basename(src), []), "-c", src, "-o", obj])
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        jobs.append(["g++", *HOST_FLAGS, "-c", src, "-o", obj])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for out in ex.map(_run, jobs):
            if verbose and out.strip():
                print(out)
    objs = [j[j.index("-o") + 1] for j in jobs]
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


def build_oracle():
    src = os.path.join(ROOT, "oracle", "oracle.cpp")
    out = os.path.join(ROOT, "oracle", "liboracle.so")
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-Wno-unknown-pragmas",
          src, "-o", tmp, "-lpthread"])
    os.replace(tmp, out)
    ref = os.path.join(ROOT, "oracle", "build_ref.sh")
    _run(["sh", ref])
    return out


def build_all(verbose=False):
    build_library(verbose)
    build_oracle()


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
    print("built", LIB)
