"""Build libpgfem.so in-tree: host C++ with g++, HIP kernels with hipcc for gfx950.

Usage: python -m pigfem_assignment_amd.build  (or import and call build()).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpgfem.so")
# compiled C++ caller of the C-ABI (pgfem.h / pgfem_host.h only), linked to the library
DRIVER_SRC = os.path.join(CSRC, "driver", "pgf_driver.cpp")
DRIVER = os.path.join(HERE, "pgf_driver")

HOST_SRC = ["host/pgf_mesh.cpp", "host/pgf_dofs.cpp", "host/pgf_problem.cpp", "host/pgf_sens.cpp",
            "host/pgf_writer.cpp", "host/pgf_batch_host.cpp"]
HIP_SRC = ["kernels/pgf_assemble.hip", "kernels/pgf_device.hip", "kernels/pgf_cnn.hip", "kernels/pgf_batch.hip",
           "kernels/pgf_sens.hip", "kernels/pgf_preprocess.hip"]
# the device set-up shares its arithmetic with the host set-up (g++, no FMA on
# x86-64): no multiply-add contraction, so both round every operation alike
HIP_EXTRA = {"kernels/pgf_preprocess.hip": ["-ffp-contract=off"]}

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
CXXFLAGS = ["-O3", "-fPIC", "-std=c++17", "-fvisibility=hidden", "-I" + os.path.join(ROOT, "include")]
# packed FP32 instructions are disabled for every device code object
HIPFLAGS = CXXFLAGS + ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    objs = []
    for s in HOST_SRC:
        o = os.path.join(OBJ, os.path.basename(s) + ".o")
        objs.append(o)
        jobs.append(["g++"] + CXXFLAGS + ["-fopenmp", "-c", os.path.join(CSRC, s), "-o", o])
    for s in HIP_SRC:
        o = os.path.join(OBJ, os.path.basename(s) + ".o")
        objs.append(o)
        jobs.append([HIPCC] + HIPFLAGS + HIP_EXTRA.get(s, []) + ["-x", "hip", "-c", os.path.join(CSRC, s), "-o", o])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs + ["-lgomp"])
    os.replace(tmp, LIB)
    tmp = DRIVER + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), DRIVER_SRC, "-o", tmp,
          "-L" + HERE, "-lpgfem", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, DRIVER)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
