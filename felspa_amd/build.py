"""Builds felspa_amd/libfelspa_amd.so in-tree: host C++ with g++ (OpenMP),
device code with hipcc for gfx950, linked into one shared library whose
C-ABI is include/fs_stokes.h + include/fs_assemble.h (+ include/fs_guess.h)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INC = os.path.join(ROOT, "include")
OUT = os.path.join(HERE, "libfelspa_amd.so")
OBJ = os.path.join(HERE, "_build")

HOST_SRC = ["assemble.cpp", "csr_util.cpp", "ordering.cpp", "blocks_io.cpp"]
HIP_SRC = ["solver.hip", "stokes.hip", "ilu_api.hip", "guess.hip", "assemble_gpu.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
             # packed FP32 instructions stay off (v_pk_*_f32)
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I", INC, "-I", CSRC]
HOST_FLAGS = ["-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fopenmp", "-fPIC", "-std=c++17", "-Wall",
              "-I", INC, "-I", CSRC]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("build step failed: %s" % os.path.basename(cmd[-1]))
    return r


def build(verbose=False, jobs=8):
    """Compile every source and link the shared library (always rebuilds)."""
    os.makedirs(OBJ, exist_ok=True)
    from concurrent.futures import ThreadPoolExecutor
    steps = []
    objs = []
    for s in HOST_SRC:
        o = os.path.join(OBJ, s + ".o")
        steps.append(["g++"] + HOST_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    for s in HIP_SRC:
        if not os.path.exists(os.path.join(CSRC, s)):
            continue
        o = os.path.join(OBJ, s + ".o")
        steps.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, steps))
    tmp = OUT + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs +
         ["-fopenmp", "-lgomp", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
