"""Build libforestclaw_amd.so in-tree: HIP kernels for gfx950 plus the C host
layer, linked with hipcc.  Run `python -m forestclaw_amd.build`.

The FP64 kernels are compiled with -ffp-contract=off (bit-faithful to the CPU
Clawpack-4.6 arithmetic); the FP32-bulk kernels allow contraction.  Packed
FP32 instructions are disabled for all device code.

Also builds the example solver libraries of user_solvers/ (an application's
own Riemann solver compiled into the step kernel through
csrc/hipclaw_user_solver.h): each source twice, FP64 and FP32-bulk, linked
into user_solvers/lib<name>.so.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libforestclaw_amd.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CC = os.environ.get("CC", "gcc")
ARCH = "gfx950"

HIP_COMMON = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
              "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
              "-I" + os.path.join(ROOT, "include")]
HIP_SOURCES = {
    "hipclaw_kernels_fp64.hip": ["-ffp-contract=off"],
    "hipclaw_kernels_fp32.hip": [],
    "hipclaw_capi.hip": [],
}
C_SOURCES = ["fclaw2d_forest.c", "fclaw2d_hip_host.c"]
USER_DIR = os.path.join(HERE, "user_solvers")
USER_SOLVERS = {"acoustics": "rp_acoustics.hip", "tracers": "rp_tracers.hip"}
USER_PRECISION = {"64": ["-ffp-contract=off"], "32": []}
C_FLAGS = ["-O2", "-fPIC", "-std=gnu99", "-Wall", "-Wno-unused-function",
           "-I" + os.path.join(ROOT, "include")]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l
                    and l.strip())
    if r.returncode != 0:
        raise RuntimeError("build step failed: %s\n%s%s" % (" ".join(cmd), r.stdout, err))
    if verbose and err:
        print(err)


def user_solver_lib(name):
    return os.path.join(USER_DIR, "lib%s.so" % name)


def build(verbose=False, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 1)
    compiles, objs = [], []
    for src, extra in HIP_SOURCES.items():
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        compiles.append([HIPCC] + HIP_COMMON + extra + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in C_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(BUILD, src.replace(".c", ".o"))
        compiles.append([CC] + C_FLAGS + ["-c", path, "-o", obj])
        objs.append(obj)
    user_objs = {}
    for name, src in USER_SOLVERS.items():
        user_objs[name] = []
        for prec, extra in USER_PRECISION.items():
            obj = os.path.join(BUILD, "user_%s_fp%s.o" % (name, prec))
            compiles.append([HIPCC] + HIP_COMMON + extra +
                            ["-I" + CSRC, "-DHIPCLAW_USER_PRECISION=" + prec,
                             "-c", os.path.join(USER_DIR, src), "-o", obj])
            user_objs[name].append(obj)
    with ThreadPoolExecutor(jobs) as ex:
        for f in [ex.submit(_run, cmd, verbose) for cmd in compiles]:
            f.result()
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs, verbose)
    os.replace(tmp, LIB)
    for name, uo in user_objs.items():
        out = user_solver_lib(name)
        _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", out + ".tmp"] + uo, verbose)
        os.replace(out + ".tmp", out)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
