"""Build libsusy_hip_U<N>.so (N = 2, 3, 4) for gfx950 with hipcc, in-tree.

Every HIP source is compiled once per gauge group (NCOL is a compile-time
constant, as in the reference's include/susy.h:16) and linked into
susy_amd/lib/libsusy_hip_U<N>.so.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
BUILDDIR = os.path.join(HERE, "_build")
SOURCES = ["lattice.hip", "dslash.hip", "fields.hip", "force.hip", "blas.hip",
           "cg.hip", "capi.hip", "surrogate.hip", "eig.hip", "meas.hip", "pfaff.hip"]
NCOLS = (2, 3, 4, 5)
ARCH = os.environ.get("SUSY_HIP_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall",
    "-Wno-unused-function", "-Wno-unused-variable",
    # packed FP32 instructions are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def lib_path(ncol):
    return os.path.join(LIBDIR, f"libsusy_hip_U{ncol}.so")


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError("command failed:\n" + " ".join(cmd) + "\n" + p.stdout)
    return p.stdout


def build(ncols=NCOLS, jobs=8, verbose=False, extra_flags=()):
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(BUILDDIR, exist_ok=True)
    tasks = []
    for n in ncols:
        for src in SOURCES:
            obj = os.path.join(BUILDDIR, f"{os.path.splitext(src)[0]}_U{n}.o")
            cmd = [HIPCC, *CFLAGS, *extra_flags, f"-DNCOL={n}", "-c",
                   os.path.join(CSRC, src), "-o", obj]
            tasks.append((n, obj, cmd))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        outs = list(ex.map(lambda t: _run(t[2]), tasks))
    if verbose:
        for (_, _, cmd), out in zip(tasks, outs):
            if out.strip():
                print(" ".join(cmd[-3:]), "\n", out)
    for n in ncols:
        objs = [t[1] for t in tasks if t[0] == n]
        _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", lib_path(n),
              *objs, "-lpthread"])
    return [lib_path(n) for n in ncols]


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
