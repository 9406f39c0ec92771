"""Builds osqp_amd/libosqp_amd.so in-tree for gfx950.

  python -m osqp_amd.build          (or osqp_amd.build.build())

C host core: gcc.  HIP kernels: hipcc --offload-arch=gfx950.  The Python front
end's marshalling helper (osqp_amd/_marshal.so, CPython C API): gcc.  Every HIP file is
compiled with -ffp-contract=off (elementwise results must equal the CPU's
correctly rounded values) and with packed-FP32 instructions disabled.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libosqp_amd.so")
MARSHAL = os.path.join(HERE, "_marshal.so")

HIP_SOURCES = ["hip/runtime.hip", "hip/vector.hip", "hip/matrix.hip", "hip/panel.hip", "hip/pcg.hip", "serve/batch.hip"]
C_SOURCES = ["host/auxil.c", "host/osqp_api.c", "host/polish.c", "host/profilers.c"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result", "-I", os.path.join(ROOT, "include")]
C_FLAGS = ["-O2", "-fPIC", "-std=gnu99", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
           "-I", os.path.join(ROOT, "include"), "-I/opt/rocm/include"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def _obj(src):
    return os.path.join(BUILD, src.replace("/", "_") + ".o")


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        if os.path.exists(path):
            jobs.append([HIPCC] + HIP_FLAGS + ["-c", path, "-o", _obj(src)])
    for src in C_SOURCES:
        jobs.append(["gcc"] + C_FLAGS + ["-c", os.path.join(CSRC, src), "-o", _obj(src)])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for out in ex.map(_run, jobs):
            if verbose and out.strip():
                print(out)
    objs = [_obj(s) for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))] + [_obj(s) for s in C_SOURCES]
    tmp = LIB + ".tmp"
    # roctx (rocprofiler-sdk) carries the profiler sections
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs +
         ["-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib", "-lrocprofiler-sdk-roctx"])
    shutil.move(tmp, LIB)
    build_marshal()
    return LIB


def build_marshal():
    """osqp_amd/_marshal.so: packs lists of QPs into osqp_batch_add_many's
    arguments without per-array Python work (osqp_amd/serve.py uses it when
    it imports; it links nothing but the interpreter's C API)."""
    import sysconfig
    tmp = MARSHAL + ".tmp"
    _run(["gcc", "-O3", "-fPIC", "-shared", "-Wall", "-I", sysconfig.get_paths()["include"], "-I",
          os.path.join(ROOT, "include"), os.path.join(CSRC, "py", "marshal.c"), "-o", tmp])
    shutil.move(tmp, MARSHAL)
    return MARSHAL


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
