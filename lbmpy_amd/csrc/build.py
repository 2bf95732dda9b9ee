"""Build ``liblbm355.so`` (the C-ABI sweep library) for gfx950 with hipcc, in-tree.

Each ``.hip`` translation unit is compiled to an object in parallel, then linked into
``lbmpy_amd/_lib/liblbm355.so``.  Run ``python -m lbmpy_amd.csrc.build`` or call
:func:`build`.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIBDIR = os.path.join(PKG, "_lib")
LIBNAME = "liblbm355.so"
SOURCES = ["sweep.hip", "sweep_eso.hip", "macroscopic.hip", "batch.hip", "closure.hip", "periodic.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # no packed FP32 instructions (v_pk_*_f32): keep arithmetic on the plain VALU path
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=fast", "-munsafe-fp-atomics",
    "-Wno-unused-result",
]


def library_path():
    return os.path.join(LIBDIR, LIBNAME)


def build(verbose=False, jobs=None):
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(LIBDIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, "-I", os.path.join(PKG, os.pardir, "include"), "-c",
                     os.path.join(HERE, src), "-o", obj])

    def run(cmd):
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed ({' '.join(cmd)}):\n{r.stdout}\n{r.stderr}")
        return r

    with ThreadPoolExecutor(max_workers=jobs or min(8, len(cmds))) as ex:
        list(ex.map(run, cmds))
    tmp = library_path() + ".tmp"
    run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, library_path())
    return library_path()


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
