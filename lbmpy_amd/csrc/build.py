"""Build ``liblbm355.so`` (the C-ABI sweep library) for gfx950 with hipcc, in-tree.

Each ``.hip`` translation unit is compiled to an object in parallel, then linked into
``lbmpy_amd/_lib/liblbm355.so``.  Run ``python -m lbmpy_amd.csrc.build`` or call
:func:`build`.
"""
import hashlib
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIBDIR = os.path.join(PKG, "_lib")
LIBNAME = "liblbm355.so"
SOURCES = ["sweep.hip", "sweep_moment_f64.hip", "sweep_moment_f32.hip", "sweep_central_f64.hip",
           "sweep_central_f32.hip", "sweep_cumulant_f64.hip", "sweep_cumulant_f32.hip",
           "sweep_cumulant_corrected_f64.hip", "sweep_cumulant_corrected_f32.hip",
           "sweep_cumulant_general_f64.hip", "sweep_cumulant_general_f32.hip", "sweep_cumulant_eq3_f64.hip",
           "sweep_cumulant_eq3_f32.hip", "sweep_cumulant_learned_f32.hip",
           "sweep_eso.hip", "macroscopic.hip", "batch.hip", "closure.hip", "periodic.hip"]
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


def source_files():
    """The files the library is built from: every .hip and .h under csrc/ and the C-ABI header."""
    names = sorted(n for n in os.listdir(HERE) if n.endswith((".hip", ".h")))
    return [os.path.join(HERE, n) for n in names] + [os.path.join(PKG, os.pardir, "include", "lbm355.h")]


def source_hash():
    """SHA-256 over the names and contents of source_files(), as compiled into lbm355_source_hash()."""
    h = hashlib.sha256()
    for p in source_files():
        h.update(os.path.basename(p).encode() + b"\0")
        with open(p, "rb") as fh:
            h.update(fh.read())
        h.update(b"\0")
    return h.hexdigest()


def build(verbose=False, jobs=None):
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(LIBDIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    cmds = []
    digest = source_hash()
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, f'-DLBM355_SOURCE_HASH="{digest}"',
                     "-I", os.path.join(PKG, os.pardir, "include"), "-c",
                     os.path.join(HERE, src), "-o", obj])

    def run(cmd):
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed ({' '.join(cmd)}):\n{r.stdout}\n{r.stderr}")
        return r

    with ThreadPoolExecutor(max_workers=jobs or min(os.cpu_count() or 8, 16, len(cmds))) as ex:
        list(ex.map(run, cmds))
    tmp = library_path() + ".tmp"
    run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, library_path())
    return library_path()


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
