"""ctypes binding of the C-ABI sweep library ``liblbm355.so`` (``include/lbm355.h``).

torch is imported first so that the library's ``libamdhip64.so.7`` dependency
resolves to the HIP runtime torch already loaded: PDF tensors, streams and the
kernels then live in one runtime.  There is no fallback: if the library is
missing the import raises.
"""
import ctypes
import os

import torch  # noqa: F401  (must precede the library load, see module docstring)

from .csrc.build import library_path, source_hash

c_int32, c_int64, c_double, c_void_p = ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p

D2Q9, D3Q19, D3Q27 = 0, 1, 2
FAMILY_MOMENT, FAMILY_CENTRAL, FAMILY_CUMULANT = 0, 1, 2
PULL, AA_EVEN, AA_ODD = 0, 1, 2
ESOTWIST_EVEN, ESOTWIST_ODD, ESOPULL_EVEN, ESOPULL_ODD, ESOPUSH_EVEN, ESOPUSH_ODD, PUSH = 3, 4, 5, 6, 7, 8, 9
# EVEN kernel code of each in-place pattern (the ODD kernel is the next code)
INPLACE_EVEN = {"aa": AA_EVEN, "esotwist": ESOTWIST_EVEN, "esopull": ESOPULL_EVEN, "esopush": ESOPUSH_EVEN}
F64, F32, F16 = 0, 1, 2
FACE_GHOST, FACE_PERIODIC, FACE_NOSLIP = 0, 1, 2
FLAG_ZERO_CENTERED, FLAG_COMPRESSIBLE, FLAG_EQ_SHIFT, FLAG_FORCE, FLAG_SMAGORINSKY = 1, 2, 4, 8, 16
FLAG_LEARNED_CLOSURE, FLAG_GALILEAN, FLAG_FOURTH_ORDER, FLAG_ADAPT_VELOCITY = 32, 64, 128, 256
FLAG_STOCHASTIC_ROUNDING, FLAG_GENERAL_CUMULANT, FLAG_HIGHER_ORDERS_RATE1 = 512, 1024, 2048
FLAG_SHEAR_PROJECTOR, FLAG_GAUSSIAN_EQUILIBRIUM = 4096, 8192

DTYPE_CODES = {"float64": F64, "float32": F32, "float16": F16}
STENCIL_NAMES = {"D2Q9": D2Q9, "D3Q19": D3Q19, "D3Q27": D3Q27}


BC_NOSLIP, BC_UBB, BC_FREESLIP, BC_FIXED_DENSITY, BC_OUTFLOW, BC_OUTFLOW_INIT = 0, 1, 2, 3, 4, 5
BC_FREESLIP_LINKS, BC_SIMPLE_OUTFLOW = 6, 7


class BoundaryDesc(ctypes.Structure):
    _fields_ = [("stencil", c_int32), ("pattern", c_int32), ("store_dtype", c_int32), ("kind", c_int32),
                ("nx", c_int32), ("ny", c_int32), ("nz", c_int32), ("ghost_layers", c_int32),
                ("periodic", c_int32 * 3), ("flags", c_int32), ("normal", c_int32 * 3), ("pad", c_int32),
                ("fused_noslip", c_int32 * 3), ("pad2", c_int32),
                ("stride_y", c_int64), ("stride_z", c_int64), ("stride_q", c_int64),
                ("velocity", c_double * 3), ("density", c_double), ("force", c_double * 3),
                ("equilibrium", c_void_p)]


class SweepDesc(ctypes.Structure):
    _fields_ = [("stencil", c_int32), ("family", c_int32), ("pattern", c_int32), ("store_dtype", c_int32),
                ("compute_dtype", c_int32), ("nx", c_int32), ("ny", c_int32), ("nz", c_int32),
                ("ghost_layers", c_int32), ("face", c_int32 * 3),
                ("stride_y", c_int64), ("stride_z", c_int64), ("stride_q", c_int64),
                ("out_stride_y", c_int64), ("out_stride_z", c_int64), ("out_stride_c", c_int64),
                ("flags", c_int32), ("n_active", c_int32),
                ("force", c_double * 3), ("smagorinsky_constant", c_double), ("omega_shear", c_double),
                ("tables", c_void_p), ("closure", c_void_p),
                ("omega_out", c_void_p), ("eddy_viscosity_out", c_void_p), ("density_in", c_void_p),
                ("velocity_in", c_void_p), ("rounding_seed", ctypes.c_uint32), ("waves_per_simd", c_int32)]


EXPORTED = ("lbm355_sweep", "lbm355_table_elements", "lbm355_noslip_links", "lbm355_reduce_conserved",
            "lbm355_count_nonfinite",
            "lbm355_set_equilibrium", "lbm355_get_macroscopic", "lbm355_stream_copy", "lbm355_desc_size",
            "lbm355_last_error", "lbm355_version", "lbm355_batch_create", "lbm355_batch_sweep",
            "lbm355_batch_info", "lbm355_batch_destroy", "lbm355_batch_sweep_static", "lbm355_closure_params_size",
            "lbm355_closure_master_size", "lbm355_closure_features", "lbm355_closure_forward",
            "lbm355_closure_train_step", "lbm355_closure_adam", "lbm355_boundary_links", "lbm355_copy_boxes",
            "lbm355_pattern_access", "lbm355_get_strain_rate", "lbm355_plane_copy", "lbm355_valu_probe",
            "lbm355_source_hash")

_lib = None


def load(path=None):
    """Load the library (once) and declare the function signatures."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    path = path or library_path()
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it with `python -m lbmpy_amd.csrc.build` "
                          "(or __graft_entry__.build())")
    lib = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
    lib.lbm355_sweep.argtypes = [ctypes.POINTER(SweepDesc), c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]
    lib.lbm355_sweep.restype = c_int32
    lib.lbm355_table_elements.argtypes = [c_int32, c_int32]
    lib.lbm355_table_elements.restype = c_int64
    lib.lbm355_noslip_links.argtypes = [c_int32, c_int32, c_int32, c_void_p, c_int64, c_int64, c_int64, c_void_p,
                                        c_int64, c_void_p]
    lib.lbm355_noslip_links.restype = c_int32
    lib.lbm355_reduce_conserved.argtypes = [ctypes.POINTER(SweepDesc), c_void_p, c_void_p, c_void_p]
    lib.lbm355_reduce_conserved.restype = c_int32
    lib.lbm355_set_equilibrium.argtypes = [ctypes.POINTER(SweepDesc), c_void_p, c_void_p, c_void_p, c_void_p,
                                           c_int32, c_void_p]
    lib.lbm355_set_equilibrium.restype = c_int32
    lib.lbm355_get_macroscopic.argtypes = [ctypes.POINTER(SweepDesc), c_void_p, c_void_p, c_void_p, c_int32,
                                           c_void_p]
    lib.lbm355_get_macroscopic.restype = c_int32
    lib.lbm355_stream_copy.argtypes = [c_void_p, c_void_p, c_int64, c_int32, c_void_p]
    lib.lbm355_stream_copy.restype = c_int32
    lib.lbm355_plane_copy.argtypes = [c_int32, c_int32, c_void_p, c_void_p, c_int64, c_int64, c_int32, c_int32,
                                      c_void_p]
    lib.lbm355_plane_copy.restype = c_int32
    lib.lbm355_valu_probe.argtypes = [c_void_p, c_int64, c_int32, c_void_p]
    lib.lbm355_valu_probe.restype = c_int32
    lib.lbm355_count_nonfinite.argtypes = [ctypes.POINTER(SweepDesc), c_void_p, c_void_p, c_void_p]
    lib.lbm355_count_nonfinite.restype = c_int32
    lib.lbm355_batch_create.argtypes = [ctypes.POINTER(SweepDesc), ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p),
                                        c_int32, ctypes.POINTER(c_void_p)]
    lib.lbm355_batch_create.restype = c_int32
    lib.lbm355_batch_sweep.argtypes = [c_void_p, c_int32, c_int32, c_int32, c_int32, c_void_p]
    lib.lbm355_batch_sweep.restype = c_int32
    lib.lbm355_batch_info.argtypes = [c_void_p, ctypes.POINTER(c_int64), ctypes.POINTER(c_int64)]
    lib.lbm355_batch_info.restype = c_int32
    lib.lbm355_batch_destroy.argtypes = [c_void_p]
    lib.lbm355_batch_destroy.restype = c_int32
    lib.lbm355_batch_sweep_static.argtypes = [ctypes.POINTER(SweepDesc), c_void_p, c_void_p, c_int32, c_int32, c_int32,
                                              c_int32, c_int32, c_void_p]
    lib.lbm355_batch_sweep_static.restype = c_int32
    lib.lbm355_closure_params_size.restype = c_int64
    lib.lbm355_closure_master_size.restype = c_int32
    lib.lbm355_closure_features.argtypes = [c_void_p, c_void_p, c_int64, c_void_p, c_void_p]
    lib.lbm355_closure_features.restype = c_int32
    lib.lbm355_closure_forward.argtypes = [c_void_p, c_void_p, c_int64, c_void_p, c_void_p]
    lib.lbm355_closure_forward.restype = c_int32
    lib.lbm355_closure_train_step.argtypes = [c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_int32,
                                              c_void_p]
    lib.lbm355_closure_train_step.restype = c_int32
    lib.lbm355_closure_adam.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_float, ctypes.c_float,
                                        ctypes.c_float, ctypes.c_float, c_int32, c_void_p, c_void_p]
    lib.lbm355_closure_adam.restype = c_int32
    lib.lbm355_boundary_links.argtypes = [ctypes.POINTER(BoundaryDesc), c_void_p, c_void_p, c_int64, c_void_p,
                                          c_void_p]
    lib.lbm355_boundary_links.restype = c_int32
    lib.lbm355_get_strain_rate.argtypes = [ctypes.POINTER(SweepDesc), c_void_p, c_void_p, c_void_p, c_double, c_int32,
                                           c_void_p]
    lib.lbm355_get_strain_rate.restype = c_int32
    lib.lbm355_copy_boxes.argtypes = [c_int32, c_void_p, c_int64, c_int64, c_int64, c_int64, c_void_p, c_int32,
                                      c_void_p]
    lib.lbm355_copy_boxes.restype = c_int32
    lib.lbm355_pattern_access.argtypes = [c_int32, c_int32, c_void_p]
    lib.lbm355_pattern_access.restype = c_int32
    lib.lbm355_desc_size.argtypes = []
    lib.lbm355_desc_size.restype = c_int64
    if lib.lbm355_desc_size() != ctypes.sizeof(SweepDesc):
        raise ImportError("lbm355_sweep_desc layout mismatch between include/lbm355.h and the ctypes binding")
    lib.lbm355_last_error.argtypes = []
    lib.lbm355_last_error.restype = ctypes.c_char_p
    lib.lbm355_version.argtypes = []
    lib.lbm355_version.restype = c_int32
    lib.lbm355_source_hash.argtypes = []
    lib.lbm355_source_hash.restype = ctypes.c_char_p
    if path == library_path():
        built = lib.lbm355_source_hash().decode()
        want = source_hash()
        if built != want:
            raise ImportError(f"{path} was built from other sources (library {built[:12]}, tree {want[:12]}): "
                              "rebuild it with `python -m lbmpy_amd.csrc.build` (or __graft_entry__.build())")
    if path == library_path():
        _lib = lib
    return lib


class LbmError(RuntimeError):
    pass


def check(rc):
    if rc != 0:
        raise LbmError(f"lbm355 error {rc}: {load().lbm355_last_error().decode()}")


def current_stream_handle():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
