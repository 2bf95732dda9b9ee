"""Streaming patterns, time steps and accessors (lbmpy advanced_streaming/utility.py:18-140).

The accessors here are numeric: the (cell offset, slot) each direction is read from and written to by one
kernel, taken from the sweep library (``lbm355_pattern_access``), i.e. the tables the HIP kernels are
compiled with.  Host code that addresses PDF arrays between kernels (periodicity slices, readouts) therefore
uses exactly the kernels' addressing.  Arrays are indexed [x, y, (z,) q]."""
import ctypes

import numpy as np

from ..creationfunctions import Timestep
from .. import _native as nat

streaming_patterns = ["push", "pull", "aa", "esotwist", "esopull", "esopush"]

__all__ = ["Timestep", "streaming_patterns", "is_inplace", "get_accessor", "get_timesteps", "kernel_code",
           "numeric_index", "numeric_offsets", "inverse_dir_index", "AccessPdfValues", "PdfAccess",
           "PdfAccessor"]


def is_inplace(streaming_pattern):
    """utility.py:60-64."""
    if streaming_pattern not in streaming_patterns:
        raise ValueError("Invalid streaming pattern", streaming_pattern)
    return streaming_pattern in ("aa", "esotwist", "esopull", "esopush")


def get_timesteps(streaming_pattern):
    """utility.py:82-83: (EVEN, ODD) for in-place patterns, (BOTH,) otherwise."""
    return (Timestep.EVEN, Timestep.ODD) if is_inplace(streaming_pattern) else (Timestep.BOTH,)


def kernel_code(streaming_pattern, timestep):
    """Kernel code (LBM355_PULL ... LBM355_PUSH) of the kernel of a pattern at a time step; in-place
    patterns need EVEN or ODD (utility.py:67-79 raises for BOTH)."""
    if streaming_pattern not in streaming_patterns:
        raise ValueError("Invalid value of parameter 'streaming_pattern'.", streaming_pattern)
    if streaming_pattern == "pull":
        return nat.PULL
    if streaming_pattern == "push":
        return nat.PUSH
    if int(timestep) == int(Timestep.BOTH):
        raise ValueError(f"Invalid timestep for streaming pattern {streaming_pattern}: {timestep}")
    return nat.INPLACE_EVEN[streaming_pattern] + (1 if int(timestep) == int(Timestep.ODD) else 0)


_TABLES = {}


def _stencil_code(stencil):
    code = nat.STENCIL_NAMES.get(stencil.name)
    if code is None or not stencil.is_gpu_layout():
        raise ValueError(f"no streaming tables for {stencil.name} in this direction order "
                         "(D2Q9, D3Q19 and D3Q27 in 'walberla' order)")
    return code


def access_table(stencil, code):
    """int32 array [Q][8]: read offset (x, y, z), read slot, write offset (x, y, z), write slot."""
    key = (stencil.name, int(code))
    t = _TABLES.get(key)
    if t is None:
        t = np.zeros((stencil.Q, 8), dtype=np.int32)
        rc = nat.load().lbm355_pattern_access(_stencil_code(stencil), int(code),
                                              t.ctypes.data_as(ctypes.c_void_p))
        if rc != stencil.Q:
            raise nat.LbmError(f"lbm355_pattern_access: {nat.load().lbm355_last_error().decode()}")
        t.setflags(write=False)
        _TABLES[key] = t
    return t


class PdfAccess:
    """One numeric field access: the cell offset and the (one-element) index tuple of the slot, what
    ``numeric_offsets`` / ``numeric_index`` read from pystencils' Field.Access."""
    __slots__ = ("offsets", "index")

    def __init__(self, offsets, index):
        self.offsets = tuple(int(o) for o in offsets)
        self.index = (int(index),)

    def __repr__(self):
        return f"PdfAccess(offsets={self.offsets}, index={self.index})"


class PdfAccessor:
    """The read and write accesses of one kernel (lbmpy's PdfFieldAccessor classes, fieldaccess.py:66-316).
    ``field`` is accepted for lbmpy's call form ``accessor.read(field, stencil)`` and not used."""

    def __init__(self, streaming_pattern, timestep):
        self.streaming_pattern = streaming_pattern
        self.timestep = timestep
        self.code = kernel_code(streaming_pattern, timestep)

    def read(self, field, stencil):
        t = access_table(stencil, self.code)
        return [PdfAccess(t[i, :stencil.D], t[i, 3]) for i in range(stencil.Q)]

    def write(self, field, stencil):
        t = access_table(stencil, self.code)
        return [PdfAccess(t[i, 4:4 + stencil.D], t[i, 7]) for i in range(stencil.Q)]


def get_accessor(streaming_pattern, timestep):
    """utility.py:67-79."""
    return PdfAccessor(streaming_pattern, timestep)


def numeric_offsets(field_access):
    return tuple(int(o) for o in field_access.offsets)


def numeric_index(field_access):
    return tuple(int(i) for i in field_access.index)


def inverse_dir_index(stencil, direction):
    return stencil.index(tuple(-d for d in stencil[direction]))


class AccessPdfValues:
    """Reads and writes populations of an array by direction, through a kernel's read ('in') or write
    ('out') accesses (utility.py:94-140).  Works on numpy arrays and torch tensors indexed [x, y, (z,) q]."""

    def __init__(self, stencil, streaming_pattern="pull", timestep=Timestep.BOTH, streaming_dir="out",
                 accessor=None):
        if streaming_dir not in ("in", "out"):
            raise ValueError("Invalid streaming direction.", streaming_dir)
        if accessor is None:
            accessor = get_accessor(streaming_pattern, timestep)
        self.accs = accessor.read(None, stencil) if streaming_dir == "in" else accessor.write(None, stencil)

    def _at(self, pos, d):
        a = self.accs[d]
        return tuple(p + o for p, o in zip(pos, a.offsets)) + a.index

    def write_pdf(self, pdf_arr, pos, d, value):
        pdf_arr[self._at(pos, d)] = value

    def read_pdf(self, pdf_arr, pos, d):
        return pdf_arr[self._at(pos, d)]

    def read_multiple(self, pdf_arr, indices):
        """PDF values for index tuples (x, y, [z,] dir)."""
        return np.array([float(self.read_pdf(pdf_arr, tuple(idx[:-1]), int(idx[-1]))) for idx in indices])

    def collect_from_index_list(self, pdf_arr, index_list):
        """Values for a boundary index list: a structured array with fields x, y, (z,) dir, or rows
        (x, y, [z,] dir)."""
        names = getattr(getattr(index_list, "dtype", None), "names", None)
        if names:
            fields = [f for f in ("x", "y", "z") if f in names] + ["dir"]
            rows = (tuple(int(r[f]) for f in fields) for r in index_list)
        else:
            rows = (tuple(int(v) for v in r) for r in index_list)
        return self.read_multiple(pdf_arr, rows)
