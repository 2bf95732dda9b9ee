"""Ghost-layer periodicity between sweeps (lbmpy advanced_streaming/communication.py:21-278).

lbmpy's own kernels never wrap: a periodic axis is served by ghost layers that ``LBMPeriodicityHandling``
fills after every kernel.  The sweeps here can wrap inside their index math instead (faces 'periodic');
this module is for code written around kernels that read ghost layers (faces 'ghost', the default for
``create_lb_function`` without ``builtin_periodicity``), as lbmpy's tests and scenarios are.

``get_communication_slices`` gives lbmpy's (source, destination) slice pairs; ``LBMPeriodicityHandling``
turns them into boxes of element offsets once per array layout and copies all of them with one call of
``lbm355_copy_boxes`` (one launch per 48 boxes) on the current stream."""
import ctypes
import itertools

import numpy as np
import torch

from .. import _native as nat
from ..slicing import get_slice_before_ghost_layer, normalize_slice, shift_slice
from .utility import Timestep, get_accessor, get_timesteps, is_inplace, numeric_index, numeric_offsets

__all__ = ["get_communication_slices", "LBMPeriodicityHandling", "periodic_pdf_gpu_copy_kernel", "GpuArrays"]


def _extend_dir(direction):
    """All directions that agree with `direction` where it is non-zero (communication.py:232-241)."""
    if not direction:
        yield ()
        return
    heads = (-1, 0, 1) if direction[0] == 0 else (direction[0],)
    for h in heads:
        for rest in _extend_dir(direction[1:]):
            yield (h,) + rest


def _length_one_to_int(slices):
    """Slices of length one become their start index, so that shifting moves them as a point
    (communication.py:248-260)."""
    if isinstance(slices, int):
        return slices
    if isinstance(slices, slice):
        if slices.stop is not None and abs(slices.start - slices.stop) == 1:
            return slices.start
        if slices.stop is None and slices.start == -1:
            return -1
        return slices
    return tuple(_length_one_to_int(s) for s in slices)


def _trim(slices, direction):
    """Drops the first (direction -1) or last (+1) layer of every slice component (communication.py:263-278)."""
    out = []
    for s, d in zip(slices, direction):
        if isinstance(s, int):
            out.append(s)
        else:
            out.append(slice(s.start + 1 if d == -1 else s.start, s.stop - 1 if d == 1 else s.stop, s.step))
    return tuple(out)


def get_communication_slices(stencil, comm_stencil=None, streaming_pattern="pull", prev_timestep=Timestep.BOTH,
                             ghost_layers=1):
    """{communication direction: [(source slice, destination slice), ...]} with the slot index as the last
    slice entry (communication.py:128-183).  For each communication direction and each stencil direction
    that points the same way along its non-zero axes: the layer before the ghost layer on that side, in the
    slot the kernel that ran at `prev_timestep` writes the direction to; for patterns other than pull,
    trimmed along the direction's tangential part and moved by the kernel's write offset; copied one
    domain period plus the ghost layers further along the communication direction."""
    D = stencil.D
    if comm_stencil is None:
        comm_stencil = itertools.product(*([-1, 0, 1],) * D)
    writes = get_accessor(streaming_pattern, prev_timestep).write(None, stencil)
    dirs = [tuple(v) for v in stencil]
    result = {}
    for comm_dir in comm_stencil:
        comm_dir = tuple(comm_dir)
        if all(d == 0 for d in comm_dir):
            continue
        pairs = []
        for streaming_dir in sorted(set(_extend_dir(comm_dir)) & set(dirs)):
            d = dirs.index(streaming_dir)
            slot = numeric_index(writes[d])[0]
            src = _length_one_to_int(get_slice_before_ghost_layer(comm_dir, ghost_layers=ghost_layers, thickness=1))
            if streaming_pattern != "pull":
                tangential = tuple(s - c for s, c in zip(streaming_dir, comm_dir))
                src = shift_slice(_trim(src, tangential), numeric_offsets(writes[d]))
            dst = shift_slice(src, tuple(c * (ghost_layers + 1) for c in comm_dir))
            pairs.append((src + (slot,), dst + (slot,)))
        result[comm_dir] = pairs
    return result


def slices_to_boxes(slices, shape, strides):
    """Boxes {source offset, destination offset, extent x, y, z} (int64 rows) of (source, destination) slice
    pairs over an array of `shape` [x, y, (z,) q] with element `strides`."""
    D = len(shape) - 1
    rows = []
    for src, dst in slices:
        slot = src[-1]
        if dst[-1] != slot:
            raise ValueError("source and destination slot differ")
        s_n = normalize_slice(src[:-1], shape[:D])
        d_n = normalize_slice(dst[:-1], shape[:D])
        s_off = d_off = slot * strides[D]
        ext = [1, 1, 1]
        for a in range(D):
            s, d = s_n[a], d_n[a]
            if isinstance(s, int) != isinstance(d, int):
                raise ValueError("source and destination slices differ in shape")
            if isinstance(s, int):
                s0, d0, n = s, d, 1
            else:
                s0, d0, n = s.start, d.start, s.stop - s.start
                if d.stop - d.start != n or n <= 0 or (s.step or 1) != 1:
                    raise ValueError(f"slices {src} / {dst} differ in shape or are empty")
            if not (0 <= s0 and s0 + n <= shape[a] and 0 <= d0 and d0 + n <= shape[a]):
                raise ValueError(f"slices {src} / {dst} reach outside the array")
            s_off += s0 * strides[a]
            d_off += d0 * strides[a]
            ext[a] = n
        rows.append((s_off, d_off, ext[0], ext[1], ext[2]))
    return np.array(rows, dtype=np.int64).reshape(-1, 5)


def periodic_pdf_gpu_copy_kernel(pdf_field, src_slice, dst_slice, domain_size=None):
    """A callable copying one slice of a PDF array onto another of the same shape (communication.py:186-229):
    ``kernel(array)``.  ``pdf_field`` is accepted for lbmpy's signature and not used (the layout comes from
    the array); ``domain_size`` is not needed for the same reason."""
    if src_slice[-1] != dst_slice[-1]:
        raise ValueError("Source and Destination PDF indices must be equal")
    pair = [(tuple(src_slice), tuple(dst_slice))]

    def kernel(arr=None, **kwargs):
        if arr is None:
            if len(kwargs) != 1:
                raise TypeError("pass the PDF array")
            arr = next(iter(kwargs.values()))
        _copy(arr, slices_to_boxes(pair, tuple(arr.shape), tuple(arr.stride())))
    return kernel


_DTYPES = {torch.float64: nat.F64, torch.float32: nat.F32, torch.float16: nat.F16}


def _copy(arr, boxes):
    if not isinstance(arr, torch.Tensor) or arr.device.type != "cuda":
        raise TypeError("periodicity copies need a GPU tensor")
    if arr.dtype not in _DTYPES:
        raise TypeError(f"unsupported dtype {arr.dtype}")
    if len(boxes) == 0:
        return
    st = arr.stride()
    D = arr.dim() - 1
    span = 1 + sum((n - 1) * s for n, s in zip(arr.shape, st))
    lib = nat.load()
    nat.check(lib.lbm355_copy_boxes(_DTYPES[arr.dtype], ctypes.c_void_p(arr.data_ptr()), span, st[0], st[1],
                                    st[2] if D == 3 else 0, boxes.ctypes.data_as(ctypes.c_void_p), len(boxes),
                                    nat.current_stream_handle()))


class GpuArrays:
    """The part of pystencils' SerialDataHandling that ``LBMPeriodicityHandling`` uses: ``periodicity`` and
    the named device arrays ``gpu_arrays``, with ``swap`` for two-array patterns."""

    def __init__(self, periodicity, **arrays):
        self.periodicity = tuple(bool(p) for p in periodicity)
        self.gpu_arrays = dict(arrays)

    def swap(self, name1, name2, gpu=True):
        a = self.gpu_arrays
        a[name1], a[name2] = a[name2], a[name1]


class LBMPeriodicityHandling:
    """communication.py:21-126.  ``data_handling`` needs ``periodicity`` and ``gpu_arrays`` (pystencils'
    SerialDataHandling, or ``GpuArrays``); the array named ``pdf_field_name`` is looked up at every call, so
    swapped arrays are followed.  ``cupy_direct_copy`` is accepted and has no effect: both of lbmpy's
    ways are one library call here.  Call after every kernel with the time step of the kernel that ran."""

    def __init__(self, stencil, data_handling, pdf_field_name, streaming_pattern="pull", ghost_layers=1,
                 cupy_direct_copy=True):
        self.stencil = stencil
        self.dim = stencil.D
        self.dh = data_handling
        self.pdf_field_name = pdf_field_name
        self.ghost_layers = ghost_layers
        self.periodicity = tuple(bool(p) for p in data_handling.periodicity)
        self.streaming_pattern = streaming_pattern
        self.inplace_pattern = is_inplace(streaming_pattern)
        self.cupy_direct_copy = cupy_direct_copy

        def is_copy_direction(direction):
            if any(d != 0 and not p for d, p in zip(direction, self.periodicity)):
                return False
            return any(d != 0 for d in direction)

        copy_directions = tuple(filter(is_copy_direction, itertools.product(*([-1, 0, 1],) * self.dim)))
        self.comm_slices = []
        for t in get_timesteps(streaming_pattern):
            per_dir = get_communication_slices(stencil, comm_stencil=copy_directions,
                                               streaming_pattern=streaming_pattern, prev_timestep=t,
                                               ghost_layers=ghost_layers)
            self.comm_slices.append(list(itertools.chain.from_iterable(per_dir.values())))
        self._boxes = {}

    def boxes(self, arr, prev_timestep=Timestep.BOTH):
        key = (tuple(arr.shape), tuple(arr.stride()))
        b = self._boxes.get(key)
        if b is None:
            b = [slices_to_boxes(s, key[0], key[1]) for s in self.comm_slices]
            self._boxes[key] = b
        return b[Timestep(int(prev_timestep)).idx]

    def __call__(self, prev_timestep=Timestep.BOTH):
        arr = self.dh.gpu_arrays[self.pdf_field_name]
        _copy(arr, self.boxes(arr, prev_timestep))
