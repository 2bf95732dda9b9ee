"""Slice helpers of pystencils (``pystencils.slicing``, pystencils >= 1.3 as pinned by lbmpy's
pyproject.toml:14), which lbmpy's periodicity handling and its tests are written with.  pystencils is a
third-party dependency that is not part of this package; these functions restate its published behaviour
for integer slices over ghost-layer-inclusive arrays:

* ``get_slice_before_ghost_layer(d)``: the last ``thickness`` interior layers on the side ``d`` points to,
  interior only along axes where ``d`` is 0 (all cells, ghosts included, with ``full_slice``);
* ``get_ghost_region_slice(d)``: the ghost layers on the side ``d`` points to;
* ``shift_slice``: every integer bound moved by an offset (``None`` bounds stay open);
* ``normalize_slice``: negative and open bounds made absolute for the given sizes.
"""


class _SliceMaker:
    def __getitem__(self, item):
        return item


make_slice = _SliceMaker()


def get_slice_before_ghost_layer(direction, ghost_layers=1, thickness=None, full_slice=False):
    thickness = thickness or ghost_layers
    inner = 0 if full_slice else ghost_layers
    out = []
    for c in direction:
        if c == -1:
            out.append(slice(ghost_layers, ghost_layers + thickness))
        elif c == 0:
            out.append(slice(inner, -inner if inner else None))
        elif c == 1:
            out.append(slice(-ghost_layers - thickness, -ghost_layers if ghost_layers else None))
        else:
            raise ValueError("Invalid direction: only -1, 0, 1 components are allowed")
    return tuple(out)


def get_ghost_region_slice(direction, ghost_layers=1, thickness=None, full_slice=False):
    thickness = thickness or ghost_layers
    if thickness > ghost_layers:
        raise ValueError("thickness exceeds the number of ghost layers")
    inner = 0 if full_slice else ghost_layers
    out = []
    for c in direction:
        if c == -1:
            out.append(slice(ghost_layers - thickness, ghost_layers))
        elif c == 0:
            out.append(slice(inner, -inner if inner else None))
        elif c == 1:
            stop = -ghost_layers + thickness
            out.append(slice(-ghost_layers, stop if stop else None))
        else:
            raise ValueError("Invalid direction: only -1, 0, 1 components are allowed")
    return tuple(out)


def _shift(comp, off):
    if comp is None or isinstance(comp, float):
        return comp
    if isinstance(comp, int):
        return comp + off
    if isinstance(comp, slice):
        return slice(_shift(comp.start, off), _shift(comp.stop, off), comp.step)
    raise ValueError(f"cannot shift {comp!r}")


def shift_slice(slices, offset):
    if hasattr(offset, "__len__"):
        return tuple(_shift(s, o) for s, o in zip(slices, offset))
    if isinstance(slices, (slice, int, float)):
        return _shift(slices, offset)
    return tuple(_shift(s, offset) for s in slices)


def normalize_slice(slices, sizes):
    if len(slices) != len(sizes):
        raise ValueError("Slice dimension does not match sizes")
    out = []
    for s, n in zip(slices, sizes):
        if isinstance(s, int):
            out.append(s + n if s < 0 else s)
            continue
        if isinstance(s, float):
            out.append(int(s * n))
            continue
        out.append(slice(_bound(s.start, n, 0), _bound(s.stop, n, n), 1 if s.step is None else s.step))
    return tuple(out)


def _bound(b, n, default):
    """One slice bound made absolute: open -> default, float -> that fraction of n, negative -> from the end."""
    if b is None:
        return default
    if isinstance(b, float):
        return int(b * n)
    return b + n if b < 0 else b
