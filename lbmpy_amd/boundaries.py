"""NoSlip boundaries for the GPU step driver.

Mirrors lbmpy ``boundaries/boundaryconditions.py:104-139`` (``NoSlip``) and the
index-list machinery of ``boundaries/boundaryhandling.py:15-199``: boundary
cells are flagged on a (ghost-layer inclusive) flag field; every link from a
fluid cell x into a flagged cell in direction d becomes an index-list entry
(x, y, z, d) executed by the ``lbm355_noslip_links`` kernel between sweeps.
Whole wall planes on the ghost layer of a non-periodic axis (the usual channel
walls, ``slice_from_direction``) are instead handled inside the sweep
(``LBM355_FACE_NOSLIP``), which reads the bounce-back value directly and saves
the separate launch; both give identical results.
"""
import numpy as np


class LbBoundary:
    inner_or_boundary = True

    def __init__(self, name=None, calculate_force_on_boundary=False):
        self.name = name or type(self).__name__
        self.calculate_force_on_boundary = calculate_force_on_boundary

    def __repr__(self):
        return self.name


class NoSlip(LbBoundary):
    """Half-way bounce-back: f_inv(i)(x_b, t + 1) = f*_i(x_b, t)."""

    def __hash__(self):
        return hash(type(self))

    def __eq__(self, other):
        return type(other) == type(self)


def slice_from_direction(direction_name, dim, normal_offset=0, tangential_offset=0):
    """Slice of the ghost layer in a compass direction ('N', 'S', 'E', 'W', 'T', 'B'),
    as pystencils' ``slice_from_direction`` returns it for ghost-layer inclusive arrays."""
    if tangential_offset != 0:
        raise NotImplementedError("tangential offsets are not supported")
    axis = {"W": 0, "E": 0, "S": 1, "N": 1, "B": 2, "T": 2}
    result = [slice(None)] * dim
    for ch in direction_name:
        a = axis[ch]
        if a >= dim:
            raise ValueError(f"direction {direction_name} needs {a + 1} dimensions")
        result[a] = (-1 - normal_offset) if ch in "ENT" else normal_offset
    return tuple(result)


class LatticeBoltzmannBoundaryHandling:
    """Flag field + NoSlip links for one domain (ghost-layer inclusive coordinates)."""

    def __init__(self, method, domain_size, periodicity, ghost_layers=1):
        self.method = method
        self.dim = len(domain_size)
        self.domain_size = tuple(domain_size)
        self.periodicity = tuple(periodicity)
        self.ghost_layers = ghost_layers
        self.shape = tuple(n + 2 * ghost_layers for n in domain_size)   # index [x, y, (z)]
        self.flags = np.zeros(self.shape, dtype=bool)
        self._dirty = True
        self._links = None
        self._faces = None

    def set_boundary(self, boundary_obj, slice_obj=None, mask_callback=None, ghost_layers=True,
                     inner_ghost_layers=True, replace=True, force_flag_value=None):
        if not isinstance(boundary_obj, NoSlip):
            raise NotImplementedError(f"boundary {boundary_obj!r} is not available on the GPU path (NoSlip is)")
        if slice_obj is None:
            slice_obj = (slice(None),) * self.dim
        if mask_callback is not None:
            g = self.ghost_layers
            coords = np.meshgrid(*[np.arange(n) - g for n in self.shape], indexing="ij")
            sub = tuple(c[slice_obj] for c in coords)
            mask = np.asarray(mask_callback(*sub), dtype=bool)
            view = self.flags[slice_obj]
            view |= mask
        else:
            self.flags[slice_obj] = True
        self._dirty = True

    def get_mask(self, slice_obj=None, boundary_name="domain", inverse=False):
        g = self.ghost_layers
        inner = self.flags[(slice(g, -g),) * self.dim] if g else self.flags
        mask = ~inner if boundary_name == "domain" else inner
        if inverse:
            mask = ~mask
        return mask if slice_obj is None else mask[slice_obj]

    # ------------------------------------------------------------------ kernel inputs
    def _analyse(self):
        if not self._dirty:
            return
        g = self.ghost_layers
        D = self.dim
        faces = []
        remaining = self.flags.copy()
        for a in range(D):
            if self.periodicity[a]:
                faces.append("periodic")
                continue
            lo = [slice(None)] * D
            hi = [slice(None)] * D
            lo[a], hi[a] = 0, -1
            if g and remaining[tuple(lo)].all() and self.flags[tuple(hi)].all():
                faces.append("noslip")
            else:
                faces.append("ghost")
        # walls fully handled by fused faces are removed from the index list
        for a in range(D):
            if faces[a] == "noslip":
                lo = [slice(None)] * D
                hi = [slice(None)] * D
                lo[a], hi[a] = 0, -1
                remaining[tuple(lo)] = False
                remaining[tuple(hi)] = False
        self._faces = faces
        self._links = self._build_links(remaining, faces)
        self._dirty = False

    def _build_links(self, walls, faces):
        g = self.ghost_layers
        D = self.dim
        st = self.method.stencil
        c = np.array(st.stencil_entries)
        if not walls.any():
            return np.zeros((0, 4), dtype=np.int32)
        inner = tuple(slice(g, g + n) for n in self.domain_size)
        fluid = np.zeros_like(self.flags)
        fluid[inner] = ~self.flags[inner]
        cells = np.argwhere(fluid)                                   # ghost coordinates
        rows = []
        for d in range(st.Q):
            nb = cells + c[d][None, :]
            skip = np.zeros(len(cells), dtype=bool)
            for a in range(D):
                n = self.domain_size[a]
                outside = (nb[:, a] < g) | (nb[:, a] >= g + n)
                if faces[a] == "periodic":
                    wrapped = nb.copy()
                    wrapped[:, a] = (nb[:, a] - g) % n + g
                    crosses = outside & walls[tuple(wrapped[:, b] for b in range(D))]
                    if crosses.any():
                        # the link kernel addresses the neighbour as x + c_d without wrapping
                        raise NotImplementedError("NoSlip cells touching a periodic face are not supported")
                    skip |= outside
                elif faces[a] == "noslip":
                    skip |= outside
            hit = ~skip & walls[tuple(nb[:, a] for a in range(D))]
            sel = cells[hit]
            if len(sel):
                block = np.zeros((len(sel), 4), dtype=np.int32)
                block[:, :D] = sel
                block[:, 3] = d
                rows.append(block)
        return np.concatenate(rows) if rows else np.zeros((0, 4), dtype=np.int32)

    @property
    def faces(self):
        self._analyse()
        return list(self._faces)

    @property
    def links(self):
        self._analyse()
        return self._links
