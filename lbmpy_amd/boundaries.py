"""Boundaries for the GPU step driver: NoSlip, UBB, FreeSlip, FixedDensity and ExtrapolationOutflow.

Mirrors lbmpy ``boundaries/boundaryconditions.py`` (``NoSlip`` :104-139, ``UBB`` :724-842,
``FreeSlip`` :360-487, ``FixedDensity`` :1044-1102, ``ExtrapolationOutflow`` :894-1042) and the index-list
machinery of ``boundaries/boundaryhandling.py:15-199``:
boundary cells are flagged on a (ghost-layer inclusive) flag field with the id of their boundary
object; every link from a fluid cell x into a boundary cell in direction d becomes an index-list
entry (x, y, z, d) of that object, executed by ``lbm355_boundary_links`` between sweeps.  Links
are found across periodic axes with wrapped coordinates, and the kernel wraps them the same way.
Whole NoSlip planes on the ghost layer of a non-periodic axis (the usual channel walls,
``slice_from_direction``) are instead handled inside the sweep (``LBM355_FACE_NOSLIP``), which reads
the bounce-back value directly and saves the separate launch; both give identical results.
"""
import numpy as np

from . import _native as nat


class LbBoundary:
    inner_or_boundary = True
    kind = None

    def __init__(self, name=None, calculate_force_on_boundary=False):
        self.name = name or type(self).__name__
        self.calculate_force_on_boundary = calculate_force_on_boundary

    def __repr__(self):
        return self.name


class NoSlip(LbBoundary):
    """Half-way bounce-back: f_inv(i)(x_b, t + 1) = f*_i(x_b, t)."""
    kind = nat.BC_NOSLIP

    def __hash__(self):
        return hash(type(self))

    def __eq__(self, other):
        return type(other) == type(self)


class UBB(LbBoundary):
    """Velocity bounce-back: f_inv(d)(x, t + 1) = f*_d(x, t) - 2 w_d rho_w (c_d . u_w) / c_s^2.

    ``velocity``: the wall velocity, a tuple, or a callable that receives the fluid cells' interior
    coordinates of the links (arrays x, y[, z]) and the link directions and returns an (n, dim) array.
    ``density``: wall density; if None it is computed from the fluid cell (compressible methods) or 1."""
    kind = nat.BC_UBB

    def __init__(self, velocity, density=None, adapt_velocity_to_force=False, dim=None, name=None,
                 data_type='double'):
        if adapt_velocity_to_force:
            raise NotImplementedError("adapt_velocity_to_force is not implemented on the GPU path")
        if callable(velocity) and not dim:
            raise ValueError("When using a velocity callback the dimension has to be specified with the dim parameter")
        self._velocity = velocity
        self._density = density
        self.dim = dim if callable(velocity) else len(velocity)
        self.data_type = data_type
        super().__init__(name)

    @property
    def velocity_is_callable(self):
        return callable(self._velocity)


class FreeSlip(LbBoundary):
    """Specular reflection at a straight wall with the given normal (pointing from the wall into the fluid):
    f_inv(d)(x_b) = f*_{inv(mirror(d))}(x + c_d + n).  Per-cell normals (normal_direction=None) are not
    available on the GPU path."""
    kind = nat.BC_FREESLIP

    def __init__(self, stencil, normal_direction=None, name=None):
        if normal_direction is None:
            raise NotImplementedError("FreeSlip needs normal_direction on the GPU path (straight walls)")
        normal_direction = tuple(int(n) for n in normal_direction)
        if len(normal_direction) - normal_direction.count(0) != 1 or any(n not in (-1, 0, 1) for n in normal_direction):
            raise ValueError("the normal direction must have exactly one entry -1 or 1, e.g. (0, 1, 0)")
        self.stencil = stencil
        self.normal_direction = normal_direction
        self.dim = len(normal_direction)
        super().__init__(name)


class FixedDensity(LbBoundary):
    """Prescribed density (pressure) at the boundary: f_inv(d)(x_b) = 2 f_eq,sym(d; rho_w, u) - f*_d(x), with
    f_eq,sym the part of the method's equilibrium even in u and u the fluid cell's velocity computed with
    rho_w (boundaryconditions.py:1044-1102).  Moment-family methods on the GPU path."""
    kind = nat.BC_FIXED_DENSITY

    def __init__(self, density, name=None):
        if name is None:
            name = "Fixed Density " + str(density)
        self.density = float(density)
        super().__init__(name)


_COMPASS = {"E": (1, 0, 0), "W": (-1, 0, 0), "N": (0, 1, 0), "S": (0, -1, 0), "T": (0, 0, 1), "B": (0, 0, -1)}


class ExtrapolationOutflow(LbBoundary):
    """Outflow by extrapolation along the normal with the previous step's populations
    (boundaryconditions.py:894-1042, Geier et al. 2015 eq. F.2 with u neglected): per link
    f_inv(d)(x_b) = c s0 + (1 - c) s1 with c = sqrt(1/3) dt/dx, s0 the previous step's f*_inv(d) at x + c_d - n and
    s1 the previous interpolated value.  ``normal_direction`` points out of the fluid (e.g. (1, 0, 0) or 'E').
    The state starts from the populations present when the links are built; dt == dx on the GPU path."""
    kind = nat.BC_OUTFLOW

    def __init__(self, normal_direction, lb_method, dt=1, dx=1, name=None, streaming_pattern='pull',
                 zeroth_timestep=None, initial_density=None, initial_velocity=None, data_type='double'):
        if dt != dx:
            raise NotImplementedError("the GPU outflow boundary uses dt == dx")
        if initial_density is not None or initial_velocity is not None:
            raise NotImplementedError("initial_density / initial_velocity are not available on the GPU path")
        dim = lb_method.stencil.D
        if isinstance(normal_direction, str):
            v = [0, 0, 0]
            for ch in normal_direction:
                v = [a + b for a, b in zip(v, _COMPASS[ch])]
            normal_direction = tuple(v[:dim])
        self.normal_direction = tuple(int(n) for n in normal_direction)
        if any(n not in (-1, 0, 1) for n in self.normal_direction):
            raise ValueError("Only -1, 0 and 1 allowed for defining the normal direction")
        self.dim = dim
        self.lb_method = lb_method
        super().__init__(name or f"Outflow: {normal_direction}")


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
    """Flag field (boundary-object ids) + index lists for one domain (ghost-layer inclusive coordinates)."""

    def __init__(self, method, domain_size, periodicity, ghost_layers=1):
        self.method = method
        self.dim = len(domain_size)
        self.domain_size = tuple(domain_size)
        self.periodicity = tuple(periodicity)
        self.ghost_layers = ghost_layers
        self.shape = tuple(n + 2 * ghost_layers for n in domain_size)   # index [x, y, (z)]
        self.flags = np.zeros(self.shape, dtype=np.int16)                # 0 fluid, k: self._objects[k - 1]
        self._objects = []
        self._dirty = True
        self._groups = None
        self._faces = None
        self.version = 0           # incremented whenever the link lists are rebuilt
        self._fuse = True

    @property
    def fuse_noslip_faces(self):
        """Whether full NoSlip planes become fused sweep faces ('noslip') rather than boundary links."""
        return self._fuse

    @fuse_noslip_faces.setter
    def fuse_noslip_faces(self, value):
        if bool(value) != self._fuse:
            self._fuse = bool(value)
            self._dirty = True

    def _id_of(self, obj):
        for k, o in enumerate(self._objects):
            if o is obj or (isinstance(o, NoSlip) and o == obj):
                return k + 1
        self._objects.append(obj)
        return len(self._objects)

    def set_boundary(self, boundary_obj, slice_obj=None, mask_callback=None, ghost_layers=True,
                     inner_ghost_layers=True, replace=True, force_flag_value=None):
        if not isinstance(boundary_obj, (NoSlip, UBB, FreeSlip, FixedDensity, ExtrapolationOutflow)):
            raise NotImplementedError(f"boundary {boundary_obj!r} is not available on the GPU path "
                                      "(NoSlip, UBB, FreeSlip, FixedDensity and ExtrapolationOutflow are)")
        if getattr(boundary_obj, "dim", self.dim) != self.dim:
            raise ValueError(f"boundary dimension {boundary_obj.dim} does not match the domain ({self.dim})")
        bid = self._id_of(boundary_obj)
        if slice_obj is None:
            slice_obj = (slice(None),) * self.dim
        view = self.flags[slice_obj]
        if mask_callback is not None:
            g = self.ghost_layers
            coords = np.meshgrid(*[np.arange(n) - g for n in self.shape], indexing="ij")
            sub = tuple(c[slice_obj] for c in coords)
            mask = np.asarray(mask_callback(*sub), dtype=bool)
            view[mask] = bid
        else:
            view[...] = bid
        self._dirty = True

    def get_mask(self, slice_obj=None, boundary_name="domain", inverse=False):
        g = self.ghost_layers
        inner = self.flags[(slice(g, -g),) * self.dim] if g else self.flags
        mask = inner == 0 if boundary_name == "domain" else inner != 0
        if inverse:
            mask = ~mask
        return mask if slice_obj is None else mask[slice_obj]

    # ------------------------------------------------------------------ kernel inputs
    def _analyse(self):
        if not self._dirty:
            return
        g = self.ghost_layers
        D = self.dim
        noslip_ids = [k + 1 for k, o in enumerate(self._objects) if isinstance(o, NoSlip)]
        is_noslip = np.isin(self.flags, noslip_ids) if noslip_ids else np.zeros(self.shape, dtype=bool)
        faces = []
        for a in range(D):
            if self.periodicity[a]:
                faces.append("periodic")
                continue
            lo = [slice(None)] * D
            hi = [slice(None)] * D
            lo[a], hi[a] = 0, -1
            if self._fuse and g and is_noslip[tuple(lo)].all() and is_noslip[tuple(hi)].all():
                faces.append("noslip")
            else:
                faces.append("ghost")
        self._faces = faces
        self._groups = self._build_links(faces)
        self.version += 1
        self._dirty = False

    def _build_links(self, faces):
        """Per boundary object: (object, links (n, 4) int32 in ghost coordinates, link velocities or None)."""
        g = self.ghost_layers
        D = self.dim
        st = self.method.stencil
        c = np.array(st.stencil_entries)
        inner = tuple(slice(g, g + n) for n in self.domain_size)
        fluid = np.zeros(self.shape, dtype=bool)
        fluid[inner] = self.flags[inner] == 0
        cells = np.argwhere(fluid)                                   # ghost coordinates
        per_id = {}
        for d in range(st.Q):
            nb = cells + c[d][None, :]
            skip = np.zeros(len(cells), dtype=bool)
            look = nb.copy()
            for a in range(D):
                n = self.domain_size[a]
                outside = (nb[:, a] < g) | (nb[:, a] >= g + n)
                if faces[a] == "periodic":
                    look[:, a] = (nb[:, a] - g) % n + g
                elif faces[a] == "noslip":
                    skip |= outside
            ids = self.flags[tuple(look[:, a] for a in range(D))]
            hit = ~skip & (ids > 0)
            for bid in np.unique(ids[hit]):
                sel = cells[hit & (ids == bid)]
                block = np.zeros((len(sel), 4), dtype=np.int32)
                block[:, :D] = sel
                block[:, 3] = d
                per_id.setdefault(int(bid), []).append(block)
        groups = []
        for bid, blocks in sorted(per_id.items()):
            obj = self._objects[bid - 1]
            links = np.concatenate(blocks)
            vel = None
            if isinstance(obj, UBB) and obj.velocity_is_callable:
                coords = [links[:, a] - g for a in range(D)]
                v = np.asarray(obj._velocity(*coords, links[:, 3]), dtype=np.float64).reshape(len(links), D)
                vel = np.zeros((len(links), 3))
                vel[:, :D] = v
            groups.append((obj, links, vel))
        return groups

    @property
    def faces(self):
        self._analyse()
        return list(self._faces)

    @property
    def link_groups(self):
        self._analyse()
        return list(self._groups)

    @property
    def links(self):
        """All links of all boundary objects (x, y, z, d), ghost coordinates."""
        self._analyse()
        if not self._groups:
            return np.zeros((0, 4), dtype=np.int32)
        return np.concatenate([l for _, l, _ in self._groups])
