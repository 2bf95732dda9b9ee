"""Boundaries for the GPU step driver: NoSlip, UBB, FreeSlip, FixedDensity and ExtrapolationOutflow.

Mirrors lbmpy ``boundaries/boundaryconditions.py`` (``NoSlip`` :104-139, ``UBB`` :724-842,
``FreeSlip`` :360-487, ``FixedDensity`` :1044-1102, ``ExtrapolationOutflow`` :894-1042) and the index-list
machinery of ``boundaries/boundaryhandling.py:15-199``:
boundary cells are flagged on a (ghost-layer inclusive) flag field with the id of their boundary
object; every link from a fluid cell x into a boundary cell in direction d becomes an index-list
entry (x, y, z, d) of that object, executed by ``lbm355_boundary_links`` between sweeps.

For ``LatticeBoltzmannStep`` (``LatticeBoltzmannBoundaryHandling.for_domain``) periodic axes wrap inside the
sweeps: links are found across them with wrapped coordinates, and the kernel wraps them the same way.
Whole NoSlip planes on the ghost layer of a non-periodic axis (the usual channel walls,
``slice_from_direction``) are handled inside the sweep there (``LBM355_FACE_NOSLIP``), which reads the
bounce-back value directly and saves the separate launch; both give identical results.  In lbmpy's form
(constructed from a data handling, called as ``bh(prev_timestep=...)`` in a user's loop) the kernels read
ghost layers filled by ``LBMPeriodicityHandling``, so links are taken from the ghost-layer flags as they are
and nothing wraps or is fused, as in lbmpy.
"""
import numpy as np

from . import _native as nat


class LbBoundary:
    """Base of the boundary objects (boundaryconditions.py:24-99): a name (the class name unless given),
    hashed by name, equal to another object of the same type with the same attributes."""
    inner_or_boundary = True
    kind = None

    def __init__(self, name=None, calculate_force_on_boundary=False):
        self._name = name
        self.calculate_force_on_boundary = calculate_force_on_boundary

    @property
    def name(self):
        return self._name if self._name else type(self).__name__

    @name.setter
    def name(self, value):
        self._name = value

    def __hash__(self):
        return hash(self.name)

    def __eq__(self, other):
        return isinstance(other, type(self)) and self.__dict__ == other.__dict__

    def __repr__(self):
        return self.name


class NoSlip(LbBoundary):
    """Half-way bounce-back: f_inv(i)(x_b, t + 1) = f*_i(x_b, t)."""
    kind = nat.BC_NOSLIP


class UBB(LbBoundary):
    """Velocity bounce-back: f_inv(d)(x, t + 1) = f*_d(x, t) - 2 w_d rho_w (c_d . u_w) / c_s^2.

    ``velocity``: the wall velocity, a tuple, or a callable.  A callable of one argument is called as lbmpy calls
    it, with the links' boundary data (``BoundaryData``: fields 'x', 'y', ('z',) 'dir', methods
    ``link_positions`` etc.), and sets ``data['vel_0']`` ...; a callable of more arguments receives the fluid
    cells' interior coordinates of the links (arrays x, y[, z]) and the link directions and returns an
    (n, dim) array.
    ``density``: wall density; if None it is computed from the fluid cell (compressible methods) or 1."""
    kind = nat.BC_UBB

    def __init__(self, velocity, density=None, adapt_velocity_to_force=False, dim=None, name=None,
                 data_type='double'):
        # with a force model, adapt_velocity_to_force sets the wall to the equilibrium input velocity of the
        # conserved-quantity computation, u_w - F / (2 rho_ref) (boundaryconditions.py:808-812,
        # conservedquantitycomputation.py:251-266): rho_ref is the link's density for compressible methods (the
        # UBB's own, or the fluid cell's from its populations, boundaryconditions.py:829-840) and 1 otherwise
        self.adapt_velocity_to_force = bool(adapt_velocity_to_force)
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
    """Specular reflection (boundaryconditions.py:360-487).  With ``normal_direction`` (a straight wall; the
    normal points from the wall into the fluid): f_inv(d)(x_b) = f*_{inv(mirror(d))}(x + c_d + n).  Without it,
    every link gets its own normal and reference direction from the boundary cells around it, as lbmpy's
    init_callback computes them (``freeslip_link_data``), and f_inv(d)(x_b) = f*_r(x + c_d + n)."""

    def __init__(self, stencil, normal_direction=None, name=None):
        self.stencil = stencil
        if normal_direction is not None:
            normal_direction = tuple(int(n) for n in normal_direction)
            if len(normal_direction) - normal_direction.count(0) != 1 or \
                    any(n not in (-1, 0, 1) for n in normal_direction):
                raise ValueError("the normal direction must have exactly one entry -1 or 1, e.g. (0, 1, 0)")
        self.normal_direction = normal_direction
        self.dim = len(stencil[0])
        super().__init__(name)

    @property
    def kind(self):
        return nat.BC_FREESLIP if self.normal_direction is not None else nat.BC_FREESLIP_LINKS


def freeslip_link_data(links, stencil, dim):
    """Per link (x, y, (z,) d) of one FreeSlip object without a given normal: (n_x, n_y, n_z, r) with n the normal
    from the wall to the fluid and r the index of the reference direction (boundaryconditions.py:402-440).  A link's
    wall axes are those along which the fluid cell's neighbour, offset by that component of c_d alone, is one of
    the object's boundary cells; n is minus c_d on those axes, and r is c_d mirrored on them and inverted.  When no
    axis qualifies (a convex corner), n = -c_d and r = d, plain bounce-back."""
    c = np.array(stencil.stencil_entries, dtype=int)
    index = {tuple(v): i for i, v in enumerate(c.tolist())}
    cells = links[:, :dim].astype(int)
    dirs = links[:, 3].astype(int)
    boundary = set(map(tuple, (cells + c[dirs, :dim]).tolist()))
    out = np.zeros((len(links), 4))
    for k in range(len(links)):
        cd = c[dirs[k], :dim]
        normal = np.zeros(dim, dtype=int)
        ref = cd.copy()
        for i in range(dim):
            sub = np.zeros(dim, dtype=int)
            sub[i] = cd[i]
            if tuple((cells[k] + sub).tolist()) in boundary:
                normal[i] = cd[i]
                ref[i] = -ref[i]
        if not normal.any():
            normal = cd.copy()
        else:
            ref = -ref
        out[k, :dim] = -normal
        out[k, 3] = index[tuple(ref.tolist())]
    return out


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


def _takes_boundary_data(callback):
    """lbmpy's UBB callbacks take one argument, the boundary data; this package's take coordinates and
    directions."""
    import inspect
    try:
        params = inspect.signature(callback).parameters.values()
    except (TypeError, ValueError):
        return False
    positional = [p for p in params if p.kind in (p.POSITIONAL_ONLY, p.POSITIONAL_OR_KEYWORD)]
    if any(p.kind == p.VAR_POSITIONAL for p in params):
        return False
    return len([p for p in positional if p.default is p.empty]) == 1


class BoundaryData:
    """What an lbmpy-style UBB velocity callback receives (pystencils' boundary data setter, which lbmpy's
    UBB.additional_data_init_callback is called with): the links of one boundary object as record fields 'x',
    'y', ('z',) 'dir' (ghost-layer coordinates of the fluid cells) and the velocity fields 'vel_0' ... to be set,
    plus the positions of the fluid-cell midpoints, of the links' wall crossings and of the boundary cells,
    in the midpoint coordinates the mask callbacks see."""

    def __init__(self, links, stencil, ghost_layers, dim):
        self._links = links
        self._c = np.array(stencil.stencil_entries)
        self._g = ghost_layers
        self.dim = dim
        self._vel = {f"vel_{a}": np.zeros(len(links)) for a in range(dim)}

    def __len__(self):
        return len(self._links)

    def __getitem__(self, key):
        axes = {"x": 0, "y": 1, "z": 2}
        if key in axes and axes[key] < self.dim:
            return self._links[:, axes[key]].copy()
        if key == "dir":
            return self._links[:, 3].copy()
        if key in self._vel:
            return self._vel[key]
        raise KeyError(key)

    def __setitem__(self, key, value):
        if key not in self._vel:
            raise KeyError(f"{key}: a UBB callback sets {', '.join(self._vel)}")
        self._vel[key][...] = value

    def link_offsets(self):
        return self._c[self._links[:, 3], :self.dim]

    def fluid_cell_positions(self, coord):
        return self._links[:, coord] - self._g + 0.5

    def link_positions(self, coord):
        return self.fluid_cell_positions(coord) + 0.5 * self._c[self._links[:, 3], coord]

    def boundary_cell_positions(self, coord):
        return self.fluid_cell_positions(coord) + self._c[self._links[:, 3], coord]


class SimpleExtrapolationOutflow(LbBoundary):
    """Outflow by copying along the normal (boundaryconditions.py:844-891, Geier et al. 2015 eq. F.1):
    f_inv(d)(x_b) = f*_inv(d)(x + c_d - n), with ``normal_direction`` pointing out of the fluid (a vector or a
    compass string such as 'E')."""
    kind = nat.BC_SIMPLE_OUTFLOW

    def __init__(self, normal_direction, stencil, name=None):
        dim = len(stencil[0])
        if isinstance(normal_direction, str):
            v = [0, 0, 0]
            for ch in normal_direction:
                v = [a + b for a, b in zip(v, _COMPASS[ch])]
            normal_direction = tuple(v[:dim])
        self.normal_direction = tuple(int(n) for n in normal_direction)
        if any(n not in (-1, 0, 1) for n in self.normal_direction):
            raise ValueError("Only -1, 0 and 1 allowed for defining the normal direction")
        self.dim = dim
        super().__init__(name or f"Simple Outflow: {self.normal_direction}")


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
    """Flag field (boundary-object ids) + index lists for one domain (ghost-layer inclusive coordinates), and
    the launches of ``lbm355_boundary_links`` that apply them (boundaries/boundaryhandling.py:15-199).

    lbmpy's form, for loops written around ``create_lb_function`` (as lbmpy's tests are):
    ``LatticeBoltzmannBoundaryHandling(lb_method, data_handling, pdf_field_name, streaming_pattern='pull')``,
    ``set_boundary(...)`` and, between kernels, ``bh(prev_timestep=...)`` with the time step of the kernel that
    ran last.  ``data_handling`` needs ``periodicity`` and ``gpu_arrays`` (pystencils' SerialDataHandling or
    ``lbmpy_amd.advanced_streaming.GpuArrays``).  As in lbmpy, the kernels of such a loop read ghost layers and
    periodicity comes from ghost copies (``LBMPeriodicityHandling``, called before the boundaries): links are
    never wrapped and no NoSlip face is fused into a sweep.  ``target`` and ``openmp`` are accepted; the
    boundaries always run on the GPU.

    ``LatticeBoltzmannStep`` builds its own with ``for_domain``: periodic axes are wrapped inside the sweeps
    and links across them are wrapped the same way."""

    def __init__(self, lb_method, data_handling, pdf_field_name, streaming_pattern="pull", name="boundary_handling",
                 flag_interface=None, target=None, openmp=False):
        from .advanced_streaming.utility import is_inplace
        is_inplace(streaming_pattern)                                   # validates the name
        arr = data_handling.gpu_arrays[pdf_field_name]
        g = 1
        if hasattr(data_handling, "ghost_layers_of_field"):
            g = int(data_handling.ghost_layers_of_field(pdf_field_name))
        D = lb_method.dim
        self._setup(lb_method, tuple(int(n) - 2 * g for n in arr.shape[:D]),
                    tuple(bool(p) for p in data_handling.periodicity)[:D], g, builtin_periodicity=False)
        self.name = name
        self._dh = data_handling
        self._pdf_name = pdf_field_name
        self._pattern = streaming_pattern

    @classmethod
    def for_domain(cls, method, domain_size, periodicity, ghost_layers=1):
        """The step driver's form: periodic axes wrap inside the sweeps."""
        self = cls.__new__(cls)
        self._setup(method, tuple(domain_size), tuple(periodicity), ghost_layers, builtin_periodicity=True)
        self.name = "boundary_handling"
        self._dh = self._pdf_name = None
        self._pattern = None
        return self

    def _setup(self, method, domain_size, periodicity, ghost_layers, builtin_periodicity):
        self.method = method
        self.dim = len(domain_size)
        self.domain_size = tuple(domain_size)
        self.periodicity = tuple(periodicity)
        self.ghost_layers = ghost_layers
        self.builtin_periodicity = builtin_periodicity
        self.shape = tuple(domain_size)                                 # lbmpy's: the domain without ghost layers
        self.flag_shape = tuple(n + 2 * ghost_layers for n in domain_size)   # index [x, y, (z)], ghosts included
        self.flags = np.zeros(self.flag_shape, dtype=np.int16)          # 0 fluid, k: self._objects[k - 1]
        self._objects = []
        self._dirty = True
        self._groups = None
        self._faces = None
        self.version = 0           # incremented whenever the link lists are rebuilt
        self._fuse = builtin_periodicity
        self._links_dev = None
        self._links_version = None
        self._fd_table = None

    @property
    def fuse_noslip_faces(self):
        """Whether full NoSlip planes become fused sweep faces ('noslip') rather than boundary links."""
        return self._fuse

    @fuse_noslip_faces.setter
    def fuse_noslip_faces(self, value):
        if value and not self.builtin_periodicity:
            raise ValueError("fused NoSlip faces belong to the step driver's sweeps")
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
        """Flags the cells of `slice_obj` (ghost-layer inclusive indices) or, with `mask_callback`, those of them
        where the callback, given the cells' midpoint coordinates, returns True (boundaryhandling.py:218-266).
        Returns the boundary object's flag value (its id here, not a bit).  A cell holds one boundary: the
        last one set wins."""
        if not isinstance(boundary_obj, (NoSlip, UBB, FreeSlip, FixedDensity, ExtrapolationOutflow,
                                         SimpleExtrapolationOutflow)):
            raise NotImplementedError(f"boundary {boundary_obj!r} is not available on the GPU path (NoSlip, UBB, "
                                      "FreeSlip, FixedDensity, ExtrapolationOutflow and SimpleExtrapolationOutflow are)")
        if getattr(boundary_obj, "dim", self.dim) != self.dim:
            raise ValueError(f"boundary dimension {boundary_obj.dim} does not match the domain ({self.dim})")
        bid = self._id_of(boundary_obj)
        if slice_obj is None:
            slice_obj = (slice(None),) * self.dim
        view = self.flags[slice_obj]
        if mask_callback is not None:
            # cell midpoints, ghost layers included: interior cell i of an axis sits at i + 0.5 (pystencils'
            # midpoint_arrays, which lbmpy's set_boundary passes to the callback)
            g = self.ghost_layers
            coords = np.meshgrid(*[np.arange(n) - g + 0.5 for n in self.flag_shape], indexing="ij")
            sub = tuple(c[slice_obj] for c in coords)
            mask = np.asarray(mask_callback(*sub), dtype=bool)
            view[mask] = bid
        else:
            view[...] = bid
        self._dirty = True
        return bid

    def get_mask(self, slice_obj=None, boundary_name="domain", inverse=False):
        """Interior mask of the fluid ('domain'), of one boundary object (the object or its name), or of every
        boundary cell (None)."""
        g = self.ghost_layers
        inner = self.flags[(slice(g, -g),) * self.dim] if g else self.flags
        if isinstance(boundary_name, str) and boundary_name == "domain":
            mask = inner == 0
        elif boundary_name is None:
            mask = inner != 0
        else:
            ids = [k + 1 for k, o in enumerate(self._objects)
                   if o is boundary_name or o == boundary_name or getattr(o, "name", None) == boundary_name]
            if not ids:
                raise ValueError(f"no boundary {boundary_name!r} is set")
            mask = np.isin(inner, ids)
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
        is_noslip = np.isin(self.flags, noslip_ids) if noslip_ids else np.zeros(self.flag_shape, dtype=bool)
        faces = []
        for a in range(D):
            if not self.builtin_periodicity:
                faces.append("ghost")
                continue
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
        fluid = np.zeros(self.flag_shape, dtype=bool)
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
            if isinstance(obj, FreeSlip) and obj.normal_direction is None:
                vel = freeslip_link_data(links, st, D)
            if isinstance(obj, UBB) and obj.velocity_is_callable:
                if _takes_boundary_data(obj._velocity):
                    data = BoundaryData(links, st, g, D)
                    obj._velocity(data)
                    v = np.stack([data[f"vel_{a}"] for a in range(D)], axis=1)
                else:
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

    # ------------------------------------------------------------------ execution
    def _device_groups(self, device):
        """Device copies of each boundary object's links (rebuilt when the flags change):
        [object, links, per-link data or None, ready]."""
        import torch
        groups = self.link_groups
        if self._links_version != self.version:
            self._links_dev = []
            c = np.array(self.method.stencil.stencil_entries)
            for obj, l, v in groups:
                if isinstance(obj, (ExtrapolationOutflow, SimpleExtrapolationOutflow)) and \
                        np.any(c[l[:, 3]] @ np.array(obj.normal_direction[:self.dim]) <= 0):
                    raise ValueError(f"{obj.name}: the normal direction {obj.normal_direction} must point out of the "
                                     "fluid, into the outflow cells")
                data = None if v is None else torch.from_numpy(np.ascontiguousarray(v)).to(device)
                if isinstance(obj, ExtrapolationOutflow):
                    data = torch.zeros((len(l), 2), dtype=torch.float64, device=device)  # set before first use
                self._links_dev.append([obj, torch.from_numpy(np.ascontiguousarray(l)).to(device), data,
                                        not isinstance(obj, ExtrapolationOutflow)])
            self._links_version = self.version
        return self._links_dev

    def apply(self, pdf, pattern_code):
        """Applies every boundary object's links to `pdf` (a [x, y, (z,) q] tensor with this domain's ghost
        layers) after the kernel with code `pattern_code` (LBM355_PULL ... LBM355_PUSH), on the current
        stream."""
        import ctypes
        import torch
        groups = self._device_groups(pdf.device)
        if not groups:
            return
        D = self.dim
        m = self.method
        if tuple(pdf.shape[:D]) != self.flag_shape:
            raise ValueError(f"PDF array {tuple(pdf.shape)} does not match the flag field {self.flag_shape}")
        dtypes = {torch.float64: nat.F64, torch.float32: nat.F32, torch.float16: nat.F16}
        if pdf.dtype not in dtypes or not pdf.is_cuda:
            raise TypeError("boundaries run on GPU tensors of float64, float32 or float16")
        lib = nat.load()
        faces = self._faces
        for group in groups:
            obj, links, vel, ready = group
            d = nat.BoundaryDesc()
            d.stencil = nat.STENCIL_NAMES[m.stencil.name]
            d.pattern = int(pattern_code)
            d.store_dtype = dtypes[pdf.dtype]
            d.kind = obj.kind
            sizes = list(self.domain_size) + [1] * (3 - D)
            d.nx, d.ny, d.nz = sizes
            d.ghost_layers = self.ghost_layers
            for ax in range(3):
                d.periodic[ax] = int(ax < D and self.builtin_periodicity and bool(self.periodicity[ax]))
                d.fused_noslip[ax] = int(ax < D and faces[ax] == "noslip")
            d.flags = (nat.FLAG_ZERO_CENTERED if m.zero_centered_pdfs else 0) | \
                (nat.FLAG_COMPRESSIBLE if m.compressible else 0)
            d.stride_y = pdf.stride(1)
            d.stride_z = pdf.stride(2) if D == 3 else 0
            d.stride_q = pdf.stride(D)
            if isinstance(obj, UBB):
                if not obj.velocity_is_callable:
                    for ax, v in enumerate(obj._velocity):
                        d.velocity[ax] = float(v)
                d.density = float(obj._density) if obj._density is not None else 0.0
                fm = m.force_model
                if obj.adapt_velocity_to_force and fm is not None:
                    # the kernel drives the wall at u_w - F / (2 rho_ref) with the link's density
                    d.flags |= nat.FLAG_ADAPT_VELOCITY
                    for ax in range(D):
                        d.force[ax] = float(fm.force[ax])
            elif isinstance(obj, FreeSlip) and obj.normal_direction is not None:
                for ax, v in enumerate(obj.normal_direction):
                    d.normal[ax] = int(v)
            elif isinstance(obj, SimpleExtrapolationOutflow):
                for ax, v in enumerate(obj.normal_direction):
                    d.normal[ax] = int(v)
            elif isinstance(obj, ExtrapolationOutflow):
                for ax, v in enumerate(obj.normal_direction):
                    d.normal[ax] = int(v)
                if not ready:
                    # the state starts from the populations present now (read through the same accessor)
                    d.kind = nat.BC_OUTFLOW_INIT
                    nat.check(lib.lbm355_boundary_links(ctypes.byref(d), ctypes.c_void_p(pdf.data_ptr()),
                                                        ctypes.c_void_p(links.data_ptr()), links.shape[0],
                                                        ctypes.c_void_p(vel.data_ptr()),
                                                        nat.current_stream_handle()))
                    d.kind = nat.BC_OUTFLOW
                    group[3] = True
            elif isinstance(obj, FixedDensity):
                if self._fd_table is None:
                    from .kernel_tables import fixed_density_table
                    self._fd_table = torch.from_numpy(fixed_density_table(m)).to(pdf.device)
                d.density = obj.density
                d.equilibrium = self._fd_table.data_ptr()
                fm = m.force_model
                if fm is not None:
                    if fm.equilibrium_velocity_shift:
                        d.flags |= nat.FLAG_EQ_SHIFT
                    for ax in range(D):
                        d.force[ax] = float(fm.force[ax])
            nat.check(lib.lbm355_boundary_links(ctypes.byref(d), ctypes.c_void_p(pdf.data_ptr()),
                                                ctypes.c_void_p(links.data_ptr()), links.shape[0],
                                                ctypes.c_void_p(vel.data_ptr() if vel is not None else 0),
                                                nat.current_stream_handle()))

    def __call__(self, prev_timestep=None, **kwargs):
        """lbmpy's form: applies the boundaries to the data handling's PDF array after the kernel that ran at
        `prev_timestep` (boundaryhandling.py:76-96)."""
        if self._dh is None:
            raise TypeError("this boundary handling belongs to a LatticeBoltzmannStep, which applies it itself")
        from .advanced_streaming.utility import kernel_code
        from .creationfunctions import Timestep
        if prev_timestep is None:
            prev_timestep = Timestep.BOTH
        self.apply(self._dh.gpu_arrays[self._pdf_name], kernel_code(self._pattern, prev_timestep))
