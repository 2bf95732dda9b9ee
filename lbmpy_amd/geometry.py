"""Geometry helpers for boundary setup and initial fields (lbmpy geometry.py:8-192).

Masks are evaluated at cell midpoints, ghost layers included (interior cell i of an axis at i + 0.5), as lbmpy's
boundary handling passes them; ``boundary_handling.shape`` is the domain without ghost layers.  The pipe inflow
profile is evaluated at the links' wall positions (fluid-cell midpoint plus half the link direction), where
lbmpy's UBB callback evaluates it (``link_positions``)."""
import numpy as np

from .boundaries import NoSlip, UBB, slice_from_direction

__all__ = ["add_box_boundary", "add_sphere", "add_pipe_inflow_boundary", "add_pipe_walls", "get_pipe_velocity_field",
           "get_shear_flow_velocity_field"]


def add_box_boundary(boundary_handling, boundary=NoSlip(), replace=True):
    """`boundary` on every face of the domain (geometry.py:8-26)."""
    borders = ["N", "S", "E", "W"] + (["T", "B"] if boundary_handling.dim == 3 else [])
    flag = None
    for d in borders:
        flag = boundary_handling.set_boundary(boundary, slice_from_direction(d, boundary_handling.dim),
                                              replace=replace)
    return flag


def add_sphere(boundary_handling, midpoint, radius, boundary=NoSlip(), replace=True):
    """`boundary` on the cells whose midpoints lie inside the sphere (disc in 2-D) (geometry.py:29-33)."""
    def inside(*coords):
        return sum((c - m) ** 2 for c, m in zip(coords, midpoint)) < radius ** 2
    return boundary_handling.set_boundary(boundary, mask_callback=inside, replace=replace)


def _radius(shape, flow_direction, diameter):
    if diameter is None:
        return int(round(min(n for i, n in enumerate(shape) if i != flow_direction) / 2))
    return int(round(diameter / 2))


def add_pipe_inflow_boundary(boundary_handling, u_max, slice_obj, flow_direction=0, diameter=None):
    """A UBB on `slice_obj` with the parabolic pipe profile u_max (1 - (r / R)^2) along `flow_direction`
    (geometry.py:36-85): r is the distance of the link's wall position from the pipe axis, measured from
    round(shape / 2) across the flow in 3-D and from R in 2-D."""
    dim = boundary_handling.dim
    shape = boundary_handling.shape
    radius = _radius(shape, flow_direction, diameter)
    c = np.array(boundary_handling.method.stencil.stencil_entries)

    def profile(*args):
        coords, dirs = args[:dim], np.asarray(args[dim])
        wall = [np.asarray(coords[a], dtype=np.float64) + 0.5 + 0.5 * c[dirs, a] for a in range(dim)]
        if dim == 3:
            n1, n2 = (flow_direction + 1) % 3, (flow_direction + 2) % 3
            r = np.sqrt((wall[n1] - int(round(shape[n1] / 2))) ** 2 + (wall[n2] - int(round(shape[n2] / 2))) ** 2)
        elif dim == 2:
            r = np.abs(wall[(flow_direction + 1) % 2] - radius)
        else:
            raise ValueError("Invalid dimension")
        v = np.zeros((len(dirs), dim))
        v[:, flow_direction] = u_max * (1 - (r / radius) ** 2)
        return v

    return boundary_handling.set_boundary(UBB(profile, dim=dim), slice_obj=slice_obj, ghost_layers=True)


def add_pipe_walls(boundary_handling, diameter, boundary=NoSlip()):
    """`boundary` outside a pipe along x of the given diameter (a number, or a callback (x line, domain shape)
    -> diameter per x) centred on shape // 2 (geometry.py:88-121)."""
    shape = boundary_handling.shape
    dim = len(shape)
    if dim not in (2, 3):
        raise ValueError("pipes are 2-D or 3-D")

    def outside(*coords):
        if callable(diameter):
            line = coords[0][:, 0, 0] if dim == 3 else coords[0][:, 0]
            d = np.asarray(diameter(line, shape))
            d = d[:, None, None] if dim == 3 else d[:, None]
        else:
            d = diameter
        centre = [shape[i] // 2 for i in range(1, dim)]
        dist_sq = sum((c - m) ** 2 for c, m in zip(coords[1:], centre))
        return dist_sq > (d / 2) ** 2

    return boundary_handling.set_boundary(boundary, mask_callback=outside)


def get_pipe_velocity_field(domain_size, u_max, flow_direction=0, diameter=None):
    """The parabolic pipe profile at the cell midpoints of `domain_size` (geometry.py:124-163); values outside
    the pipe are not meaningful."""
    radius = _radius(domain_size, flow_direction, diameter)
    grid = np.meshgrid(*[np.arange(n) + 0.5 for n in domain_size], indexing="ij")
    dist_sq = 0
    for i, n in enumerate(domain_size):
        if i != flow_direction:
            dist_sq = dist_sq + (grid[i] - int(round(n / 2))) ** 2
    u = np.zeros(tuple(domain_size) + (len(domain_size),))
    u[..., flow_direction] = u_max * (1 - dist_sq / radius ** 2)
    return u


def get_shear_flow_velocity_field(domain_size, u_max, random_y_factor=0.01):
    """Two shear layers: x velocity u_max, -u_max in the middle third of the last axis, and a random transverse
    part of amplitude u_max * random_y_factor to break the symmetry (geometry.py:166-190)."""
    height = domain_size[2] if len(domain_size) == 3 else domain_size[1]
    velocity = np.empty(tuple(domain_size) + (len(domain_size),))
    velocity[..., 0] = u_max
    velocity[..., height // 3: height // 3 * 2, 0] = -u_max
    for j in range(1, len(domain_size)):
        velocity[..., j] = random_y_factor * u_max * (np.random.rand(*domain_size) - 0.5)
    return velocity
