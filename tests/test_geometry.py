"""lbmpy_amd.geometry (lbmpy geometry.py) without a GPU: pipe walls, spheres, the pipe inflow profile and the
initial velocity fields, on the step driver's boundary handling (flags only)."""
import numpy as np

from lbmpy_amd import LBMConfig, LBStencil, Stencil, Method, create_lb_method
from lbmpy_amd.boundaries import LatticeBoltzmannBoundaryHandling, NoSlip, slice_from_direction
from lbmpy_amd.geometry import (add_box_boundary, add_pipe_inflow_boundary, add_pipe_walls, add_sphere,
                                get_pipe_velocity_field, get_shear_flow_velocity_field)


def _bh(shape, periodic=None):
    st = "D2Q9" if len(shape) == 2 else "D3Q19"
    m = create_lb_method(LBMConfig(stencil=LBStencil(Stencil[st]), method=Method.SRT, relaxation_rate=1.5))
    return LatticeBoltzmannBoundaryHandling.for_domain(m, shape, periodic or (False,) * len(shape))


def test_pipe_walls_2d_are_the_ghost_rows():
    bh = _bh((20, 10), (True, False))
    add_pipe_walls(bh, 10)
    inner = bh.flags[1:-1, 1:-1]
    assert not inner.any()                                       # midpoints 0.5..9.5 lie within 5 of y = 5
    assert bh.flags[:, 0].all() and bh.flags[:, -1].all()


def test_pipe_walls_3d_cross_section():
    bh = _bh((6, 10, 10), (True, False, False))
    add_pipe_walls(bh, 10)
    f = bh.flags > 0
    sec = f[2]
    assert np.array_equal(sec, sec.T) and np.array_equal(sec, sec[::-1]) and np.array_equal(sec, sec[:, ::-1])
    assert all(np.array_equal(f[i], sec) for i in range(f.shape[0]))
    assert not sec[5, 5] and sec[1, 1] and sec[0, 5]
    # cells whose midpoint lies within the radius are fluid, the others wall (4 x 4 corner block checked)
    for j in range(1, 5):
        for k in range(1, 5):
            r2 = (j - 1 + 0.5 - 5) ** 2 + (k - 1 + 0.5 - 5) ** 2
            assert sec[j, k] == (r2 > 25)


def test_sphere_and_box():
    bh = _bh((12, 12))
    add_sphere(bh, (6, 6), 3)
    inner = bh.flags[1:-1, 1:-1] > 0
    assert inner.sum() == 32 and np.array_equal(inner, inner.T) and np.array_equal(inner, inner[::-1])
    add_box_boundary(bh, NoSlip())
    assert (bh.flags[0] > 0).all() and (bh.flags[:, -1] > 0).all()


def test_pipe_inflow_profile_matches_pipe_velocity_field():
    bh = _bh((20, 10), (False, False))
    add_pipe_inflow_boundary(bh, 0.05, slice_from_direction("W", 2))
    bh.set_boundary(NoSlip(), slice_from_direction("N", 2))
    bh.set_boundary(NoSlip(), slice_from_direction("S", 2))
    groups = bh.link_groups
    ubb = [(o, l, v) for o, l, v in groups if v is not None]
    assert len(ubb) == 1
    _, links, vel = ubb[0]
    c = np.array(LBStencil(Stencil.D2Q9).stencil_entries)
    west = c[links[:, 3], 0] == -1
    assert west.all()
    straight = c[links[:, 3], 1] == 0
    field = get_pipe_velocity_field((20, 10), 0.05)
    y = links[straight, 1] - 1
    np.testing.assert_allclose(vel[straight, 0], field[0, y, 0], rtol=1e-14)
    assert np.all(vel[:, 1] == 0) and vel[:, 0].max() <= 0.05 and vel[:, 0].min() >= 0


def test_velocity_fields():
    u = get_pipe_velocity_field((8, 10, 10), 0.1)
    assert u.shape == (8, 10, 10, 3) and np.all(u[..., 1:] == 0)
    assert abs(u[0, 4, 4, 0] - 0.1 * (1 - 0.5 / 25)) < 1e-15
    np.random.seed(0)
    s = get_shear_flow_velocity_field((9, 12), 0.04)
    assert np.all(s[:, 4:8, 0] == -0.04) and np.all(s[:, :4, 0] == 0.04) and np.all(np.abs(s[..., 1]) <= 0.0002)


def test_lbmpy_style_ubb_callback_equals_coordinate_callback():
    """lbmpy's callback form (boundary data with link_positions and vel_i fields, as in lbmpy's
    add_pipe_inflow_boundary) gives the same link velocities as the coordinate form."""
    from lbmpy_amd.boundaries import UBB
    u_max, radius = 0.05, 5

    def lbmpy_style(boundary_data):
        boundary_data["vel_1"] = 0.0
        y = boundary_data.link_positions(1)
        boundary_data["vel_0"] = u_max * (1 - ((y - radius) / radius) ** 2)

    c = np.array(LBStencil(Stencil.D2Q9).stencil_entries)

    def coordinate_style(x, y, d):
        v = np.zeros((len(d), 2))
        v[:, 0] = u_max * (1 - ((y + 0.5 + 0.5 * c[d, 1] - radius) / radius) ** 2)
        return v

    vels = []
    for cb in (lbmpy_style, coordinate_style):
        bh = _bh((20, 10), (False, False))
        bh.set_boundary(UBB(cb, dim=2), slice_from_direction("W", 2))
        (_, links, vel), = [(o, l, v) for o, l, v in bh.link_groups if v is not None]
        vels.append(vel)
    np.testing.assert_array_equal(vels[0], vels[1])
    assert vels[0][:, 0].max() > 0.04
