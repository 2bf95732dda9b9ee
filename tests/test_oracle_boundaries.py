"""Pins of the oracle's UBB and FreeSlip boundaries (oracle.lbm_oracle.OracleDomain.apply_links).

* UBB with zero wall velocity is NoSlip, exactly, for pull and AA;
* Couette flow: NoSlip bottom, UBB lid moving at U, periodic x -> the linear profile
  u(y) = U (y + 1/2) / H at steady state (half-way walls at y = -1/2 and y = H - 1/2);
* FreeSlip: a force-driven half channel (NoSlip bottom, FreeSlip top) equals the lower half of the full,
  mirror-symmetric NoSlip channel after every step, and reaches u(y) = F/(2 nu) y_b (2H - y_b),
  y_b = y + 1/2 (the half of a Poiseuille parabola with zero gradient at the free-slip plane).
A planted mistake (FreeSlip without the mirror, i.e. bounce-back) fails the symmetry pin.
"""
import numpy as np
import pytest

from lbmpy_amd import LBMConfig, LBStencil, Stencil, Method, create_lb_method
from oracle.lbm_oracle import OracleDomain
from tests.lbm_helpers import seeded_state


def _plane(shape_g, axis, index):
    f = np.zeros(shape_g, dtype=bool)
    sl = [slice(None)] * len(shape_g)
    sl[axis] = index
    f[tuple(sl)] = True
    return f


def _run(m, shape, boundaries, pattern, f0, steps):
    dom = OracleDomain(m, shape, faces={0: "periodic", 1: "ghost"}, pattern=pattern, boundaries=boundaries)
    if pattern == "aa":
        dom.set_pdfs_interior(f0[dom.inv])
    else:
        dom.set_pdfs_interior(f0)
    for _ in range(steps):
        dom.step()
    return dom.pdfs_in_natural_layout(), dom


@pytest.mark.parametrize("pattern", ["pull", "aa"])
def test_ubb_with_zero_velocity_is_noslip(pattern):
    m = create_lb_method(LBMConfig(stencil=LBStencil(Stencil.D2Q9), method=Method.TRT, relaxation_rates=[1.7, 1.2],
                                   compressible=True, streaming_pattern=pattern))
    shape = (6, 5)
    gs = (8, 7)
    f0 = seeded_state(m, shape, seed=41)
    walls = _plane(gs, 1, 0) | _plane(gs, 1, -1)
    a, _ = _run(m, shape, [(walls, {"kind": "noslip"})], pattern, f0, 10)
    b, _ = _run(m, shape, [(walls, {"kind": "ubb", "velocity": (0.0, 0.0), "density": None})], pattern, f0, 10)
    np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("pattern", ["pull", "aa"])
def test_couette_profile(pattern):
    U = 0.02
    m = create_lb_method(LBMConfig(stencil=LBStencil(Stencil.D2Q9), method=Method.SRT, relaxation_rate=1.0,
                                   streaming_pattern=pattern))
    H = 12
    shape = (3, H)
    gs = (5, H + 2)
    f0 = seeded_state(m, shape, seed=0, du=0.0, drho=0.0, noise=0.0)
    bnd = [(_plane(gs, 1, 0), {"kind": "noslip"}), (_plane(gs, 1, -1), {"kind": "ubb", "velocity": (U, 0.0),
                                                                         "density": None})]
    f, dom = _run(m, shape, bnd, pattern, f0, 3000)
    ux = (f * dom.c[:, 0][:, None, None]).sum(0)            # (y, x), incompressible: u = j
    want = U * (np.arange(H) + 0.5) / H
    assert np.abs(ux[:, 1] - want).max() <= 1e-9 * U


def _half_channel(pattern, freeslip_fn="freeslip", steps=40):
    F = 1e-5
    kw = dict(stencil=LBStencil(Stencil.D2Q9), method=Method.TRT, relaxation_rates=[1.4, 1.1], force=(F, 0.0),
              streaming_pattern=pattern)
    m = create_lb_method(LBMConfig(**kw))
    H = 6
    full = (3, 2 * H)
    f0_full = seeded_state(m, full, seed=3, du=0.0, drho=0.0, noise=0.0)
    gs_full = (5, 2 * H + 2)
    walls = _plane(gs_full, 1, 0) | _plane(gs_full, 1, -1)
    ref, _ = _run(m, full, [(walls, {"kind": "noslip"})], pattern, f0_full, steps)
    gs = (5, H + 2)
    bnd = [(_plane(gs, 1, 0), {"kind": "noslip"}), (_plane(gs, 1, -1), {"kind": freeslip_fn, "normal": (0, -1)})]
    half, _ = _run(m, (3, H), bnd, pattern, f0_full[:, :H], steps)
    return ref[:, :H], half, m


@pytest.mark.parametrize("pattern", ["pull", "aa"])
def test_freeslip_equals_mirror_symmetric_channel(pattern):
    ref, half, _ = _half_channel(pattern)
    assert np.abs(ref - half).max() <= 1e-15


def test_planted_mistake_freeslip_as_noslip():
    ref, half, _ = _half_channel("pull", freeslip_fn="noslip")
    assert np.abs(ref - half).max() > 1e-8


def test_freeslip_half_parabola():
    F = 1e-5
    m = create_lb_method(LBMConfig(stencil=LBStencil(Stencil.D2Q9), method=Method.TRT, relaxation_rate=1.0,
                                   force=(F, 0.0)))
    H = 8
    gs = (5, H + 2)
    bnd = [(_plane(gs, 1, 0), {"kind": "noslip"}), (_plane(gs, 1, -1), {"kind": "freeslip", "normal": (0, -1)})]
    f0 = seeded_state(m, (3, H), seed=0, du=0.0, drho=0.0, noise=0.0)
    f, dom = _run(m, (3, H), bnd, "pull", f0, 3000)
    ux = (f * dom.c[:, 0][:, None, None]).sum(0)[:, 1] - F / 2.0      # post-collision -> physical velocity
    nu = (1.0 - 0.5) / 3.0
    yb = np.arange(H) + 0.5
    want = F / (2 * nu) * yb * (2 * H - yb)
    assert np.abs(ux - want).max() <= 1e-6 * want.max()


def _fd_setup(pattern, rho_in, rho_out, shape=(20, 10), compressible=False, omega=1.0):
    m = create_lb_method(LBMConfig(stencil=LBStencil(Stencil.D2Q9), method=Method.SRT, relaxation_rate=omega,
                                   compressible=compressible, streaming_pattern=pattern))
    gs = (shape[0] + 2, shape[1] + 2)
    walls = _plane(gs, 1, 0) | _plane(gs, 1, -1)
    inl = _plane(gs, 0, 0) & ~walls
    outl = _plane(gs, 0, -1) & ~walls
    bnd = [(walls, {"kind": "noslip"}), (inl, {"kind": "fixeddensity", "density": rho_in}),
           (outl, {"kind": "fixeddensity", "density": rho_out})]
    return m, bnd


@pytest.mark.parametrize("pattern", ["pull", "aa"])
def test_fixed_density_keeps_rest_state(pattern):
    m, bnd = _fd_setup(pattern, 1.0, 1.0)
    f0 = seeded_state(m, (20, 10), seed=0, du=0.0, drho=0.0, noise=0.0)
    dom = OracleDomain(m, (20, 10), faces={0: "ghost", 1: "ghost"}, pattern=pattern, boundaries=bnd)
    dom.set_pdfs_interior(f0[dom.inv] if pattern == "aa" else f0)
    for _ in range(20):
        dom.step()
    assert np.abs(dom.pdfs_in_natural_layout() - f0).max() <= 1e-15


def test_fixed_density_pressure_driven_channel():
    """Poiseuille flow driven by FixedDensity planes (incompressible, Delta rho = 2e-3): at steady state the
    mid-channel profile is u(y) = G / (2 nu) y_b (H - y_b) with G = c_s^2 |d rho / dx| the density gradient
    the flow has there, the gradient is uniform along the channel, the mean gradient carries the imposed
    density drop over the channel length (within 10%: entrance effects steepen it near the ends), and the
    density deviation is antisymmetric about the channel middle (rho_in - 1 = 1 - rho_out)."""
    nx, H = 24, 10
    m, bnd = _fd_setup("pull", 1.001, 0.999, shape=(nx, H))
    f0 = seeded_state(m, (nx, H), seed=0, du=0.0, drho=0.0, noise=0.0)
    dom = OracleDomain(m, (nx, H), faces={0: "ghost", 1: "ghost"}, pattern="pull", boundaries=bnd)
    dom.set_pdfs_interior(f0)
    for _ in range(5000):
        dom.step()
    f = dom.pdfs_in_natural_layout()
    ux = (f * dom.c[:, 0][:, None, None]).sum(0)
    rho = 1.0 + f.sum(0)                                            # zero-centred storage
    rmid = rho[H // 2]
    grad = np.gradient(rmid)[nx // 4: 3 * nx // 4]
    assert np.abs(grad - grad.mean()).max() <= 0.02 * abs(grad.mean())
    G = abs(grad.mean()) / 3.0
    nu = (1.0 - 0.5) / 3.0
    yb = np.arange(H) + 0.5
    want = G / (2 * nu) * yb * (H - yb)
    np.testing.assert_allclose(ux[:, nx // 2], want, rtol=0.02)
    assert abs(abs(grad.mean()) * nx - 0.002) <= 0.1 * 0.002
    np.testing.assert_allclose(rmid - 1.0, -(rmid[::-1] - 1.0), atol=5e-6)


@pytest.mark.parametrize("pattern", ["pull", "aa"])
def test_outflow_passes_uniform_flow(pattern):
    """UBB inlet (W) at U and extrapolation outflow (E), periodic y: the uniform equilibrium flow at U is a steady
    state of the scheme -- the outflow's extrapolation reproduces it exactly."""
    U = 0.03
    m = create_lb_method(LBMConfig(stencil=LBStencil(Stencil.D2Q9), method=Method.TRT, relaxation_rates=[1.6, 1.2],
                                   streaming_pattern=pattern))
    nx, ny = 12, 6
    gs = (nx + 2, ny + 2)
    bnd = [(_plane(gs, 0, 0), {"kind": "ubb", "velocity": (U, 0.0), "density": None}),
           (_plane(gs, 0, -1), {"kind": "outflow", "normal": (1, 0)})]
    from oracle.lbm_oracle import equilibrium_pdfs
    u = np.zeros((2, ny, nx))
    u[0] = U
    f0 = equilibrium_pdfs(m, np.ones((ny, nx)), u)
    dom = OracleDomain(m, (nx, ny), faces={0: "ghost", 1: "periodic"}, pattern=pattern, boundaries=bnd)
    dom.set_pdfs_interior(f0[dom.inv] if pattern == "aa" else f0)
    for _ in range(30):
        dom.step()
    assert np.abs(dom.pdfs_in_natural_layout() - f0).max() <= 1e-15


def _uniform_flow_with(pattern, outlet_spec, steps=30):
    U = 0.03
    m = create_lb_method(LBMConfig(stencil=LBStencil(Stencil.D2Q9), method=Method.TRT, relaxation_rates=[1.6, 1.2],
                                   streaming_pattern=pattern))
    nx, ny = 12, 6
    gs = (nx + 2, ny + 2)
    bnd = [(_plane(gs, 0, 0), {"kind": "ubb", "velocity": (U, 0.0), "density": None}), (_plane(gs, 0, -1), outlet_spec)]
    from oracle.lbm_oracle import equilibrium_pdfs
    u = np.zeros((2, ny, nx))
    u[0] = U
    f0 = equilibrium_pdfs(m, np.ones((ny, nx)), u)
    dom = OracleDomain(m, (nx, ny), faces={0: "ghost", 1: "periodic"}, pattern=pattern, boundaries=bnd)
    dom.set_pdfs_interior(f0[dom.inv] if pattern == "aa" else f0)
    for _ in range(steps):
        dom.step()
    return np.abs(dom.pdfs_in_natural_layout() - f0).max()


@pytest.mark.parametrize("pattern", ["pull", "aa"])
def test_simple_outflow_passes_uniform_flow(pattern):
    """SimpleExtrapolationOutflow (boundaryconditions.py:844-891) copies along the normal, so the uniform
    equilibrium flow through a UBB inlet stays exact; copying from the outflow cell itself instead of the fluid
    cell upstream (a planted mistake: normal 0) breaks it."""
    assert _uniform_flow_with(pattern, {"kind": "simple_outflow", "normal": (1, 0)}) <= 1e-15
    assert _uniform_flow_with(pattern, {"kind": "simple_outflow", "normal": (0, 0)}, steps=4) > 1e-6
