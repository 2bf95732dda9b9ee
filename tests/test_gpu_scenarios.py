"""Scenarios (lbmpy scenarios.py) on the GPU, against the FP64 oracle with the same boundaries:

* pressure-driven channels, the cases of lbmpy's CPU/GPU equivalence test (tests/test_cpu_gpu_equivalence.py:
  D2Q9 SRT incompressible, D2Q9 MRT compressible, D3Q27 TRT with pipe walls; 13 steps), max-abs <= 1e-12;
* a lid-driven cavity (D2Q9 SRT) and a fully periodic shear flow (tests/test_builtin_periodicity.py's
  setup), <= 1e-12;
* the central-moment cavities of tests/test_central_moment.py:15-28 stay finite for 1000 steps;
* the force-driven duct (64 x 32 x 32 D3Q19 TRT) is the hand-built step of the channel test, bit for bit;
* the velocity-driven pipe runs with a positive inflow.
"""
import numpy as np
import pytest

from lbmpy_amd import LBMConfig, LBStencil, Stencil, Method
from lbmpy_amd.boundaries import NoSlip, UBB, FixedDensity, slice_from_direction
from lbmpy_amd.geometry import get_shear_flow_velocity_field
from lbmpy_amd.scenarios import create_channel, create_fully_periodic_flow, create_lid_driven_cavity
from oracle.lbm_oracle import OracleDomain

pytestmark = pytest.mark.gpu


def _oracle(step):
    bh = step.boundary_handling
    specs = []
    for k, obj in enumerate(bh._objects):
        flags = bh.flags == k + 1
        if not flags.any():
            continue
        if isinstance(obj, NoSlip):
            spec = {"kind": "noslip"}
        elif isinstance(obj, FixedDensity):
            spec = {"kind": "fixeddensity", "density": obj.density}
        elif isinstance(obj, UBB) and not obj.velocity_is_callable:
            spec = {"kind": "ubb", "velocity": tuple(obj._velocity), "density": obj._density}
        else:
            raise NotImplementedError(obj)
        specs.append((flags, spec))
    D = step.dim
    faces = {a: ("periodic" if step.periodicity[a] else "ghost") for a in range(D)}
    dom = OracleDomain(step.method, bh.shape, faces=faces, pattern="pull", boundaries=specs)
    return dom


def _natural(step):
    """Post-collision populations of a pull step, (Q, [z,] y, x)."""
    D = step.dim
    a = step.pdf_array[(slice(1, -1),) * D].cpu().numpy()
    return np.transpose(a, (D,) + tuple(reversed(range(D))))


def _compare(step, steps):
    dom = _oracle(step)
    dom.set_pdfs_natural(_natural(step))
    step.run(steps)
    for _ in range(steps):
        dom.step()
    fluid = ~step.boundary_handling.get_mask(None, None).T
    got, want = _natural(step)[:, fluid], dom.pdfs_in_natural_layout()[:, fluid]
    assert np.abs(got - want).max() <= 1e-12


CHANNELS = [((17, 12), "D2Q9", Method.SRT, False), ((18, 20), "D2Q9", Method.MRT, True),
            ((7, 11, 18), "D3Q27", Method.TRT, False)]


@pytest.mark.parametrize("case", CHANNELS, ids=lambda c: f"{c[1]}-{c[2].name}")
def test_pressure_driven_channel_matches_oracle(case):
    ds, st, meth, comp = case
    lc = LBMConfig(stencil=LBStencil(Stencil[st]), method=meth, compressible=comp,
                   relaxation_rates=[1.95, 1.9, 1.92, 1.92])
    step = create_channel(domain_size=ds, pressure_difference=0.001, lbm_config=lc)
    start = _natural(step)
    _compare(step, 13)
    assert np.abs(_natural(step) - start).max() > 1e-6
    assert np.isfinite(step.velocity_slice()).all()


def test_lid_driven_cavity_matches_oracle():
    step = create_lid_driven_cavity((14, 12), lid_velocity=0.05, method=Method.SRT, relaxation_rate=1.6)
    _compare(step, 30)
    u = step.velocity[:, :]
    assert u[7, -1, 0] > 0.005                                   # dragged along by the lid


def test_fully_periodic_flow_matches_oracle_and_builtin_periodicity():
    np.random.seed(1)
    u0 = get_shear_flow_velocity_field((16, 16), 0.05, 0.1)
    a = create_fully_periodic_flow(initial_velocity=u0, relaxation_rate=1.8)
    b = create_fully_periodic_flow(initial_velocity=u0, relaxation_rate=1.8, periodicity_in_kernel=True)
    _compare(a, 20)
    b.run(20)
    np.testing.assert_array_equal(a.velocity[:, :], b.velocity[:, :])


def test_central_moment_cavities_stay_finite():
    sc2 = create_lid_driven_cavity((20, 20), method=Method.CENTRAL_MOMENT, relaxation_rate=1.8,
                                   equilibrium_order=4, compressible=True, force=(-1e-10, 0))
    sc3 = create_lid_driven_cavity((20, 20, 3), method=Method.CENTRAL_MOMENT, relaxation_rate=1.8,
                                   equilibrium_order=4, compressible=True, force=(-1e-10, 0, 0))
    sc2.run(1000)
    sc3.run(1000)
    assert np.isfinite(np.max(sc2.velocity[:, :])) and np.isfinite(np.max(sc3.velocity[:, :, :]))


def test_force_driven_duct_equals_hand_built_step():
    from lbmpy_amd import LatticeBoltzmannStep
    kw = dict(stencil=LBStencil(Stencil.D3Q19), method=Method.TRT, relaxation_rates=[1.97, 1.6], compressible=True)
    sc = create_channel(domain_size=(64, 32, 32), force=1e-7, duct=True, **kw)
    ref = LatticeBoltzmannStep(domain_size=(64, 32, 32), periodicity=(True, False, False), force=(1e-7, 0, 0), **kw)
    for d in ("N", "S", "T", "B"):
        ref.boundary_handling.set_boundary(NoSlip(), slice_from_direction(d, 3))
    sc.run(50)
    ref.run(50)
    np.testing.assert_array_equal(sc.pdf_array.cpu().numpy(), ref.pdf_array.cpu().numpy())
    assert sc.velocity[32, 16, 16][0] > 0


def test_velocity_driven_pipe_runs():
    sc = create_channel(domain_size=(24, 12, 12), u_max=0.02, method=Method.SRT, relaxation_rate=1.6)
    sc.run(40)
    u = sc.velocity_slice((2, slice(None), slice(None)))
    assert np.isfinite(u).all() and u[6, 6, 0] > 0
