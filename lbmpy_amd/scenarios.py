"""Predefined scenarios on ``LatticeBoltzmannStep`` (lbmpy scenarios.py:37-191): a fully periodic flow from a
velocity field, the lid-driven cavity, and channels driven by a body force, a pressure difference or a
parabolic inflow, with rectangular (duct) or circular cross-section.

The step runs on the GPU sweep; periodicity is always built into the sweep, so ``periodicity_in_kernel`` only
documents the caller's choice (lbmpy's two variants give the same flow, tests/test_builtin_periodicity.py).
pystencils data handlings and the waLBerla-parallel variants are not used here: pass ``domain_size``.

>>> from lbmpy_amd.enums import Method
>>> scenario = create_channel(domain_size=(20, 10), force=1e-5, method=Method.SRT, relaxation_rate=1.9)
>>> scenario.run(100)
"""
import numpy as np

from .boundaries import UBB, FixedDensity, NoSlip, slice_from_direction
from .geometry import add_pipe_inflow_boundary, add_pipe_walls

__all__ = ["create_fully_periodic_flow", "create_lid_driven_cavity", "create_channel"]


def _no_data_handling(data_handling, parallel):
    if data_handling is not None:
        raise NotImplementedError("pystencils data handlings are not used on the GPU path; pass the domain size")
    if parallel:
        raise NotImplementedError("the waLBerla-parallel scenarios are not part of this package")


def _step(**kwargs):
    from .lbstep import LatticeBoltzmannStep
    return LatticeBoltzmannStep(**kwargs)


def create_fully_periodic_flow(initial_velocity, periodicity_in_kernel=False, lbm_kernel=None,
                               data_handling=None, parallel=False, **kwargs):
    """A fully periodic domain of ``initial_velocity.shape[:-1]`` cells, initialised at equilibrium with that
    velocity and density 1 (scenarios.py:37-69)."""
    _no_data_handling(data_handling, parallel)
    initial_velocity = np.asarray(initial_velocity, dtype=np.float64)
    domain_size = initial_velocity.shape[:-1]
    step = _step(domain_size=domain_size, periodicity=True, name="periodic_scenario", lbm_kernel=lbm_kernel,
                 **kwargs)
    np.copyto(step.velocity[...], initial_velocity)
    step.set_pdf_fields_from_macroscopic_values()
    return step


def create_lid_driven_cavity(domain_size=None, lid_velocity=0.005, lbm_kernel=None, parallel=False,
                             data_handling=None, **kwargs):
    """A closed box whose north face (the lid) moves with ``lid_velocity`` along x: UBB on N, NoSlip on the
    other faces (scenarios.py:72-102)."""
    _no_data_handling(data_handling, parallel)
    if domain_size is None:
        raise ValueError("give the domain size")
    step = _step(domain_size=domain_size, periodicity=False, lbm_kernel=lbm_kernel, name="ldc", **kwargs)
    dim = step.dim
    step.boundary_handling.set_boundary(UBB(velocity=[lid_velocity, 0, 0][:dim]), slice_from_direction("N", dim))
    for direction in ("W", "E", "S") if dim == 2 else ("W", "E", "S", "T", "B"):
        step.boundary_handling.set_boundary(NoSlip(), slice_from_direction(direction, dim))
    return step


def create_channel(domain_size=None, force=None, pressure_difference=None, u_max=None, diameter_callback=None,
                   duct=False, wall_boundary=NoSlip(), parallel=False, data_handling=None, **kwargs):
    """A channel along x (scenarios.py:105-191), driven by exactly one of:

    * ``force``: a body force along x, periodic in x;
    * ``pressure_difference``: FixedDensity 1 + difference at the inflow (W) and 1 at the outflow (E);
    * ``u_max``: a parabolic UBB inflow (W) and FixedDensity 1 at the outflow (E).

    ``wall_boundary`` goes on the faces across the flow; without ``duct`` a circular pipe wall (of a
    ``diameter_callback``-given or the smallest cross-section's diameter) is added as well."""
    _no_data_handling(data_handling, parallel)
    if domain_size is None:
        raise ValueError("give the domain size")
    if [bool(p) for p in (force, pressure_difference, u_max)].count(True) != 1:
        raise ValueError("Please specify exactly one of the parameters 'force', 'pressure_difference' or 'u_max'")
    dim = len(domain_size)
    if dim not in (2, 3):
        raise ValueError("channels are 2-D or 3-D")
    if duct and diameter_callback is not None:
        raise ValueError("For duct flows, passing a diameter callback does not make sense.")
    if u_max and duct:
        raise NotImplementedError("Velocity inflow for duct flows not yet implemented")
    periodicity = (True, False, False)[:dim] if force else (False,) * dim
    if force:
        kwargs["force"] = tuple([force, 0, 0][:dim])
        step = _step(domain_size=domain_size, periodicity=periodicity, name="force_driven_channel", **kwargs)
    elif pressure_difference:
        step = _step(domain_size=domain_size, periodicity=periodicity, name="pressure_driven_channel", **kwargs)
        step.boundary_handling.set_boundary(FixedDensity(1.0 + pressure_difference), slice_from_direction("W", dim))
        step.boundary_handling.set_boundary(FixedDensity(1.0), slice_from_direction("E", dim))
    else:
        step = _step(domain_size=domain_size, periodicity=periodicity, name="velocity_driven_channel", **kwargs)
        diameter = diameter_callback(np.array([0]), domain_size)[0] if diameter_callback else min(domain_size[1:])
        add_pipe_inflow_boundary(step.boundary_handling, u_max, slice_from_direction("W", dim), flow_direction=0,
                                 diameter=diameter)
        step.boundary_handling.set_boundary(FixedDensity(1.0), slice_from_direction("E", dim))
    for direction in ("N", "S", "T", "B") if dim == 3 else ("N", "S"):
        step.boundary_handling.set_boundary(wall_boundary, slice_from_direction(direction, dim))
    if not duct:
        diameter = min(step.boundary_handling.shape[1:])
        add_pipe_walls(step.boundary_handling, diameter_callback if diameter_callback else diameter, wall_boundary)
    return step
