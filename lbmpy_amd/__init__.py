"""lbmpy_amd -- lbmpy's fused stream-collide sweep, rebuilt for AMD Instinct MI355X (gfx950).

The public names mirror lbmpy's (``LBMConfig``, ``LBMOptimisation``,
``create_lb_function``, ``create_lb_method``, ``LatticeBoltzmannStep``,
``Stencil``, ``Method``, ``ForceModel``, ``LBStencil`` ...) so a lbmpy setup
with ``CreateKernelConfig(target=Target.GPU)`` runs on the HIP kernels of
``csrc/`` through the C ABI in ``include/lbm355.h``.

Importing the package does not touch the GPU; the native library is loaded
when a sweep is created.
"""
from .enums import Stencil, Method, CollisionSpace, ForceModel, SubgridScaleModel, Target
from .stencils import LBStencil
from .creationfunctions import (LBMConfig, LBMOptimisation, CreateKernelConfig, Timestep,
                                create_lb_method, create_lb_collision_rule, create_lb_update_rule,
                                create_lb_ast, create_lb_function, update_with_default_parameters)
from .relaxationrates import (relaxation_rate_from_lattice_viscosity, lattice_viscosity_from_relaxation_rate,
                              relaxation_rate_from_magic_number)
from .equilibrium import get_weights
from . import forcemodels

__all__ = ["Stencil", "Method", "CollisionSpace", "ForceModel", "SubgridScaleModel", "Target", "LBStencil",
           "LBMConfig", "LBMOptimisation", "CreateKernelConfig", "Timestep", "create_lb_method",
           "create_lb_collision_rule", "create_lb_update_rule", "create_lb_ast", "create_lb_function",
           "update_with_default_parameters", "relaxation_rate_from_lattice_viscosity",
           "lattice_viscosity_from_relaxation_rate", "relaxation_rate_from_magic_number", "forcemodels", "get_weights",
           "LatticeBoltzmannStep", "pdf_initialization_assignments", "macroscopic_values_getter",
           "macroscopic_values_setter", "strain_rate_tensor_getter", "compile_macroscopic_values_getter",
           "compile_macroscopic_values_setter", "create_lid_driven_cavity", "create_fully_periodic_flow",
           "create_channel"]

# lazy like LatticeBoltzmannStep: they import torch and the native library
_MACRO = ("pdf_initialization_assignments", "macroscopic_values_getter", "macroscopic_values_setter",
          "strain_rate_tensor_getter", "compile_macroscopic_values_getter", "compile_macroscopic_values_setter")


def __getattr__(name):
    # lazy: the step driver imports torch and the native library
    if name == "LatticeBoltzmannStep":
        from .lbstep import LatticeBoltzmannStep
        return LatticeBoltzmannStep
    if name in _MACRO:
        from . import macroscopic_value_kernels
        return getattr(macroscopic_value_kernels, name)
    if name in ("create_lid_driven_cavity", "create_fully_periodic_flow", "create_channel"):
        from . import scenarios
        return getattr(scenarios, name)
    raise AttributeError(name)
