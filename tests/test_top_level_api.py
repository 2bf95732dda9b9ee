"""The package's top level offers lbmpy's top-level names (src/lbmpy/__init__.py:1-56) that belong to the sweep
and its callers; get_weights gives lbmpy's weight tables (tests/test_weights.py: sum 1, isotropy)."""
from fractions import Fraction

import numpy as np
import pytest

import lbmpy_amd
from lbmpy_amd import LBStencil, Stencil

LBMPY_TOP_LEVEL = ["create_lb_ast", "create_lb_collision_rule", "create_lb_function", "create_lb_method",
                   "create_lb_update_rule", "LBMConfig", "LBMOptimisation", "Stencil", "Method", "ForceModel",
                   "CollisionSpace", "SubgridScaleModel", "LatticeBoltzmannStep", "pdf_initialization_assignments",
                   "macroscopic_values_getter", "strain_rate_tensor_getter", "compile_macroscopic_values_getter",
                   "compile_macroscopic_values_setter", "get_weights", "relaxation_rate_from_lattice_viscosity",
                   "lattice_viscosity_from_relaxation_rate", "relaxation_rate_from_magic_number",
                   "create_lid_driven_cavity", "create_fully_periodic_flow", "LBStencil"]


@pytest.mark.parametrize("name", LBMPY_TOP_LEVEL)
def test_lbmpy_top_level_name_exists(name):
    assert name in lbmpy_amd.__all__
    assert getattr(lbmpy_amd, name) is not None


@pytest.mark.parametrize("stencil", ["D2Q9", "D3Q19", "D3Q27"])
def test_get_weights(stencil):
    st = LBStencil(Stencil[stencil])
    w = lbmpy_amd.get_weights(st)
    assert sum(w) == 1 and all(isinstance(x, Fraction) for x in w)
    c = np.array(st.stencil_entries, dtype=float)
    wf = np.array([float(x) for x in w])
    np.testing.assert_allclose(c.T @ wf, 0, atol=1e-15)
    np.testing.assert_allclose((c.T * wf) @ c, np.eye(st.D) / 3, atol=1e-15)
    with pytest.warns(UserWarning):
        lbmpy_amd.get_weights(st, c_s_sq=Fraction(1, 2))
