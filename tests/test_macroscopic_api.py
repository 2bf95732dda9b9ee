"""Argument handling of the GPU macroscopic setter / getter (no GPU needed): the streaming-pattern / timestep
rules of lbmpy's accessors (advanced_streaming/utility.py:52-68), the conserved-quantity names of
compile_macroscopic_values_* (macroscopic_value_kernels.py:114-118), and the parts that are not available."""
import pytest

from lbmpy_amd import LBMConfig, LBStencil, Stencil, Method, create_lb_method, Target, Timestep
from lbmpy_amd import macroscopic_value_kernels as mvk


@pytest.fixture
def method():
    return create_lb_method(LBMConfig(stencil=LBStencil(Stencil.D2Q9), method=Method.SRT, relaxation_rate=1.5))


def test_inplace_pattern_needs_even_or_odd(method):
    with pytest.raises(ValueError, match="Invalid timestep"):
        mvk.compile_macroscopic_values_getter(method, ["density"], streaming_pattern="esotwist")
    mvk.compile_macroscopic_values_getter(method, ["density"], streaming_pattern="esotwist",
                                          previous_timestep=Timestep.ODD)


def test_unknown_pattern_and_quantity(method):
    with pytest.raises(ValueError):
        mvk.compile_macroscopic_values_getter(method, ["density"], streaming_pattern="zigzag")
    with pytest.raises(ValueError, match="No such conserved quantity"):
        mvk.compile_macroscopic_values_getter(method, ["pressure"])
    with pytest.raises(ValueError, match="No such conserved quantity"):
        mvk.compile_macroscopic_values_setter(method, {"temperature": 1.0})


def test_unavailable_parts(method):
    with pytest.raises(NotImplementedError):
        mvk.compile_macroscopic_values_getter(method, ["density"], target=Target.CPU)
    with pytest.raises(NotImplementedError):
        mvk.compile_macroscopic_values_getter(method, ["density"], iteration_slice=(slice(1, 3),) * 2)
    with pytest.raises(NotImplementedError):
        mvk.macroscopic_values_setter(method, 1.0, (0.0, 0.0), None, set_pre_collision_pdfs=True)
    with pytest.raises(ValueError):
        mvk.macroscopic_values_getter(method, None, None, None)


def test_ghost_free_domain_must_be_periodic(method):
    with pytest.raises(ValueError):
        mvk.compile_macroscopic_values_setter(method, {"density": 1.0}, ghost_layers=0, periodicity=(True, False))
