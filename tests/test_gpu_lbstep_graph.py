"""LatticeBoltzmannStep.run replays a captured two-step HIP graph (the GPU form of lbmpy's fixed two-step time loop,
lbstep.py:283-309).  Every replay must give bit for bit what the same steps give when run one by one from Python:
several run() calls in a row, with the program's own eager work and synchronisations between them, parameter and
boundary changes that must trigger a new capture, an outflow boundary whose state starts on its first application,
and stochastic rounding, whose per-launch seeds keep it on the step-by-step path."""
import numpy as np
import pytest
import torch

from lbmpy_amd import LBMConfig, LBStencil, Stencil, Method, CreateKernelConfig, Target
from lbmpy_amd.boundaries import NoSlip, UBB, slice_from_direction

pytestmark = pytest.mark.gpu


def _cavity(pattern, use_graphs, data_type="float32", lc_kw=None, kernel_params=None, cfg_kw=None):
    from lbmpy_amd.lbstep import LatticeBoltzmannStep
    kw = dict(stencil=LBStencil(Stencil.D3Q19), compressible=True, streaming_pattern=pattern)
    kw.update(lc_kw or dict(method=Method.TRT, relaxation_rates=[1.8, 1.1]))
    cfg = CreateKernelConfig(target=Target.GPU, data_type=data_type, **(cfg_kw or {}))
    step = LatticeBoltzmannStep(domain_size=(24, 20, 16), periodicity=(False, False, True), lbm_config=LBMConfig(**kw),
                                config=cfg, kernel_params=kernel_params or {})
    step.use_graphs = use_graphs
    step.boundary_handling.set_boundary(UBB((0.05, 0.0, 0.0)), slice_from_direction("N", 3))
    for w in ("S", "W", "E"):
        step.boundary_handling.set_boundary(NoSlip(), slice_from_direction(w, 3))
    return step


def _same(a, b):
    torch.cuda.synchronize()
    return torch.equal(a.pdf_array, b.pdf_array) and np.array_equal(a.velocity[:, :, :], b.velocity[:, :, :])


@pytest.mark.parametrize("pattern", ["pull", "aa", "esotwist"])
def test_graph_replays_equal_step_by_step(pattern):
    g = _cavity(pattern, True)
    e = _cavity(pattern, False)
    print(f"graph replay {pattern}", flush=True)
    for n in (7, 5, 10, 1, 2, 12):
        g.run(n)
        e.run(n)
        # the program's own work between the runs: a GPU reduction, a host read of the macroscopic fields
        assert g.nonfinite_count() == 0 and e.nonfinite_count() == 0
        assert _same(g, e), f"after run({n})"
    assert g._graph is not None, "the run did not go through a captured graph"
    assert g.time_steps_run == e.time_steps_run == 37
    assert np.abs(g.velocity[:, :, :]).max() > 1e-3


def test_graph_recaptures_when_parameters_or_boundaries_change():
    import sympy as sp
    kw = dict(lc_kw=dict(method=Method.SRT, relaxation_rate=sp.Symbol("omega")), kernel_params={"omega": 1.7})
    g = _cavity("aa", True, **kw)
    e = _cavity("aa", False, **kw)
    g.run(8)
    e.run(8)
    assert _same(g, e)
    first = g._graph[1]
    for s in (g, e):
        s.kernel_params["omega"] = 1.3
    g.run(8)
    e.run(8)
    assert _same(g, e)
    assert g._graph[1] is not first
    # a new wall: the link lists are rebuilt, and so is the graph
    for s in (g, e):
        s.boundary_handling.set_boundary(NoSlip(), slice_from_direction("B", 3))
    g.run(6)
    e.run(6)
    assert _same(g, e)


def test_graph_with_outflow_state():
    from lbmpy_amd.boundaries import ExtrapolationOutflow
    steps = []
    for use in (True, False):
        s = _cavity("aa", use)
        s.boundary_handling.set_boundary(ExtrapolationOutflow((1, 0, 0), s.method), slice_from_direction("E", 3))
        steps.append(s)
    g, e = steps
    for n in (9, 4, 11):
        g.run(n)
        e.run(n)
        assert _same(g, e)
    assert g._graph is not None


def test_stochastic_rounding_runs_step_by_step():
    g = _cavity("aa", True, data_type="float16", cfg_kw=dict(stochastic_rounding=True))
    e = _cavity("aa", False, data_type="float16", cfg_kw=dict(stochastic_rounding=True))
    g.run(6)
    e.run(6)
    assert g._graph is None
    assert _same(g, e)


def test_graph_benchmark_run_counts_steps():
    s = _cavity("pull", True)
    mlups = s.benchmark_run(20)
    assert mlups > 0
    assert s.time_steps_run == 22     # the two steps run before the capture are counted too
    print(f"benchmark_run 24x20x16 D3Q19 TRT pull through the graph: {mlups:.0f} MLUPS", flush=True)
