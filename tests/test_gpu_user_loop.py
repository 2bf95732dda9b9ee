"""Loops written around create_lb_function, as lbmpy's own tests are, on the GPU: kernels that read ghost
layers, LBMPeriodicityHandling, LatticeBoltzmannBoundaryHandling called as bh(prev_timestep=...), and the
macroscopic setter / getter.

* lbmpy's single-cell NoSlip test (tests/advanced_streaming/test_advanced_streaming_noslip.py:20-60): after the
  boundary, the next kernel reads at the cell, in direction d, the value the previous kernel wrote for inv(d),
  for every stencil, pattern and parity;
* lbmpy's periodic pipe (tests/advanced_streaming/test_periodic_pipe_with_force.py, force-driven SRT flow in a
  NoSlip pipe, periodic along it): every pattern gives the same velocity field after 100 steps;
* a smaller pipe whose pull run matches the FP64 oracle (built-in periodicity, the same NoSlip cells) to 1e-12,
  and whose other patterns feed every next kernel bit-identical values to pull after each step.
"""
import numpy as np
import pytest
import torch

from lbmpy_amd import LBMConfig, LBStencil, Stencil, Method, ForceModel, create_lb_function, create_lb_method, \
    CreateKernelConfig, Target
from lbmpy_amd.advanced_streaming import (AccessPdfValues, GpuArrays, LBMPeriodicityHandling, get_timesteps,
                                          inverse_dir_index, is_inplace, streaming_patterns, Timestep)
from lbmpy_amd.advanced_streaming.utility import access_table, kernel_code
from lbmpy_amd.boundaries import LatticeBoltzmannBoundaryHandling, NoSlip
from lbmpy_amd.macroscopic_value_kernels import macroscopic_values_getter, macroscopic_values_setter
from lbmpy_amd.sweep import create_pdf_array
from oracle.lbm_oracle import OracleDomain

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prev_timestep", [Timestep.EVEN, Timestep.ODD])
@pytest.mark.parametrize("pattern", streaming_patterns)
@pytest.mark.parametrize("stencil", ["D2Q9", "D3Q19", "D3Q27"])
def test_noslip_single_cell(stencil, pattern, prev_timestep):
    st = LBStencil(Stencil[stencil])
    D, Q = st.D, st.Q
    prev_access = AccessPdfValues(st, pattern, prev_timestep, "out")
    next_access = AccessPdfValues(st, pattern, prev_timestep.next(), "in")
    host = create_pdf_array((1,) * D, Q, 1, "float64", device="cpu")
    pos = (1,) * D
    for d in range(Q):
        prev_access.write_pdf(host, pos, d, float(d))
    pdfs = create_pdf_array((1,) * D, Q, 1, "float64")
    pdfs.copy_(host)
    method = create_lb_method(LBMConfig(stencil=st, method=Method.SRT))
    bh = LatticeBoltzmannBoundaryHandling(method, GpuArrays((False,) * D, pdfs=pdfs), "pdfs",
                                          streaming_pattern=pattern)
    # every cell but the single interior one (midpoint 0.5 on each axis)
    bh.set_boundary(NoSlip(), mask_callback=lambda *xs: np.logical_or.reduce([np.abs(x - 0.5) > 0.25 for x in xs]))
    assert len(bh.links) == Q - 1
    bh(prev_timestep=prev_timestep)
    out = pdfs.cpu()
    reflected = [float(next_access.read_pdf(out, pos, d)) for d in range(Q)]
    assert reflected == [float(inverse_dir_index(st, d)) for d in range(Q)]


class PeriodicPipeFlow:
    """Force-driven SRT flow in a NoSlip pipe along x, periodic in x (test_periodic_pipe_with_force.py:30-175)."""

    def __init__(self, stencil, pattern, length=60, radius=15, force=1e-4):
        self.stencil, self.pattern = stencil, pattern
        D, Q = stencil.D, stencil.Q
        self.dim = D
        self.radius = radius
        self.domain_size = (length,) + (2 * radius,) * (D - 1)
        self.timesteps = get_timesteps(pattern)
        self.zeroth = self.timesteps[0]
        kw = dict(stencil=stencil, method=Method.SRT, relaxation_rate=1.0, force_model=ForceModel.GUO,
                  force=(force,) + (0.0,) * (D - 1))
        cfg = CreateKernelConfig(target=Target.GPU, data_type="float64", ghost_layers=1)
        if is_inplace(pattern):
            self.kernels = [create_lb_function(lbm_config=LBMConfig(**kw, streaming_pattern=pattern, timestep=t),
                                               config=cfg) for t in self.timesteps]
        else:
            self.kernels = [create_lb_function(lbm_config=LBMConfig(**kw, streaming_pattern=pattern), config=cfg)]
        self.method = create_lb_method(LBMConfig(**kw))
        pdfs = create_pdf_array(self.domain_size, Q, 1, "float64")
        self.dh = GpuArrays((True,) + (False,) * (D - 1), pdfs=pdfs, pdfs_tmp=torch.zeros_like(pdfs))
        self.velocity = torch.zeros(tuple(pdfs.shape[:-1]) + (D,), dtype=torch.float64, device="cuda")
        self.init_kernel = macroscopic_values_setter(self.method, 1.0, (0.0,) * D, pdfs, streaming_pattern=pattern,
                                                     previous_timestep=self.zeroth).compile()
        self.getters = [macroscopic_values_getter(self.method, None, self.velocity, pdfs, streaming_pattern=pattern,
                                                  previous_timestep=t).compile() for t in self.timesteps]
        self.periodicity = LBMPeriodicityHandling(stencil, self.dh, "pdfs", streaming_pattern=pattern)
        self.wall = NoSlip()
        self.bh = LatticeBoltzmannBoundaryHandling(self.method, self.dh, "pdfs", streaming_pattern=pattern)
        self.bh.set_boundary(boundary_obj=self.wall, mask_callback=self.mask_callback)
        self.current = self.zeroth

    def mask_callback(self, x, y, z=None):
        y = y - self.radius
        z = z - self.radius if z is not None else 0
        return np.sqrt(y ** 2 + z ** 2) >= self.radius

    def init(self):
        self.current = self.zeroth
        self.init_kernel(pdfs=self.dh.gpu_arrays["pdfs"])

    def boundaries_for_next(self):
        # order matters: first communicate, then boundaries (periodicity would overwrite reflected values)
        self.periodicity(self.current)
        self.bh(prev_timestep=self.current)

    def step(self):
        self.boundaries_for_next()
        self.current = self.current.next()
        k = self.kernels[self.current.idx] if is_inplace(self.pattern) else self.kernels[0]
        if is_inplace(self.pattern):
            k(src=self.dh.gpu_arrays["pdfs"])
        else:
            k(src=self.dh.gpu_arrays["pdfs"], dst=self.dh.gpu_arrays["pdfs_tmp"])
            self.dh.swap("pdfs", "pdfs_tmp")
        self.getters[self.current.idx](pdfs=self.dh.gpu_arrays["pdfs"])

    def run(self, n):
        for _ in range(n):
            self.step()

    def trimmed_velocity(self):
        u = self.velocity[tuple(slice(1, -1) for _ in range(self.dim))].cpu().numpy().copy()
        u[self.bh.get_mask(None, self.wall)] = np.nan
        return u


_PIPES = {}


@pytest.mark.parametrize("pattern", streaming_patterns)
@pytest.mark.parametrize("stencil", ["D2Q9", "D3Q19", "D3Q27"])
def test_periodic_pipe(stencil, pattern):
    st = LBStencil(Stencil[stencil])
    flow = PeriodicPipeFlow(st, pattern)
    flow.init()
    flow.run(100)
    u = flow.trimmed_velocity()
    fluid = ~np.isnan(u[..., 0])
    assert np.all(u[..., 0][fluid] > 0)
    # the pipe and its flow are uniform along x
    np.testing.assert_allclose(u[..., 0][fluid].reshape(u.shape[0], -1),
                               np.broadcast_to(u[0, ..., 0][fluid[0]], (u.shape[0], fluid[0].sum())), rtol=1e-12)
    for (st_name, other), prev_u in _PIPES.items():
        if st_name == stencil:
            np.testing.assert_allclose(u, prev_u, rtol=0, atol=1e-15,
                                       err_msg=f"velocity for {pattern} differs from {other}")
    _PIPES[(stencil, pattern)] = u


def _next_reads(arr, st, shape, code):
    tab = access_table(st, code)
    D = st.D
    cells = np.stack(np.meshgrid(*[np.arange(n) for n in shape], indexing="ij"), axis=-1).reshape(-1, D)
    return torch.stack([arr[tuple(torch.from_numpy(cells[:, a] + tab[i, a] + 1).cuda() for a in range(D))
                            + (int(tab[i, 3]),)] for i in range(st.Q)])


SMALL = dict(length=12, radius=5, force=1e-4)


def _oracle_layout(arr):
    """[x, y, (z,) q] -> the oracle's (q, [z,] y, x)."""
    D = arr.ndim - 1
    return np.transpose(arr, (D,) + tuple(reversed(range(D))))


@pytest.mark.parametrize("stencil", ["D2Q9", "D3Q19"])
def test_small_pipe_pull_matches_oracle(stencil):
    st = LBStencil(Stencil[stencil])
    D = st.D
    flow = PeriodicPipeFlow(st, "pull", **SMALL)
    flow.init()
    inner = tuple(slice(1, -1) for _ in range(D))
    faces = {0: "periodic", 1: "ghost", 2: "ghost"}
    dom = OracleDomain(flow.method, flow.domain_size, faces={a: faces[a] for a in range(D)}, pattern="pull",
                       boundaries=[(flow.bh.flags > 0, {"kind": "noslip"})])
    f0 = _oracle_layout(flow.dh.gpu_arrays["pdfs"][inner].cpu().numpy())
    dom.set_pdfs_natural(f0)
    for _ in range(25):
        flow.step()
        dom.step()
    got = _oracle_layout(flow.dh.gpu_arrays["pdfs"][inner].cpu().numpy())
    want = dom.pdfs_in_natural_layout()
    fluid = ~flow.bh.get_mask(None, flow.wall).T
    assert np.abs(got[:, fluid] - want[:, fluid]).max() <= 1e-12
    assert np.abs(want[:, fluid] - f0[:, fluid]).max() > 1e-6


@pytest.mark.parametrize("pattern", [p for p in streaming_patterns if p != "pull"])
@pytest.mark.parametrize("stencil", ["D2Q9", "D3Q27"])
def test_small_pipe_patterns_feed_pull_values(stencil, pattern):
    st = LBStencil(Stencil[stencil])
    ref = PeriodicPipeFlow(st, "pull", **SMALL)
    flow = PeriodicPipeFlow(st, pattern, **SMALL)
    ref.init()
    flow.init()
    shape = flow.domain_size
    fluid = torch.from_numpy(~flow.bh.get_mask(None, flow.wall).reshape(-1)).cuda()
    for step in range(6):
        ref.boundaries_for_next()
        flow.boundaries_for_next()
        a = _next_reads(ref.dh.gpu_arrays["pdfs"], st, shape, kernel_code("pull", Timestep.BOTH))
        b = _next_reads(flow.dh.gpu_arrays["pdfs"], st, shape, kernel_code(pattern, flow.current.next()))
        assert torch.equal(a[:, fluid], b[:, fluid]), (pattern, step)
        _kernel_step(ref)
        _kernel_step(flow)


def _kernel_step(flow):
    flow.current = flow.current.next()
    if is_inplace(flow.pattern):
        flow.kernels[flow.current.idx](src=flow.dh.gpu_arrays["pdfs"])
    else:
        flow.kernels[0](src=flow.dh.gpu_arrays["pdfs"], dst=flow.dh.gpu_arrays["pdfs_tmp"])
        flow.dh.swap("pdfs", "pdfs_tmp")
