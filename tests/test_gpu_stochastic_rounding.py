"""Stochastic rounding of FP16 storage (CreateKernelConfig(stochastic_rounding=True)).

A uniform flow at u = 0.02 in a periodic D3Q19 TRT box under a body force of 1e-6 gains momentum F per cell and
step.  Its populations (about 3e-3 from the weights) have FP16 ulps near 1e-6, so the per-step increment
3 w |c.F| (at most 1.7e-7) is below half an ulp: round-to-nearest storage drops it (the sweep warns), while
stochastic rounding keeps it on average.  The FP64 momentum sum after 200 steps must be within 1% of 200 F N
with stochastic rounding, and below 20% of it with round-to-nearest.  Runs with one seed repeat bit for bit;
another seed gives another run.
"""
import warnings

import numpy as np
import pytest
import torch

from lbmpy_amd import LBMConfig, LBStencil, Stencil, Method, create_lb_function, CreateKernelConfig, Target
from lbmpy_amd.macroscopic_value_kernels import macroscopic_values_setter
from lbmpy_amd.sweep import create_pdf_array

pytestmark = pytest.mark.gpu

SHAPE = (64, 32, 16)
FORCE = 1e-6


def _kernel(stochastic, seed=0, pattern="pull"):
    cfg = CreateKernelConfig(target=Target.GPU, data_type="float16", compute_type="float32", ghost_layers=1,
                             stochastic_rounding=stochastic, rounding_seed=seed)
    lc = LBMConfig(stencil=LBStencil(Stencil.D3Q19), method=Method.TRT, relaxation_rate=1.6,
                   force=(FORCE, 0.0, 0.0), streaming_pattern=pattern)
    with warnings.catch_warnings(record=True) as caught:
        warnings.simplefilter("always")
        k = create_lb_function(lbm_config=lc, config=cfg)
    k.set_faces(["periodic"] * 3)
    return k, [w for w in caught if issubclass(w.category, RuntimeWarning)]


def _moving_state(k, **setter_kwargs):
    """Equilibrium at rho = 1, u = (0.02, 0, 0) written by the GPU setter."""
    pdf = create_pdf_array(SHAPE, 19, 1, "float16")
    rho = torch.ones(tuple(n + 2 for n in SHAPE), dtype=torch.float64, device="cuda")
    u = torch.zeros(tuple(n + 2 for n in SHAPE) + (3,), dtype=torch.float64, device="cuda")
    u[..., 0] = 0.02
    macroscopic_values_setter(k.method, rho, u, pdf, periodicity=(True,) * 3, **setter_kwargs).compile()()
    return pdf


def _run(stochastic, seed=0, steps=200):
    k, caught = _kernel(stochastic, seed)
    src = _moving_state(k)
    dst = torch.empty_like(src)
    p0 = k.conserved(src)[1]
    for _ in range(steps):
        k(src=src, dst=dst)
        src, dst = dst, src
    torch.cuda.synchronize()
    return src, k.conserved(src)[1] - p0, caught


def test_round_to_nearest_drops_the_force_and_warns():
    steps = 200
    _, gain, caught = _run(False, steps=steps)
    expect = steps * FORCE * np.prod(SHAPE)
    print(f"round to nearest: momentum gain {gain:.4e} of {expect:.4e}", flush=True)
    assert caught, "FP16 storage with a force below half an ulp must warn"
    assert abs(gain) < 0.2 * expect


def test_stochastic_rounding_keeps_the_force():
    steps = 200
    _, gain, caught = _run(True, seed=3, steps=steps)
    expect = steps * FORCE * np.prod(SHAPE)
    print(f"stochastic rounding: momentum gain {gain:.4e} of {expect:.4e}", flush=True)
    assert not caught
    assert abs(gain - expect) <= 0.01 * expect


def test_stochastic_rounding_is_reproducible_per_seed():
    a, _, _ = _run(True, seed=11, steps=20)
    b, _, _ = _run(True, seed=11, steps=20)
    c, _, _ = _run(True, seed=12, steps=20)
    assert torch.equal(a, b)
    assert not torch.equal(a, c)


def test_stochastic_rounding_in_place_patterns():
    """AA (both kernels) with stochastic rounding gains the force like pull does."""
    steps = 200
    from lbmpy_amd import Timestep
    cfg = CreateKernelConfig(target=Target.GPU, data_type="float16", compute_type="float32", ghost_layers=1,
                             stochastic_rounding=True, rounding_seed=5)
    kernels = [create_lb_function(lbm_config=LBMConfig(stencil=LBStencil(Stencil.D3Q19), method=Method.TRT,
                                                       relaxation_rate=1.6, force=(FORCE, 0.0, 0.0),
                                                       streaming_pattern="aa", timestep=t), config=cfg)
               for t in (Timestep.ODD, Timestep.EVEN)]
    for k in kernels:
        k.set_faces(["periodic"] * 3)
    # the layout an EVEN kernel leaves (slot i holds f_inv(i)), which the first (ODD) kernel reads
    pdf = _moving_state(kernels[0], streaming_pattern="aa", previous_timestep=Timestep.EVEN)
    p0 = kernels[0].conserved(pdf)[1]
    for t in range(steps):
        kernels[t % 2](src=pdf)
    torch.cuda.synchronize()
    # after an even number of steps the slots again hold f_inv(i): momentum sums carry -c
    gain = -(kernels[0].conserved(pdf)[1] - p0)
    expect = steps * FORCE * np.prod(SHAPE)
    print(f"AA stochastic rounding: momentum gain {gain:.4e} of {expect:.4e}", flush=True)
    assert abs(gain - expect) <= 0.01 * expect
