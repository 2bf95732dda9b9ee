"""Host-side sweep options that need no GPU: when FP16 storage drops a body force, the stochastic-rounding
option's validation and its per-launch seeds, and the per-cell field options' validation."""
import numpy as np
import pytest

from lbmpy_amd import LBMConfig, LBStencil, Stencil, Method, CreateKernelConfig, Target, create_lb_method
from lbmpy_amd.sweep import force_rounded_away_in_fp16, launch_seed, _mix32


def _method(force):
    return create_lb_method(LBMConfig(stencil=LBStencil(Stencil.D3Q19), method=Method.MRT, relaxation_rate=1.8,
                                      force=force, compressible=False))


def test_fp16_force_loss_criterion():
    # the D3Q19 MRT channel's 1e-6: increments 3 w F <= 1.7e-7 against half-ulps of 4.8e-7 / 2.4e-7
    assert force_rounded_away_in_fp16(_method((1e-6, 0, 0)))
    # ten times larger survives round-to-nearest
    assert not force_rounded_away_in_fp16(_method((1e-5, 0, 0)))
    assert not force_rounded_away_in_fp16(_method((0.0, 0.0, 0.0)))
    m = create_lb_method(LBMConfig(stencil=LBStencil(Stencil.D3Q19), method=Method.SRT, relaxation_rate=1.8))
    assert not force_rounded_away_in_fp16(m)


def test_criterion_against_numpy_fp16_rounding():
    """The criterion says exactly what numpy's FP16 round-to-nearest does to a population at lattice velocity
    0.01 that receives the force increment."""
    for F in (1e-7, 1e-6, 3e-6, 1e-5):
        m = _method((F, 0, 0))
        c = np.array(list(m.stencil), dtype=np.float64)
        w = np.array([float(v) for v in m.weights])
        dev16 = (3 * w * 0.01).astype(np.float16).astype(np.float64)        # a stored FP16 population
        incr = 3 * w * np.abs(c[:, 0] * F)
        sel = incr > 0
        kept = (dev16[sel] + incr[sel]).astype(np.float16).astype(np.float64) != dev16[sel]
        assert force_rounded_away_in_fp16(m) == (not kept.any())


def test_stochastic_rounding_is_an_fp16_storage_option():
    with pytest.raises(ValueError):
        CreateKernelConfig(target=Target.GPU, data_type="float32", stochastic_rounding=True)
    cfg = CreateKernelConfig(target=Target.GPU, data_type="float16", stochastic_rounding=True, rounding_seed=7)
    assert cfg.compute_type == "float32" and cfg.rounding_seed == 7


def test_launch_seeds_differ_per_launch_pattern_and_seed():
    seeds = {launch_seed(s, n, p) for s in (0, 1) for n in range(64) for p in (0, 1, 2)}
    assert len(seeds) == 2 * 64 * 3
    assert all(0 <= v < 2 ** 32 for v in seeds)


def test_mix32_is_a_bijection_on_a_sample():
    xs = np.arange(1 << 16, dtype=np.int64)
    assert len({_mix32(int(v)) for v in xs}) == len(xs)
