"""The orthogonal MRT whose only rates off 1 are the shear moments' (lbmpy's MRT given its shear rate) relaxes them
through their projector in closed form, (9/2) w_i (c_i c_i) : Pi_dev(f_eq - f), instead of one rank-one correction
per shear moment.  It must agree with the rank-one form and with the FP64 oracle, in every storage form, on 3-D
stencils, and methods with other rate patterns must keep the rank-one form."""
import numpy as np
import pytest

from lbmpy_amd import LBMConfig, LBStencil, Stencil, Method, create_lb_method
from lbmpy_amd import _native as nat
from tests.lbm_helpers import seeded_state, oracle_run, gpu_kernels, gpu_run

pytestmark = pytest.mark.gpu


def _flags(kernel, q, dtype):
    from lbmpy_amd.sweep import create_pdf_array
    return kernel.descriptor(create_pdf_array((8, 8, 8), q, 1, dtype), {}).flags


def _kw(stencil, pattern, compressible=False):
    return dict(stencil=LBStencil(stencil), method=Method.MRT, relaxation_rate=1.8, compressible=compressible,
                force=(1e-5, -2e-5, 0.0), streaming_pattern=pattern)


@pytest.mark.parametrize("stencil", [Stencil.D3Q19, Stencil.D3Q27])
@pytest.mark.parametrize("pattern", ["pull", "aa"])
@pytest.mark.parametrize("compressible", [False, True])
def test_closed_form_against_oracle_fp64(stencil, pattern, compressible):
    kw = _kw(stencil, pattern, compressible)
    m = create_lb_method(LBMConfig(**kw))
    shape = (9, 8, 7)
    faces = {0: "periodic", 1: "noslip", 2: "periodic"}
    f0 = seeded_state(m, shape, seed=21, du=0.03)
    ref, _ = oracle_run(m, shape, faces, pattern, f0, 12)
    ks = gpu_kernels(kw)
    assert all(_flags(k, m.stencil.Q, "float64") & nat.FLAG_SHEAR_PROJECTOR for k in ks)
    print(f"MRT shear projector vs oracle: {stencil.name} {pattern} compressible={compressible}", flush=True)
    got = gpu_run(ks, m, shape, faces, f0, 12)
    assert np.abs(got - ref).max() <= 1e-12


@pytest.mark.parametrize("data_type,tol", [("float64", 1e-15), ("float32", 2e-7), ("float16", 2e-3)])
@pytest.mark.parametrize("stencil", [Stencil.D3Q19, Stencil.D3Q27])
def test_closed_form_matches_rank_one_form(stencil, data_type, tol):
    kw = _kw(stencil, "pull")
    m = create_lb_method(LBMConfig(**kw))
    shape = (64, 6, 5)
    faces = {0: "periodic", 1: "noslip", 2: "periodic"}
    f0 = seeded_state(m, shape, seed=22, du=0.04, noise=3e-4)
    fast = gpu_kernels(kw, data_type=data_type, compute_type="float32" if data_type == "float16" else None)
    slow = gpu_kernels(kw, data_type=data_type, compute_type="float32" if data_type == "float16" else None)
    for k in slow:
        k.shortcuts = False
    assert not any(_flags(k, m.stencil.Q, data_type) & nat.FLAG_SHEAR_PROJECTOR for k in slow)
    print(f"MRT shear projector vs rank-one corrections: {stencil.name} {data_type}", flush=True)
    a = gpu_run(fast, m, shape, faces, f0, 10, data_type=data_type)
    b = gpu_run(slow, m, shape, faces, f0, 10, data_type=data_type)
    assert np.abs(a - b).max() <= tol


def test_other_rate_patterns_keep_the_rank_one_form():
    q = 19
    trt = gpu_kernels(dict(stencil=LBStencil(Stencil.D3Q19), method=Method.TRT, relaxation_rates=[1.8, 1.1]))
    assert not _flags(trt[0], q, "float64") & nat.FLAG_SHEAR_PROJECTOR
    rates = [0.0] * 4 + [1.8] * 5 + [1.3] + [1.0] * 9       # a bulk rate of its own
    mrt = gpu_kernels(dict(stencil=LBStencil(Stencil.D3Q19), method=Method.MRT, relaxation_rates=rates))
    assert not _flags(mrt[0], q, "float64") & nat.FLAG_SHEAR_PROJECTOR
    d2 = gpu_kernels(dict(stencil=LBStencil(Stencil.D2Q9), method=Method.MRT, relaxation_rate=1.8))
    from lbmpy_amd.sweep import create_pdf_array
    assert not d2[0].descriptor(create_pdf_array((8, 8), 9, 1, "float64"), {}).flags & nat.FLAG_SHEAR_PROJECTOR
