"""The central-moment sweep with lbmpy's default methods (the continuous Maxwellian, compressible or incompressible,
given the shear rate) evaluates the equilibrium's central moments in closed form (rho gamma_e, or gamma_e + drho
times a rest Gaussian's) and the relaxation's rank-one sum too (conserved moments at rate 0, shear moments through
the deviatoric projector).  It must agree with the general
table-driven form and with the FP64 oracle, with and without the Guo force; other methods keep the general form."""
import numpy as np
import pytest

from lbmpy_amd import LBMConfig, LBStencil, Stencil, Method, ForceModel, create_lb_method
from lbmpy_amd import _native as nat
from tests.lbm_helpers import seeded_state, oracle_run, gpu_kernels, gpu_run

pytestmark = pytest.mark.gpu

D3Q27 = LBStencil(Stencil.D3Q27)
CLOSED = nat.FLAG_GAUSSIAN_EQUILIBRIUM | nat.FLAG_SHEAR_PROJECTOR


def _flags(kernel, dtype="float64"):
    from lbmpy_amd.sweep import create_pdf_array
    return kernel.descriptor(create_pdf_array((8, 8, 8), 27, 1, dtype), {}).flags


CASES = [("pull", None), ("aa", (ForceModel.GUO, (1e-5, -2e-5, 1e-5))), ("aa", (ForceModel.SIMPLE, (2e-5, 0, 0)))]


def _case_kw(pattern, forcing, compressible=True):
    kw = dict(stencil=D3Q27, method=Method.CENTRAL_MOMENT, relaxation_rate=1.8, compressible=compressible,
              streaming_pattern=pattern)
    if forcing is not None:
        kw.update(force_model=forcing[0], force=forcing[1])
    return kw


@pytest.mark.parametrize("pattern,forcing", CASES)
@pytest.mark.parametrize("compressible", [True, False])
def test_closed_forms_against_oracle_fp64(pattern, forcing, compressible):
    kw = _case_kw(pattern, forcing, compressible)
    m = create_lb_method(LBMConfig(**kw))
    shape = (8, 7, 6)
    faces = {0: "periodic", 1: "noslip", 2: "periodic"}
    f0 = seeded_state(m, shape, seed=31, du=0.03)
    ref, _ = oracle_run(m, shape, faces, pattern, f0, 10)
    ks = gpu_kernels(kw)
    assert all(_flags(k) & CLOSED == CLOSED for k in ks)
    print(f"central closed forms vs oracle: {pattern} {forcing} compressible={compressible}", flush=True)
    got = gpu_run(ks, m, shape, faces, f0, 10)
    assert np.abs(got - ref).max() <= 1e-12


@pytest.mark.parametrize("data_type,tol", [("float64", 1e-15), ("float32", 2e-7)])
@pytest.mark.parametrize("pattern,forcing", CASES)
@pytest.mark.parametrize("compressible", [True, False])
def test_closed_forms_match_general_form(pattern, forcing, data_type, tol, compressible):
    kw = _case_kw(pattern, forcing, compressible)
    m = create_lb_method(LBMConfig(**kw))
    shape = (64, 5, 4)
    f0 = seeded_state(m, shape, seed=32, du=0.04, noise=3e-4)
    fast = gpu_kernels(kw, data_type=data_type)
    slow = gpu_kernels(kw, data_type=data_type)
    for k in slow:
        k.shortcuts = False
    assert not any(_flags(k, data_type) & CLOSED for k in slow)
    faces = {0: "periodic", 1: "periodic", 2: "periodic"}
    print(f"central closed forms vs general form: {pattern} {forcing} {data_type}", flush=True)
    a = gpu_run(fast, m, shape, faces, f0, 8, data_type=data_type)
    b = gpu_run(slow, m, shape, faces, f0, 8, data_type=data_type)
    assert np.abs(a - b).max() <= tol


@pytest.mark.parametrize("pattern", ["pull", "aa"])
def test_d3q19_closed_forms_against_oracle_fp64(pattern):
    kw = dict(stencil=LBStencil(Stencil.D3Q19), method=Method.CENTRAL_MOMENT, relaxation_rate=1.8, compressible=True,
              streaming_pattern=pattern, force_model=ForceModel.GUO, force=(1e-5, 0, -1e-5))
    m = create_lb_method(LBMConfig(**kw))
    shape = (8, 7, 6)
    faces = {0: "periodic", 1: "noslip", 2: "periodic"}
    f0 = seeded_state(m, shape, seed=33, du=0.03)
    ref, _ = oracle_run(m, shape, faces, pattern, f0, 10)
    ks = gpu_kernels(kw)
    from lbmpy_amd.sweep import create_pdf_array
    assert all(k.descriptor(create_pdf_array((8, 8, 8), 19, 1, "float64"), {}).flags & CLOSED == CLOSED for k in ks)
    print(f"D3Q19 central closed forms vs oracle: {pattern}", flush=True)
    got = gpu_run(ks, m, shape, faces, f0, 10)
    assert np.abs(got - ref).max() <= 1e-12


def test_other_central_methods_keep_the_general_form():
    rates = [0.0] * 4 + [1.8] * 5 + [1.3] + [1.0] * 17          # a bulk rate of its own
    (k,) = gpu_kernels(dict(stencil=D3Q27, method=Method.CENTRAL_MOMENT, relaxation_rates=rates, compressible=True))
    assert _flags(k) & nat.FLAG_GAUSSIAN_EQUILIBRIUM
    assert not _flags(k) & nat.FLAG_SHEAR_PROJECTOR
    (k,) = gpu_kernels(dict(stencil=LBStencil(Stencil.D2Q9), method=Method.CENTRAL_MOMENT, relaxation_rate=1.8,
                            compressible=True))
    from lbmpy_amd.sweep import create_pdf_array
    assert not k.descriptor(create_pdf_array((8, 8), 9, 1, "float64"), {}).flags & CLOSED
