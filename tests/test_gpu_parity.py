"""GPU parity: HIP sweep vs the FP64 oracle on the same seeded inputs.

FP64 tolerance: max-abs PDF difference <= 1e-12 after N steps (the bar the
reference's own CPU/GPU equivalence test sets is 1e-14 after 13 steps,
``tests/test_cpu_gpu_equivalence.py:13-28``; over hundreds of steps round-off
of two different operation orders accumulates, 1e-12 is the stated bound).
"""
import numpy as np
import pytest

from lbmpy_amd import LBMConfig, LBStencil, Stencil, Method, create_lb_method

from tests.lbm_helpers import seeded_state, oracle_run, gpu_kernels, gpu_run

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12

MOMENT_CASES = [
    # (stencil, method, extra config, faces, pattern)
    ("D2Q9", Method.SRT, dict(relaxation_rate=1.6), ("periodic", "noslip"), "pull"),
    ("D2Q9", Method.TRT, dict(relaxation_rates=[1.9, 1.2], compressible=True), ("periodic", "periodic"), "aa"),
    ("D2Q9", Method.MRT, dict(relaxation_rate=1.7, force=(1e-4, -2e-5)), ("periodic", "noslip"), "pull"),
    ("D3Q19", Method.SRT, dict(relaxation_rate=1.8, compressible=True), ("periodic",) * 3, "aa"),
    ("D3Q19", Method.TRT, dict(relaxation_rates=[1.97, 1.6], compressible=True, force=(1e-5, 0, 0)),
     ("periodic", "noslip", "noslip"), "pull"),
    ("D3Q19", Method.MRT, dict(relaxation_rate=1.8, force=(1e-5, 0, 0)), ("periodic", "noslip", "periodic"), "pull"),
    ("D3Q19", Method.MRT_RAW, dict(relaxation_rates=[0, 0, 0, 0] + [1.3] * 15), ("periodic",) * 3, "aa"),
    ("D3Q27", Method.SRT, dict(relaxation_rate=1.5, force=(2e-5, 1e-5, 0)), ("periodic",) * 3, "pull"),
    ("D3Q27", Method.TRT, dict(relaxation_rate=1.9, continuous_equilibrium=False), ("noslip", "periodic", "periodic"),
     "pull"),
    ("D3Q27", Method.SRT, dict(relaxation_rate=1.7, zero_centered=False, compressible=True), ("periodic",) * 3, "aa"),
]


def _shape(st):
    return (12, 8) if st == "D2Q9" else (10, 8, 6)


@pytest.mark.parametrize("case", MOMENT_CASES, ids=lambda c: f"{c[0]}-{c[1].name}-{c[4]}")
def test_moment_family_fp64(case):
    st, meth, extra, faces, pattern = case
    kw = dict(stencil=LBStencil(Stencil[st]), method=meth, streaming_pattern=pattern, **extra)
    m = create_lb_method(LBMConfig(**kw))
    shape = _shape(st)
    f0 = seeded_state(m, shape, seed=1)
    steps = 24
    fdict = {a: faces[a] for a in range(m.dim)}
    ref, _ = oracle_run(m, shape, fdict, pattern, f0, steps)
    print(f"case {case[0]} {case[1].name} {pattern} shape={shape}", flush=True)
    got = gpu_run(gpu_kernels(kw), m, shape, fdict, f0, steps)
    err = np.abs(got - ref).max()
    assert err <= FP64_TOL, err


@pytest.mark.parametrize("pattern", ["pull", "aa"])
@pytest.mark.parametrize("force", [(0, 0, 0), (2e-5, -1e-5, 3e-5)])
def test_cumulant_d3q27_fp64(pattern, force):
    kw = dict(stencil=LBStencil(Stencil.D3Q27), method=Method.CUMULANT, relaxation_rate=1.85, compressible=True,
              streaming_pattern=pattern, force=force)
    m = create_lb_method(LBMConfig(**kw))
    shape = (8, 6, 10)
    f0 = seeded_state(m, shape, seed=2, du=0.04)
    steps = 20
    fdict = {0: "periodic", 1: "periodic", 2: "periodic"}
    ref, _ = oracle_run(m, shape, fdict, pattern, f0, steps)
    print(f"cumulant {pattern} force={force}", flush=True)
    got = gpu_run(gpu_kernels(kw), m, shape, fdict, f0, steps)
    err = np.abs(got - ref).max()
    assert err <= FP64_TOL, err


def test_cumulant_d3q27_fp64_all_rates():
    """Every cumulant group relaxed with its own rate (Q rates) exercises each branch of the
    polynomial relaxation."""
    rates = [0.0, 0, 0, 0, 1.7, 1.6, 1.5, 1.4, 1.3, 1.2, 1.1, 1.0, 0.9, 0.8, 0.7, 0.6, 0.5, 0.4, 0.3,
             0.35, 0.45, 0.55, 0.65, 0.75, 0.85, 0.95, 1.05]
    kw = dict(stencil=LBStencil(Stencil.D3Q27), method=Method.CUMULANT, relaxation_rates=rates, compressible=True,
              streaming_pattern="pull", force=(1e-5, 2e-5, -1e-5))
    m = create_lb_method(LBMConfig(**kw))
    shape = (6, 6, 6)
    f0 = seeded_state(m, shape, seed=3, du=0.05, noise=5e-4)
    fdict = {0: "periodic", 1: "periodic", 2: "periodic"}
    ref, _ = oracle_run(m, shape, fdict, "pull", f0, 10)
    got = gpu_run(gpu_kernels(kw), m, shape, fdict, f0, 10)
    assert np.abs(got - ref).max() <= FP64_TOL


# nx a multiple of 64 with periodic x selects the LDS row-shift kernels (blockDim.x == nx)
ROWLDS_CASES = [
    ("D2Q9", Method.TRT, dict(relaxation_rates=[1.8, 1.1], force=(1e-4, 0)), ("periodic", "noslip"), "pull", (64, 6)),
    ("D2Q9", Method.SRT, dict(relaxation_rate=1.6), ("periodic", "periodic"), "aa", (128, 5)),
    ("D3Q19", Method.TRT, dict(relaxation_rates=[1.97, 1.6], compressible=True, force=(1e-5, 0, 0)),
     ("periodic", "noslip", "noslip"), "pull", (64, 5, 4)),
    ("D3Q19", Method.MRT, dict(relaxation_rate=1.8, force=(1e-5, 0, 0)), ("periodic", "noslip", "periodic"), "aa",
     (64, 4, 5)),
    ("D3Q27", Method.CUMULANT, dict(relaxation_rate=1.9, compressible=True, force=(1e-5, 2e-5, 0)),
     ("periodic", "periodic", "periodic"), "aa", (64, 4, 3)),
    ("D3Q27", Method.CUMULANT, dict(relaxation_rate=1.7, compressible=True), ("periodic", "noslip", "periodic"),
     "pull", (64, 3, 4)),
    ("D3Q27", Method.CUMULANT, dict(relaxation_rate=1.7, compressible=True), ("periodic", "periodic", "noslip"),
     "aa", (64, 3, 4)),
]


@pytest.mark.parametrize("case", ROWLDS_CASES, ids=lambda c: f"{c[0]}-{c[1].name}-{c[4]}-{'x'.join(map(str, c[5]))}")
def test_row_lds_kernels_fp64(case):
    st, meth, extra, faces, pattern, shape = case
    kw = dict(stencil=LBStencil(Stencil[st]), method=meth, streaming_pattern=pattern, **extra)
    m = create_lb_method(LBMConfig(**kw))
    f0 = seeded_state(m, shape, seed=4)
    steps = 16
    fdict = {a: faces[a] for a in range(m.dim)}
    ref, _ = oracle_run(m, shape, fdict, pattern, f0, steps)
    print(f"rowlds case {st} {meth.name} {pattern} {shape}", flush=True)
    got = gpu_run(gpu_kernels(kw), m, shape, fdict, f0, steps)
    assert np.abs(got - ref).max() <= FP64_TOL


@pytest.mark.parametrize("pattern", ["aa", "pull"])
def test_cumulant_smagorinsky_fp64(pattern):
    """Smagorinsky model fused into the cumulant sweep (turbulence_models.py:25-84), Cs = 0.12."""
    from lbmpy_amd import SubgridScaleModel
    from oracle.lbm_oracle import OracleDomain
    kw = dict(stencil=LBStencil(Stencil.D3Q27), method=Method.CUMULANT, relaxation_rate=1.99, compressible=True,
              streaming_pattern=pattern, subgrid_scale_model=SubgridScaleModel.SMAGORINSKY, force=(1e-5, 0, 0))
    m = create_lb_method(LBMConfig(**kw))
    shape = (8, 6, 6)
    f0 = seeded_state(m, shape, seed=6, du=0.05, noise=1e-3)
    fdict = {0: "periodic", 1: "periodic", 2: "periodic"}
    dom = OracleDomain(m, shape, faces=fdict, pattern=pattern)
    dom.coll.smagorinsky = 0.12
    dom.set_pdfs_interior(f0[dom.inv] if pattern == "aa" else f0)
    for _ in range(12):
        dom.step()
    ref = dom.pdfs_in_natural_layout()
    got = gpu_run(gpu_kernels(kw), m, shape, fdict, f0, 12)
    assert np.abs(got - ref).max() <= FP64_TOL


CM_CASES = [
    ("D2Q9", dict(relaxation_rate=1.7), ("periodic", "noslip"), "pull", (12, 8)),
    ("D2Q9", dict(relaxation_rate=1.6, compressible=True, force=(1e-4, 2e-5)), ("periodic", "periodic"), "aa", (64, 6)),
    ("D3Q19", dict(relaxation_rate=1.8, force=(1e-5, 0, 0)), ("periodic", "noslip", "periodic"), "pull", (10, 8, 6)),
    ("D3Q19", dict(relaxation_rate=1.75, compressible=True), ("periodic",) * 3, "aa", (64, 4, 3)),
    ("D3Q27", dict(relaxation_rate=1.9, compressible=True, force=(1e-5, 2e-5, -1e-5)), ("periodic",) * 3, "pull",
     (8, 6, 10)),
    ("D3Q27", dict(relaxation_rate=1.7, force=(2e-5, 0, 0), force_model=None), ("periodic", "noslip", "noslip"), "aa",
     (10, 6, 5)),
    ("D3Q27", dict(relaxation_rates=[0, 0, 0, 0, 1.7, 1.7, 1.7, 1.6, 1.6, 1.1, 1.2, 1.2, 1.2, 1.3, 1.3, 1.3, 1.4,
                                     1.5, 1.5, 0.9, 0.8, 0.8, 0.8, 0.7, 0.7, 0.7, 0.6], compressible=True,
                   force=(1e-5, 0, 0)), ("periodic",) * 3, "pull", (6, 6, 6)),
]


@pytest.mark.parametrize("case", CM_CASES, ids=lambda c: f"{c[0]}-{c[3]}-{'x'.join(map(str, c[4]))}")
def test_central_moment_family_fp64(case):
    """Central-moment methods (cascaded moment sets, compressible and incompressible, Guo force)."""
    st, extra, faces, pattern, shape = case
    extra = {k: v for k, v in extra.items() if not (k == "force_model" and v is None)}
    kw = dict(stencil=LBStencil(Stencil[st]), method=Method.CENTRAL_MOMENT, streaming_pattern=pattern, **extra)
    m = create_lb_method(LBMConfig(**kw))
    f0 = seeded_state(m, shape, seed=8)
    fdict = {a: faces[a] for a in range(m.dim)}
    ref, _ = oracle_run(m, shape, fdict, pattern, f0, 16)
    print(f"central-moment case {st} {pattern} {shape}", flush=True)
    got = gpu_run(gpu_kernels(kw), m, shape, fdict, f0, 16)
    assert np.abs(got - ref).max() <= FP64_TOL
