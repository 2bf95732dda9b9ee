"""FP16 storage with FP32 arithmetic: the sweep without ghost layers (periodic wrap and fused walls in the index
math) gives bit-identical populations to the sweep with a ghost layer on the same problem.  Both load the same
FP16 values, run the same per-cell collision and round the same results to FP16, so the comparison is exact.
(These cases once compared a row-staged half2 FP16 kernel with the per-element one; the per-element kernel
without the LDS row shift proved faster and is the only FP16 sweep now.)"""
import numpy as np
import pytest

from lbmpy_amd import LBMConfig, LBStencil, Stencil, Method, create_lb_method
from tests.lbm_helpers import seeded_state, gpu_kernels, gpu_run

pytestmark = pytest.mark.gpu

CASES = [
    ("D3Q19", Method.TRT, dict(relaxation_rates=[1.9, 1.2], force=(1e-4, 0, 0)), ("periodic", "noslip", "noslip"),
     "pull", (128, 6, 5)),
    ("D3Q19", Method.MRT, dict(relaxation_rate=1.8, force=(1e-4, 0, 0)), ("periodic", "noslip", "periodic"), "aa",
     (128, 5, 4)),
    ("D3Q27", Method.SRT, dict(relaxation_rate=1.95, compressible=True), ("periodic",) * 3, "aa", (256, 4, 3)),
    ("D3Q27", Method.TRT, dict(relaxation_rates=[1.9, 1.1], force=(1e-4, 0, 0)),
     ("periodic", "noslip", "periodic"), "pull", (128, 4, 4)),
    ("D3Q27", Method.TRT, dict(relaxation_rates=[1.7, 1.3]), ("periodic", "periodic", "noslip"), "aa", (128, 3, 5)),
    ("D2Q9", Method.SRT, dict(relaxation_rate=1.6, force=(1e-4, 0)), ("periodic", "noslip"), "pull", (128, 8)),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}-{c[1].name}-{c[4]}")
def test_fp16_sweep_without_ghost_layers_equals_with(case):
    st, meth, extra, faces, pattern, shape = case
    kw = dict(stencil=LBStencil(Stencil[st]), method=meth, streaming_pattern=pattern, zero_centered=True, **extra)
    m = create_lb_method(LBMConfig(**kw))
    f0 = seeded_state(m, shape, seed=31, du=0.03, noise=1e-3)
    fdict = {a: faces[a] for a in range(m.dim)}
    print(f"fp16 ghost layers {st} {meth.name} {pattern} {shape}", flush=True)
    runs = []
    for g in (0, 1):
        ks = gpu_kernels(kw, data_type="float16", compute_type="float32", ghost_layers=g)
        runs.append(gpu_run(ks, m, shape, fdict, f0, 8, data_type="float16", ghost_layers=g))
    assert np.isfinite(runs[0]).all()
    np.testing.assert_array_equal(runs[0], runs[1])
    # the sweeps moved the state (not a no-op comparison)
    assert np.abs(runs[0] - f0).max() > 1e-4
