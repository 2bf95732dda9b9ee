"""The C restatement of the oracle (oracle/lbm_oracle_c.c) agrees with the numpy oracle.

Both are FP64 and run the same generic algorithm; they differ in operation order only, so the bound
is round-off (1e-13 absolute on PDFs of order 1e-1 after a few steps)."""
import os

import numpy as np
import pytest

from lbmpy_amd import LBMConfig, LBStencil, Stencil, Method, create_lb_method
from tests.lbm_helpers import seeded_state, oracle_run

c_oracle = pytest.importorskip("oracle.c_oracle")
if not os.path.exists(c_oracle.LIB):
    pytest.skip("C oracle not built (python -m oracle.build)", allow_module_level=True)

CASES = [
    ("D2Q9", Method.SRT, dict(relaxation_rate=1.6)),
    ("D2Q9", Method.TRT, dict(relaxation_rates=[1.9, 1.2], compressible=True, force=(1e-4, -2e-5))),
    ("D3Q19", Method.TRT, dict(relaxation_rates=[1.97, 1.6], compressible=True, force=(1e-5, 0, 0))),
    ("D3Q19", Method.MRT, dict(relaxation_rate=1.8, force=(1e-5, 2e-6, 0))),
    ("D3Q27", Method.SRT, dict(relaxation_rate=1.5, force=(2e-5, 1e-5, 0), force_model="luo")),
    ("D3Q27", Method.CUMULANT, dict(relaxation_rate=1.9848, compressible=True)),
    ("D3Q27", Method.CUMULANT, dict(relaxation_rate=1.8, compressible=True, force=(2e-5, -1e-5, 3e-5))),
]


def _method(st, meth, extra, pattern="pull"):
    from lbmpy_amd import ForceModel
    extra = dict(extra)
    if extra.get("force_model") == "luo":
        extra["force_model"] = ForceModel.LUO
    return create_lb_method(LBMConfig(stencil=LBStencil(Stencil[st]), method=meth, streaming_pattern=pattern,
                                      **extra))


def _shape(st):
    return (7, 5) if st == "D2Q9" else (5, 4, 3)


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}-{c[1].name}")
def test_collision_matches_numpy_oracle(case):
    from oracle.lbm_oracle import OracleCollision
    m = _method(*case)
    f = seeded_state(m, _shape(case[0]), seed=21, du=0.05, noise=1e-3).reshape(m.stencil.Q, -1)
    ref, _, _ = OracleCollision(m)(f)
    got = c_oracle.COracle(m).collide(f)
    assert np.abs(got - ref).max() <= 1e-14


@pytest.mark.parametrize("pattern", ["pull", "aa"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}-{c[1].name}")
def test_periodic_sweeps_match_numpy_oracle(case, pattern):
    m = _method(*case, pattern=pattern)
    shape = _shape(case[0])
    f0 = seeded_state(m, shape, seed=22, du=0.03, noise=2e-4)
    steps = 6
    ref, _ = oracle_run(m, shape, {a: "periodic" for a in range(m.dim)}, pattern, f0, steps)
    co = c_oracle.COracle(m)
    inv = np.array(m.stencil.inverse_indices())
    if pattern == "pull":
        a, b = f0.copy(), np.zeros_like(f0)
        for _ in range(steps):
            co.sweep(a, b, "pull", nthreads=2)
            a, b = b, a
        got = a
    else:
        a = np.ascontiguousarray(f0[inv])
        for t in range(steps):
            co.sweep(a, a, "aa_odd" if t % 2 == 0 else "aa_even", nthreads=2)
        got = a[inv]
    assert np.abs(got - ref).max() <= 1e-13
