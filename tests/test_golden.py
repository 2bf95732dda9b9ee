"""Committed golden vectors (``tests/golden/*.npz``): the oracle must reproduce them on the CPU,
and the HIP sweep must match them on the GPU (FP64, max-abs <= 1e-12)."""
import os

import numpy as np
import pytest

from tests.golden.make_golden import CASES, build_method, faces_of, lbm_config_kwargs, load_case
from tests.lbm_helpers import seeded_state, oracle_run, gpu_kernels, gpu_run

GOLDEN_DIR = os.path.join(os.path.dirname(__file__), "golden")
GOLDEN = [os.path.join(GOLDEN_DIR, c[0] + ".npz") for c in CASES]
IDS = [c[0] for c in CASES]


def test_every_case_has_a_fixture():
    for p in GOLDEN:
        assert os.path.exists(p), p


@pytest.mark.parametrize("path", GOLDEN, ids=IDS)
def test_oracle_reproduces_golden(path):
    meta, f0, expected = load_case(path)
    m = build_method(meta)
    again = seeded_state(m, tuple(meta["shape"]), seed=meta["seed"], du=0.03, noise=2e-4)
    np.testing.assert_array_equal(again, f0)
    ref, _ = oracle_run(m, tuple(meta["shape"]), faces_of(meta), meta["pattern"], f0, meta["steps"])
    assert np.abs(ref - expected).max() <= 1e-13
    # the state has moved away from the start: the fixture pins collision and streaming, not a no-op
    assert np.abs(expected - f0).max() > 1e-4


@pytest.mark.gpu
@pytest.mark.parametrize("path", GOLDEN, ids=IDS)
def test_gpu_matches_golden(path):
    meta, f0, expected = load_case(path)
    m = build_method(meta)
    print(f"golden case {meta['name']}", flush=True)
    got = gpu_run(gpu_kernels(lbm_config_kwargs(meta)), m, tuple(meta["shape"]), faces_of(meta), f0,
                  meta["steps"])
    assert np.abs(got - expected).max() <= 1e-12
